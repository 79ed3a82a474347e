"""tsqr_amd -- MI355X (gfx950) tall-skinny and recursive blocked Householder QR.

The compute lives in libtsqr_amd.so (hand-written HIP kernels behind the C
ABI declared in include/tsqr_amd.h).  This module is a thin ctypes binding
that takes torch tensors only as holders of device memory and streams:

    tsqr(A, R=None, work=None)   A (m x n, column-major, on the GPU) is
                                 overwritten by Q; returns (A, R)
    qr(A, R=None, work=None)     recursive blocked QR, same convention
    tsqr_mixed(A, R=None, work=None)
                                 float32 TSQR on the BF16 matrix cores
                                 (CholeskyQR with a residual check and an
                                 optional second pass), same convention
    tsqr_r(A, R=None, work=None) R only; A is read, never written
    check(A, Q, R)               device-side FP64 backward and
                                 orthogonality errors of A = QR
    Batch(mats, method="plain"|"mixed")
                                 many uneven TSQRs (n <= 64) planned once,
                                 run as a few launches or one hipGraph

A column-major m x n tensor is one whose strides are (1, lda), for example
``colmajor(torch.randn(m, n))``.  Errors from the library raise TsqrError
with the library's status code.  There is no CPU fallback: if the shared
library is missing the import raises.
"""
from __future__ import annotations

from ._lib import (  # noqa: F401
    LIB_PATH,
    TSQR_DOUBLE,
    TSQR_FLOAT,
    TsqrError,
    lib,
    qr_workspace_bytes,
    tsqr_workspace_bytes,
)
from .api import Batch, check, colmajor, empty_colmajor, qr, qr_mixed, tsqr, tsqr_mixed, tsqr_r  # noqa: F401

__all__ = [
    "Batch",
    "LIB_PATH",
    "check",
    "TsqrError",
    "colmajor",
    "empty_colmajor",
    "lib",
    "qr",
    "qr_mixed",
    "qr_workspace_bytes",
    "tsqr",
    "tsqr_mixed",
    "tsqr_r",
    "tsqr_workspace_bytes",
]
