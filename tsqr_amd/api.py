"""torch-facing wrappers over the C ABI (device memory and streams only)."""
from __future__ import annotations

import torch

from ._lib import TSQR_DOUBLE, TSQR_FLOAT, TsqrError, lib, qr_workspace_bytes, tsqr_workspace_bytes


def colmajor(x: torch.Tensor) -> torch.Tensor:
    """Column-major copy of a 2-D tensor (strides (1, rows))."""
    return x.t().contiguous().t()


def empty_colmajor(m: int, n: int, dtype=torch.float32, device="cuda") -> torch.Tensor:
    return torch.empty((n, m), dtype=dtype, device=device).t()


def _code(dtype) -> int:
    if dtype == torch.float32:
        return TSQR_FLOAT
    if dtype == torch.float64:
        return TSQR_DOUBLE
    raise TypeError(f"unsupported dtype {dtype}")


def _ld(x: torch.Tensor, name: str) -> int:
    if x.dim() != 2 or x.stride(0) != 1 or not x.is_cuda:
        raise ValueError(f"{name} must be a column-major 2-D GPU tensor (strides (1, ld))")
    return x.stride(1) if x.shape[1] > 1 else max(x.shape[0], 1)


def _run(kind: str, A: torch.Tensor, R, work, stream):
    m, n = A.shape
    code = _code(A.dtype)
    lda = _ld(A, "A")
    if R is None:
        R = empty_colmajor(n, n, A.dtype, A.device)
    ldr = _ld(R, "R")
    need = (qr_workspace_bytes if kind == "qr" else tsqr_workspace_bytes)(m, n, code)
    if work is None:
        work = torch.empty(max(need, 1), dtype=torch.uint8, device=A.device)
    nbytes = work.numel() * work.element_size()
    if stream is None:
        stream = torch.cuda.current_stream(A.device)
    suf = "s" if code == TSQR_FLOAT else "d"
    if kind == "tsqr_mixed":
        if code != TSQR_FLOAT:
            raise TypeError("the mixed-precision TSQR takes float32 input")
        suf = "s"
    fn = getattr(lib(), f"{kind}_{suf}_stream")
    st = fn(m, n, A.data_ptr(), lda, R.data_ptr(), ldr, work.data_ptr(), nbytes, stream.cuda_stream)
    if st != 0:
        raise TsqrError(st, f"{kind}_{suf}_stream(m={m}, n={n})")
    return A, R


def tsqr(A: torch.Tensor, R: torch.Tensor | None = None, work: torch.Tensor | None = None, stream=None):
    """TSQR of a column-major m x n GPU tensor, n <= 32 (TallShinnyQR.h:86-149).
    A is overwritten by the explicit thin Q; returns (A, R)."""
    return _run("tsqr", A, R, work, stream)


def qr(A: torch.Tensor, R: torch.Tensor | None = None, work: torch.Tensor | None = None, stream=None):
    """Recursive blocked QR (block_QR.h:72-95); same conventions as tsqr."""
    return _run("qr", A, R, work, stream)


def tsqr_mixed(A: torch.Tensor, R: torch.Tensor | None = None, work: torch.Tensor | None = None, stream=None):
    """Mixed-precision TSQR of a float32 column-major m x n GPU tensor, n <= 32:
    BF16x3 CholeskyQR with an FP32 residual check and an optional second pass.
    Same conventions as tsqr."""
    return _run("tsqr_mixed", A, R, work, stream)
