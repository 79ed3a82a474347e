"""Helpers for the GPU parity tests: move numpy matrices to column-major
device buffers, call the C ABI of libtsqr_amd.so, and bring Q and R back."""
from __future__ import annotations

import ctypes

import numpy as np


def torch_dtype(np_dtype):
    import torch

    return torch.float32 if np.dtype(np_dtype) == np.float32 else torch.float64


def to_device_colmajor(A: np.ndarray, lda: int | None = None, offset_elems: int = 0):
    """Returns (buffer, ptr, lda) with A stored column-major at ptr; rows
    m..lda-1 of each column are filled with NaN so that any read or write
    outside the matrix shows up in the results."""
    import torch

    m, n = A.shape
    lda = m if lda is None else lda
    buf = torch.full((offset_elems + n * lda,), float("nan"), dtype=torch_dtype(A.dtype), device="cuda")
    view = buf[offset_elems:].view(n, lda)
    view[:, :m] = torch.from_numpy(np.ascontiguousarray(A.T))
    return buf, view.data_ptr(), lda


def from_device_colmajor(buf, m: int, n: int, lda: int, offset_elems: int = 0) -> np.ndarray:
    v = buf[offset_elems:].view(n, lda)[:, :m]
    return v.cpu().numpy().T.copy()


class LeafPath:
    """Context manager selecting the level-0 leaf path for one dtype."""

    def __init__(self, dtype, path: int, gate: float | None = None):
        from tsqr_amd._lib import TSQR_DOUBLE, TSQR_FLOAT

        self.code = TSQR_FLOAT if np.dtype(dtype) == np.float32 else TSQR_DOUBLE
        if gate is None:  # the gate each path is meant for
            gate = 1e5 if (self.code == TSQR_DOUBLE and path == 3) else 1024.0
        self.path, self.gate = path, gate

    def __enter__(self):
        from tsqr_amd._lib import TSQR_LEAF_AUTO, TSQR_LEAF_CHOLQR2, TSQR_FLOAT, set_leaf_path

        set_leaf_path(self.code, self.path, self.gate)
        # the library's defaults: FP32 AUTO (gate 1024), FP64 CHOLQR2 (gate 1e5)
        self.default = (TSQR_LEAF_AUTO, 1024.0) if self.code == TSQR_FLOAT else (TSQR_LEAF_CHOLQR2, 1e5)
        return self

    def __exit__(self, *exc):
        from tsqr_amd._lib import set_leaf_path

        set_leaf_path(self.code, *self.default)


LAST_LEAF_STATS = {}
LAST_MIXED_STATS = {}


def call(kind: str, A: np.ndarray, lda: int | None = None, offset_elems: int = 0, work_bytes: int | None = None,
         ref_shaped: bool = False, ldwork: int = 0):
    """Runs tsqr / qr / tsqr_mixed (kind) on a copy of A through the C ABI.
    Returns (status, Q, R, padding_untouched); through the stream entry
    points the Gram / Householder superleaf counts of tsqr land in
    LAST_LEAF_STATS and the outcome of tsqr_mixed in LAST_MIXED_STATS."""
    import torch

    import tsqr_amd
    from tsqr_amd._lib import TSQR_DOUBLE, TSQR_FLOAT

    L = tsqr_amd.lib()
    m, n = A.shape
    code = TSQR_FLOAT if A.dtype == np.float32 else TSQR_DOUBLE
    suf = "s" if code == TSQR_FLOAT else "d"
    buf, pA, lda = to_device_colmajor(A, lda, offset_elems)
    Rbuf = torch.full((n * n,), float("nan"), dtype=torch_dtype(A.dtype), device="cuda")
    need = (L.qr_workspace_bytes if kind.startswith("qr") else L.tsqr_workspace_bytes)(m, n, code)
    if ref_shaped:
        elem = 4 if code == TSQR_FLOAT else 8
        work = torch.empty(max(ldwork * n * elem, 1), dtype=torch.uint8, device="cuda")
        st = getattr(L, f"{kind}_{suf}")(m, n, pA, lda, Rbuf.data_ptr(), n, work.data_ptr(), ldwork)
    else:
        nbytes = need if work_bytes is None else work_bytes
        work = torch.empty(max(nbytes, 1), dtype=torch.uint8, device="cuda")
        st = getattr(L, f"{kind}_{'s' if 'mixed' in kind else suf}_stream")(m, n, pA, lda, Rbuf.data_ptr(), n,
                                                                            work.data_ptr(), nbytes, None)
        if st == 0 and kind == "tsqr":
            from tsqr_amd._lib import leaf_stats

            (LAST_LEAF_STATS["gram"], LAST_LEAF_STATS["householder"],
             LAST_LEAF_STATS["flat"]) = leaf_stats(work.data_ptr(), m, n, code)
        if st == 0 and kind == "tsqr_mixed":
            from tsqr_amd._lib import mixed_stats

            LAST_MIXED_STATS.clear()
            LAST_MIXED_STATS.update(mixed_stats(work.data_ptr(), m, n))
    torch.cuda.synchronize()
    Q = from_device_colmajor(buf, m, n, lda, offset_elems)
    R = Rbuf.view(n, n).cpu().numpy().T.copy()
    pad_ok = True
    if lda > m:
        pad = buf[offset_elems:].view(n, lda)[:, m:]
        pad_ok = bool(torch.isnan(pad).all().item())
    if offset_elems:
        pad_ok = pad_ok and bool(torch.isnan(buf[:offset_elems]).all().item())
    return st, Q, R, pad_ok
