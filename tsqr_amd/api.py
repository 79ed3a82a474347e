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
    need = (qr_workspace_bytes if kind.startswith("qr") else tsqr_workspace_bytes)(m, n, code)  # tsqr_r: as tsqr
    if work is None:
        work = torch.empty(max(need, 1), dtype=torch.uint8, device=A.device)
    nbytes = work.numel() * work.element_size()
    if stream is None:
        stream = torch.cuda.current_stream(A.device)
    suf = "s" if code == TSQR_FLOAT else "d"
    if kind in ("tsqr_mixed", "qr_mixed"):
        if code != TSQR_FLOAT:
            raise TypeError("the mixed-precision TSQR takes float32 input")
        suf = "s"
    if kind == "tsqr_r":
        fn = getattr(lib(), f"tsqr_r_{suf}_stream")
    else:
        fn = getattr(lib(), f"{kind}_{suf}_stream")
    st = fn(m, n, A.data_ptr(), lda, R.data_ptr(), ldr, work.data_ptr(), nbytes, stream.cuda_stream)
    if st != 0:
        raise TsqrError(st, f"{kind}_{suf}_stream(m={m}, n={n})")
    return A, R


def tsqr(A: torch.Tensor, R: torch.Tensor | None = None, work: torch.Tensor | None = None, stream=None):
    """TSQR of a column-major m x n GPU tensor, n <= 32 (TallShinnyQR.h:86-149).
    A is overwritten by the explicit thin Q; returns (A, R)."""
    return _run("tsqr", A, R, work, stream)


def tsqr_r(A: torch.Tensor, R: torch.Tensor | None = None, work: torch.Tensor | None = None, stream=None):
    """R factor only of a column-major m x n GPU tensor, n <= 32
    (tsqr_r_s / tsqr_r_d): A is read and left unchanged; returns R."""
    return _run("tsqr_r", A, R, work, stream)[1]


def check(A: torch.Tensor, Q: torch.Tensor, R: torch.Tensor, stream=None) -> dict:
    """Device-side FP64 check of A = QR (tsqr_check_s / tsqr_check_d, n <= 32):
    backward_fro = ||A - QR||_F / ||A||_F, orth_fro = ||I - Q^T Q||_F and
    their spectral-norm counterparts backward_2, orth_2.  Waits for the
    result."""
    from ._lib import check as _check

    m, n = A.shape
    if Q.shape != (m, n) or R.shape != (n, n) or Q.dtype != A.dtype or R.dtype != A.dtype:
        raise ValueError("A, Q must be m x n and R n x n, all of one dtype")
    if stream is None:
        stream = torch.cuda.current_stream(A.device)
    return _check(m, n, _code(A.dtype), A.data_ptr(), _ld(A, "A"), Q.data_ptr(), _ld(Q, "Q"), R.data_ptr(),
                  _ld(R, "R"), stream.cuda_stream)


def qr(A: torch.Tensor, R: torch.Tensor | None = None, work: torch.Tensor | None = None, stream=None):
    """Recursive blocked QR (block_QR.h:72-95); same conventions as tsqr."""
    return _run("qr", A, R, work, stream)


def tsqr_mixed(A: torch.Tensor, R: torch.Tensor | None = None, work: torch.Tensor | None = None, stream=None):
    """Mixed-precision TSQR of a float32 column-major m x n GPU tensor, n <= 32:
    BF16x3 CholeskyQR with an FP32 residual check and an optional second pass.
    Same conventions as tsqr."""
    return _run("tsqr_mixed", A, R, work, stream)


def qr_mixed(A: torch.Tensor, R: torch.Tensor | None = None, work: torch.Tensor | None = None, stream=None):
    """Mixed-precision recursive blocked QR of a float32 column-major GPU
    tensor: mixed TSQR panels and a reorthogonalised coupling.  Same
    conventions as qr."""
    return _run("qr_mixed", A, R, work, stream)


class Batch:
    """A planned batch of independent TSQRs (tsqr_batch_* in the C ABI).

    mats: column-major m_i x n_i GPU tensors (n_i <= 64), all float32 or all
    float64; each is overwritten by its Q when the batch runs.  Rs: the
    n_i x n_i R tensors (allocated if None).  method: "plain" (Householder)
    or "mixed" (float32 only).  The plan holds the tensors' pointers, so the
    tensors must stay alive and in place while the plan is used."""

    def __init__(self, mats, Rs=None, method: str = "plain"):
        import ctypes

        from ._lib import TSQR_BATCH_MIXED, TSQR_BATCH_PLAIN

        if not mats:
            raise ValueError("empty batch")
        dtype = mats[0].dtype
        code = _code(dtype)
        self.mats = list(mats)
        self.Rs = list(Rs) if Rs is not None else [empty_colmajor(a.shape[1], a.shape[1], dtype, a.device)
                                                     for a in self.mats]
        k = len(self.mats)
        i32 = ctypes.c_int * k
        vp = ctypes.c_void_p * k
        m = i32(*[a.shape[0] for a in self.mats])
        n = i32(*[a.shape[1] for a in self.mats])
        lda = i32(*[_ld(a, "A") for a in self.mats])
        ldr = i32(*[_ld(r, "R") for r in self.Rs])
        A = vp(*[a.data_ptr() for a in self.mats])
        R = vp(*[r.data_ptr() for r in self.Rs])
        meth = {"plain": TSQR_BATCH_PLAIN, "mixed": TSQR_BATCH_MIXED}[method]
        self._h = ctypes.c_void_p()
        st = lib().tsqr_batch_create(code, meth, k, m, n, A, lda, R, ldr, ctypes.byref(self._h))
        if st != 0:
            raise TsqrError(st, f"tsqr_batch_create(count={k}, method={method})")
        nbytes = int(lib().tsqr_batch_workspace_bytes(self._h))
        self.work = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=self.mats[0].device)

    def _stream(self, stream):
        return (stream if stream is not None else torch.cuda.current_stream(self.work.device)).cuda_stream

    def run(self, stream=None):
        st = lib().tsqr_batch_run(self._h, self.work.data_ptr(), self.work.numel(), self._stream(stream))
        if st != 0:
            raise TsqrError(st, "tsqr_batch_run")

    def graph_launch(self, stream=None):
        st = lib().tsqr_batch_graph_launch(self._h, self.work.data_ptr(), self.work.numel(), self._stream(stream))
        if st != 0:
            raise TsqrError(st, "tsqr_batch_graph_launch")

    def set_mixed(self, gate: float = 1024.0, tol: float = 8e-7):
        st = lib().tsqr_batch_set_mixed(self._h, gate, tol)
        if st != 0:
            raise TsqrError(st, "tsqr_batch_set_mixed")

    def stats(self) -> dict:
        import ctypes

        v = [ctypes.c_int(0) for _ in range(4)]
        st = lib().tsqr_batch_stats(self._h, self.work.data_ptr(), *[ctypes.byref(x) for x in v])
        if st != 0:
            raise TsqrError(st, "tsqr_batch_stats")
        out = dict(zip(("cholesky", "householder", "second_pass", "check_failed"), (x.value for x in v)))
        f = [ctypes.c_int(0) for _ in range(2)]
        st = lib().tsqr_batch_q_forms(self._h, self.work.data_ptr(), *[ctypes.byref(x) for x in f])
        if st != 0:
            raise TsqrError(st, "tsqr_batch_q_forms")
        out["bf16_q_pass1"], out["bf16_q_pass2"] = f[0].value, f[1].value
        return out

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().tsqr_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
