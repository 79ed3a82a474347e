"""ctypes loader for libtsqr_amd.so (include/tsqr_amd.h)."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtsqr_amd.so")

TSQR_OK = 0
TSQR_ERR_SCRATCH = -6
TSQR_FLOAT = 1
TSQR_DOUBLE = 2
TSQR_LEAF_AUTO = 0
TSQR_LEAF_HOUSEHOLDER = 1
TSQR_LEAF_GRAM_TREE = 2
TSQR_LEAF_CHOLQR2 = 3
TSQR_BATCH_PLAIN = 0
TSQR_BATCH_MIXED = 1

# Every symbol include/tsqr_amd.h declares.
EXPORTS = (
    "tsqr_status_string",
    "tsqr_workspace_bytes",
    "qr_workspace_bytes",
    "tsqr_s",
    "tsqr_d",
    "tsqr_s_stream",
    "tsqr_d_stream",
    "tsqr_s_stream_noalloc",
    "qr_s",
    "qr_d",
    "qr_s_stream",
    "qr_d_stream",
    "tsqr_set_leaf_path",
    "tsqr_leaf_stats",
    "tsqr_timing_enable",
    "tsqr_timing_read",
    "tsqr_mixed_s",
    "tsqr_mixed_s_stream",
    "tsqr_set_mixed",
    "tsqr_mixed_stats",
    "tsqr_mixed_q_forms",
    "qr_mixed_s",
    "qr_mixed_s_stream",
    "tsqr_set_leaf_tiles",
    "tsqr_leaf_tiles",
    "tsqr_set_gemm_waves",
    "tsqr_set_tree_arity",
    "tsqr_tree_arity",
    "tsqr_set_stream_shape",
    "tsqr_batch_create",
    "tsqr_batch_workspace_bytes",
    "tsqr_batch_run",
    "tsqr_batch_graph_launch",
    "tsqr_batch_set_mixed",
    "tsqr_batch_stats",
    "tsqr_batch_q_forms",
    "tsqr_batch_destroy",
    "tsqr_check_s",
    "tsqr_check_d",
    "tsqr_r_s",
    "tsqr_r_d",
    "tsqr_r_s_stream",
    "tsqr_r_d_stream",
    "tsqr_solve_workspace_bytes",
    "tsqr_lstsq_s_stream",
    "tsqr_lstsq_d_stream",
    "tsqr_apply_qt_s_stream",
    "tsqr_apply_qt_d_stream",
)


class TsqrError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        super().__init__(f"{where}: status {status} ({status_string(status)})")


_lib = None


def lib() -> ctypes.CDLL:
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m tsqr_amd.build` "
            "(there is no CPU fallback)"
        )
    L = ctypes.CDLL(LIB_PATH)
    i32, vp, sz = ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t
    L.tsqr_status_string.argtypes = [i32]
    L.tsqr_status_string.restype = ctypes.c_char_p
    for f in (L.tsqr_workspace_bytes, L.qr_workspace_bytes):
        f.argtypes = [i32, i32, i32]
        f.restype = sz
    for name in ("tsqr_s", "tsqr_d", "qr_s", "qr_d", "tsqr_mixed_s", "qr_mixed_s", "tsqr_r_s", "tsqr_r_d"):
        f = getattr(L, name)
        f.argtypes = [i32, i32, vp, i32, vp, i32, vp, i32]
        f.restype = i32
    for name in ("tsqr_s_stream", "tsqr_d_stream", "qr_s_stream", "qr_d_stream", "tsqr_mixed_s_stream",
                 "qr_mixed_s_stream", "tsqr_r_s_stream", "tsqr_r_d_stream"):
        f = getattr(L, name)
        f.argtypes = [i32, i32, vp, i32, vp, i32, vp, sz, vp]
        f.restype = i32
    L.tsqr_s_stream_noalloc.argtypes = [i32, i32, vp, i32, vp, i32, vp]
    L.tsqr_s_stream_noalloc.restype = i32
    L.tsqr_set_leaf_path.argtypes = [i32, i32, ctypes.c_double]
    L.tsqr_set_leaf_path.restype = i32
    L.tsqr_leaf_stats.argtypes = [vp, i32, i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.tsqr_leaf_stats.restype = i32
    L.tsqr_set_mixed.argtypes = [ctypes.c_double, ctypes.c_double]
    L.tsqr_set_mixed.restype = i32
    pi = ctypes.POINTER(i32)
    L.tsqr_mixed_stats.argtypes = [vp, i32, i32, pi, pi, pi, ctypes.POINTER(ctypes.c_double)]
    L.tsqr_mixed_stats.restype = i32
    L.tsqr_mixed_q_forms.argtypes = [vp, i32, i32, pi, pi]
    L.tsqr_mixed_q_forms.restype = i32
    ip = ctypes.POINTER(i32)
    L.tsqr_batch_create.argtypes = [i32, i32, i32, ip, ip, ctypes.POINTER(vp), ip, ctypes.POINTER(vp), ip,
                                    ctypes.POINTER(vp)]
    L.tsqr_batch_create.restype = i32
    L.tsqr_batch_workspace_bytes.argtypes = [vp]
    L.tsqr_batch_workspace_bytes.restype = sz
    for name in ("tsqr_batch_run", "tsqr_batch_graph_launch"):
        f = getattr(L, name)
        f.argtypes = [vp, vp, sz, vp]
        f.restype = i32
    L.tsqr_batch_set_mixed.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    L.tsqr_batch_set_mixed.restype = i32
    L.tsqr_batch_stats.argtypes = [vp, vp, ip, ip, ip, ip]
    L.tsqr_batch_stats.restype = i32
    L.tsqr_batch_q_forms.argtypes = [vp, vp, ip, ip]
    L.tsqr_batch_q_forms.restype = i32
    L.tsqr_batch_destroy.argtypes = [vp]
    L.tsqr_batch_destroy.restype = i32
    L.tsqr_set_leaf_tiles.argtypes = [i32, i32, i32, i32]
    L.tsqr_set_leaf_tiles.restype = i32
    L.tsqr_leaf_tiles.argtypes = [i32, i32, i32]
    L.tsqr_leaf_tiles.restype = i32
    L.tsqr_set_gemm_waves.argtypes = [i32]
    L.tsqr_set_gemm_waves.restype = i32
    L.tsqr_set_tree_arity.argtypes = [i32, i32, i32]
    L.tsqr_set_tree_arity.restype = i32
    L.tsqr_tree_arity.argtypes = [i32, i32, i32]
    L.tsqr_tree_arity.restype = i32
    L.tsqr_set_stream_shape.argtypes = [i32, i32]
    L.tsqr_set_stream_shape.restype = i32
    for name in ("tsqr_check_s", "tsqr_check_d"):
        f = getattr(L, name)
        f.argtypes = [i32, i32, vp, i32, vp, i32, vp, i32, ctypes.POINTER(ctypes.c_double), vp]
        f.restype = i32
    L.tsqr_solve_workspace_bytes.argtypes = [i32, i32, i32, i32]
    L.tsqr_solve_workspace_bytes.restype = sz
    for name in ("tsqr_lstsq_s_stream", "tsqr_lstsq_d_stream"):
        f = getattr(L, name)
        f.argtypes = [i32, i32, i32, vp, i32, vp, i32, vp, i32, vp, sz, vp]
        f.restype = i32
    for name in ("tsqr_apply_qt_s_stream", "tsqr_apply_qt_d_stream"):
        f = getattr(L, name)
        f.argtypes = [i32, i32, i32, vp, i32, vp, i32, vp, i32, vp, i32, vp, sz, vp]
        f.restype = i32
    L.tsqr_timing_enable.argtypes = [i32]
    L.tsqr_timing_enable.restype = i32
    L.tsqr_timing_read.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), i32]
    L.tsqr_timing_read.restype = i32
    _lib = L
    # the autotuned superleaf table for this GPU, if one has been written
    # (tsqr_amd/autotune.py); TSQR_AMD_NO_TUNING=1 keeps the built-in rule
    if not os.environ.get("TSQR_AMD_NO_TUNING"):
        from .autotune import DEFAULT_FILE, apply, load

        if os.path.exists(DEFAULT_FILE):
            apply(load(DEFAULT_FILE))
    return L


PHASES = ("leaf_r", "tree_up", "down", "leaf_q", "qr_gemm_tn", "qr_gemm_nn")


def timing_read() -> dict:
    """Summed milliseconds and interval counts per phase since the last read."""
    n = len(PHASES)
    ms = (ctypes.c_double * n)()
    cnt = (ctypes.c_int * n)()
    st = lib().tsqr_timing_read(ms, cnt, n)
    if st != TSQR_OK:
        raise TsqrError(st, "tsqr_timing_read")
    return {PHASES[i]: (ms[i], cnt[i]) for i in range(n)}


def set_leaf_path(dtype_code: int, path: int, gate: float = 1024.0) -> None:
    st = lib().tsqr_set_leaf_path(dtype_code, path, gate)
    if st != TSQR_OK:
        raise TsqrError(st, "tsqr_set_leaf_path")


def leaf_stats(work_ptr: int, m: int, n: int, dtype_code: int) -> tuple:
    """(Gram superleaves, Householder superleaves, flat path taken) of the
    last call on work."""
    g, h, f = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    st = lib().tsqr_leaf_stats(work_ptr, m, n, dtype_code, ctypes.byref(g), ctypes.byref(h), ctypes.byref(f))
    if st != TSQR_OK:
        raise TsqrError(st, "tsqr_leaf_stats")
    return g.value, h.value, f.value


def set_mixed(gate: float = 1024.0, tol: float = 8e-7) -> None:
    st = lib().tsqr_set_mixed(gate, tol)
    if st != TSQR_OK:
        raise TsqrError(st, "tsqr_set_mixed")


def mixed_stats(work_ptr: int, m: int, n: int) -> dict:
    """Outcome of the last tsqr_mixed_s* call on work: CholeskyQR path taken
    (flat), passes, check status and the final ||Q^T Q - I||_F estimate."""
    f, p, c = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    r = ctypes.c_double(0.0)
    st = lib().tsqr_mixed_stats(work_ptr, m, n, ctypes.byref(f), ctypes.byref(p), ctypes.byref(c), ctypes.byref(r))
    if st != TSQR_OK:
        raise TsqrError(st, "tsqr_mixed_stats")
    b1, b2 = ctypes.c_int(0), ctypes.c_int(0)
    st = lib().tsqr_mixed_q_forms(work_ptr, m, n, ctypes.byref(b1), ctypes.byref(b2))
    if st != TSQR_OK:
        raise TsqrError(st, "tsqr_mixed_q_forms")
    return {"flat": f.value, "passes": p.value, "check": c.value, "residual": r.value,
            "bf16_q": [b1.value, b2.value]}


def status_string(status: int) -> str:
    try:
        return lib().tsqr_status_string(status).decode()
    except ImportError:
        return "unknown"


def tsqr_workspace_bytes(m: int, n: int, dtype_code: int) -> int:
    return int(lib().tsqr_workspace_bytes(m, n, dtype_code))


def qr_workspace_bytes(m: int, n: int, dtype_code: int) -> int:
    return int(lib().qr_workspace_bytes(m, n, dtype_code))


CHECK_KEYS = ("backward_fro", "orth_fro", "backward_2", "orth_2")


def check(m: int, n: int, dtype_code: int, A_ptr: int, lda: int, Q_ptr: int, ldq: int, R_ptr: int, ldr: int,
          stream_ptr: int = 0) -> dict:
    """Device-side FP64 check of a factorisation (tsqr_check_s/d): the
    relative backward error ||A - QR|| / ||A|| and the orthogonality
    ||I - Q^T Q||, Frobenius and spectral, all pointers on the device."""
    norms = (ctypes.c_double * 4)()
    fn = lib().tsqr_check_s if dtype_code == TSQR_FLOAT else lib().tsqr_check_d
    st = fn(m, n, A_ptr, lda, Q_ptr, ldq, R_ptr, ldr, norms, stream_ptr or None)
    if st != TSQR_OK:
        raise TsqrError(st, "tsqr_check")
    return dict(zip(CHECK_KEYS, list(norms)))
