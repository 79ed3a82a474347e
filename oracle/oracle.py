"""CPU oracle for the TSQR / recursive QR path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
import this module, and only as the checker or as the timed CPU baseline.
The HIP library (tsqr_amd) never imports, links or falls back to it.

Contents
  gen_uniform(m, n, dtype, seed)  the reference harness input stream
                                  (test_tsqr.cu:24-28), column-major
  gen_normal(m, n, seed)          seeded N(0,1) inputs for the general tests
                                  (numpy MT19937 + standard_normal; no
                                  reference counterpart)
  gen_normal_stream(m, n, dtype)  the recursive-QR configuration's N(0,1)
                                  input: std::mt19937_64(0) + Box-Muller in
                                  FP64, column-major (C, pinned by
                                  tests/golden/normal_stream.json)
  tsqr_ref(A, bs, num_sm)         C restatement of tsqr<T> (TallShinnyQR.h:86-149,
                                  kernelQR.h:69-497), float32 or float64
  qr_ref(A, bs, num_sm, nb)       C restatement of qr<T> (block_QR.h:72-95)
  qr_tile_ref(A)                  C restatement of one qr_kernel tile
                                  (kernelQR.h:69-291)
  lapack_qr(A)                    FP64 LAPACK dgeqrf + dorgqr (numpy)
  metrics(...)                    the reference's checks (utils.h:317-455):
                                  all_close(QR, A) with abs 1e-4 / rel 1e-5,
                                  spectral |A-QR|_2/|A|_2 and |I-Q^TQ|_2/n,
                                  plus Frobenius backward error and
                                  orthogonality and R / Q against LAPACK after
                                  diagonal-sign normalisation.

Parity pins (tests/test_oracle.py): generator fixtures from libstdc++
(tests/golden/uniform_stream.json), closed-form reflector KATs, LAPACK.
Not pinned: the reference's exact rounding (warp-butterfly summation order,
kernelQR.h:110-122), which only an NVIDIA GPU run could show.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build() -> str:
    """Compiles liboracle.so with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, u32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_uint32
        for suf in ("s", "d"):
            getattr(L, f"oracle_gen_uniform_{suf}").argtypes = [vp, ctypes.c_int64, ctypes.c_int64, u32]
            getattr(L, f"oracle_gen_uniform_{suf}").restype = None
            getattr(L, f"oracle_gen_normal_{suf}").argtypes = [vp, ctypes.c_int64, ctypes.c_uint64]
            getattr(L, f"oracle_gen_normal_{suf}").restype = None
            f = getattr(L, f"oracle_qr_tile_{suf}")
            f.argtypes = [i32, i32, vp, i64, vp, i64, vp, i64]
            f.restype = None
            f = getattr(L, f"oracle_tsqr_{suf}")
            f.argtypes = [i32, i32, vp, i64, vp, i64, i32, i32]
            f.restype = i32
            f = getattr(L, f"oracle_qr_{suf}")
            f.argtypes = [i32, i32, vp, i64, vp, i64, i32, i32, i32]
            f.restype = i32
        L.oracle_mt64_raw.argtypes = [vp, ctypes.c_int64, ctypes.c_uint64]
        L.oracle_mt64_raw.restype = None
        L.oracle_fnv1a64.argtypes = [vp, ctypes.c_int64]
        L.oracle_fnv1a64.restype = ctypes.c_uint64
        _lib = L
    return _lib


def _suf(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return "s"
    if dt == np.float64:
        return "d"
    raise TypeError(f"unsupported dtype {dt}")


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------------------
# inputs
# ---------------------------------------------------------------------------

def gen_uniform_flat(count: int, dtype=np.float32, seed: int = 0, first: int = 0) -> np.ndarray:
    """Values first..first+count of the harness stream (test_tsqr.cu:24-28)."""
    out = np.empty(count, dtype=dtype)
    getattr(lib(), f"oracle_gen_uniform_{_suf(dtype)}")(_ptr(out), first, count, seed)
    return out


def gen_uniform(m: int, n: int, dtype=np.float32, seed: int = 0) -> np.ndarray:
    """m x n matrix filled column-major from the harness stream; returned as a
    Fortran-ordered array (A[i, j] = stream[i + j*m])."""
    flat = gen_uniform_flat(m * n, dtype, seed)
    return flat.reshape(n, m).T


def gen_normal(m: int, n: int, dtype=np.float32, seed: int = 0) -> np.ndarray:
    """Seeded iid N(0,1), Fortran order (no reference counterpart)."""
    rng = np.random.Generator(np.random.MT19937(seed))
    return np.asfortranarray(rng.standard_normal((m, n)).astype(dtype))


def mt64_raw(count: int, seed: int = 0) -> np.ndarray:
    """The first `count` outputs of std::mt19937_64(seed) (C restatement)."""
    out = np.empty(count, dtype=np.uint64)
    lib().oracle_mt64_raw(_ptr(out), count, seed)
    return out


def fnv1a64(a: np.ndarray) -> str:
    """FNV-1a-64 over the bytes of a contiguous array, as hex (the fixtures'
    array hash)."""
    a = np.ascontiguousarray(a)
    return "%016x" % lib().oracle_fnv1a64(_ptr(a), a.nbytes)


def gen_normal_stream(m: int, n: int, dtype=np.float32, seed: int = 0) -> np.ndarray:
    """m x n N(0,1) matrix of the recursive-QR configuration, filled
    column-major from std::mt19937_64(seed) by Box-Muller in FP64 (float32:
    the FP64 values rounded); Fortran-ordered."""
    out = np.empty(m * n, dtype=dtype)
    getattr(lib(), f"oracle_gen_normal_{_suf(dtype)}")(_ptr(out), m * n, seed)
    return out.reshape(n, m).T


def gen_conditioned(m: int, n: int, cond: float, dtype=np.float32, seed: int = 0) -> np.ndarray:
    """A = U diag(s) V^T with U, V orthonormal from seeded Gaussians and s
    geometrically spaced from 1 down to 1/cond."""
    rng = np.random.Generator(np.random.MT19937(seed))
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = cond ** (-np.arange(n) / max(n - 1, 1))
    return np.asfortranarray(((U * s) @ V.T).astype(dtype))


# ---------------------------------------------------------------------------
# restatement of the reference algorithm (C, liboracle.so)
# ---------------------------------------------------------------------------

def qr_tile_ref(A: np.ndarray):
    """One qr_kernel tile (kernelQR.h:69-291): returns (Q h x n, R n x n)."""
    A = np.asfortranarray(A)
    h, n = A.shape
    Q = np.zeros((h, n), dtype=A.dtype, order="F")
    R = np.zeros((n, n), dtype=A.dtype, order="F")
    getattr(lib(), f"oracle_qr_tile_{_suf(A.dtype)}")(h, n, _ptr(A), h, _ptr(Q), h, _ptr(R), n)
    return Q, R


def tsqr_ref(A: np.ndarray, bs: int = 256, num_sm: int = 108):
    """tsqr<T> (TallShinnyQR.h:86-149) with NUM_SM=num_sm, BLOCK_SIZE=bs.
    Returns (Q, R); raises ValueError outside the reference's domain."""
    Q = np.array(A, dtype=A.dtype, order="F", copy=True)
    m, n = Q.shape
    R = np.zeros((n, n), dtype=A.dtype, order="F")
    st = getattr(lib(), f"oracle_tsqr_{_suf(A.dtype)}")(m, n, _ptr(Q), m, _ptr(R), n, bs, num_sm)
    if st != 0:
        raise ValueError(f"shape {m}x{n} outside the reference domain (status {st})")
    return Q, R


def qr_ref(A: np.ndarray, bs: int = 256, num_sm: int = 108, nb: int = 32):
    """qr<T> (block_QR.h:72-95).  Returns (Q, R) with R21 = 0."""
    Q = np.array(A, dtype=A.dtype, order="F", copy=True)
    m, n = Q.shape
    R = np.zeros((n, n), dtype=A.dtype, order="F")
    st = getattr(lib(), f"oracle_qr_{_suf(A.dtype)}")(m, n, _ptr(Q), m, _ptr(R), n, bs, num_sm, nb)
    if st != 0:
        raise ValueError(f"shape {m}x{n} outside the reference domain (status {st})")
    return Q, R


# ---------------------------------------------------------------------------
# FP64 LAPACK and metrics
# ---------------------------------------------------------------------------

def lapack_qr(A: np.ndarray):
    """FP64 LAPACK dgeqrf + dorgqr (numpy.linalg.qr, reduced mode)."""
    return np.linalg.qr(np.asarray(A, dtype=np.float64), mode="reduced")


def lapack_r(A: np.ndarray) -> np.ndarray:
    return np.linalg.qr(np.asarray(A, dtype=np.float64), mode="r")


def sign_normalise(Q, R):
    """Returns (Q D, D R) with D = diag(sign(R_ii)) (sign(0) = +1)."""
    d = np.where(np.diag(R) < 0, -1.0, 1.0)
    Qn = None if Q is None else np.asarray(Q, dtype=np.float64) * d[None, :]
    return Qn, np.asarray(R, dtype=np.float64) * d[:, None]


def _spectral(M: np.ndarray) -> float:
    """||M||_2 from the n x n Gram matrix (M is m x n, m >= n)."""
    G = M.T @ M
    return float(np.sqrt(max(np.linalg.eigvalsh(G)[-1], 0.0)))


def all_close(X, Xref, abs_tol=1e-4, rel_tol=1e-5) -> bool:
    """utils.h:147-179: |X - Xref| <= max(abs_tol, |Xref|*rel_tol), no NaN."""
    d = np.abs(np.asarray(X, np.float64) - np.asarray(Xref, np.float64))
    thr = np.maximum(abs_tol, np.abs(np.asarray(Xref, np.float64)) * rel_tol)
    return bool(np.all(d <= thr) and not np.isnan(d).any())


def metrics(A, Q, R, R_oracle=None, Q_oracle=None) -> dict:
    """Error figures of one factorisation, all in FP64 on the host.

    backward      ||A - QR||_F / ||A||_F
    orth          ||I - Q^T Q||_F
    spec_backward ||A - QR||_2 / ||A||_2        (utils.h:380, printed there)
    spec_orth     ||I - Q^T Q||_2 / n           (utils.h:380, printed there)
    all_close     utils.h:352-357 criterion on QR vs A
    r_err         ||D_g R - D_o R_o||_F / ||R_o||_F   (if R_oracle given)
    q_err         ||Q D_g - Q_o D_o||_F / sqrt(n)      (if Q_oracle given)
    """
    A = np.asarray(A, np.float64)
    Q = np.asarray(Q, np.float64)
    R = np.asarray(R, np.float64)
    n = A.shape[1]
    QR = Q @ R
    E = A - QR
    I_QtQ = np.eye(n) - Q.T @ Q
    out = {
        "backward": float(np.linalg.norm(E) / np.linalg.norm(A)),
        "orth": float(np.linalg.norm(I_QtQ)),
        "spec_backward": _spectral(E) / _spectral(A),
        "spec_orth": float(np.linalg.norm(I_QtQ, 2)) / n,
        "all_close": all_close(QR, A),
        "r_lower_zero": bool(np.all(np.tril(R, -1) == 0)),
    }
    if R_oracle is not None:
        _, Rn = sign_normalise(None, R)
        _, Ron = sign_normalise(None, R_oracle)
        out["r_err"] = float(np.linalg.norm(Rn - Ron) / np.linalg.norm(Ron))
        if Q_oracle is not None:
            Qn, _ = sign_normalise(Q, R)
            Qon, _ = sign_normalise(Q_oracle, R_oracle)
            out["q_err"] = float(np.linalg.norm(Qn - Qon) / np.sqrt(n))
    return out


def unit_roundoff(dtype) -> float:
    return float(np.finfo(np.dtype(dtype)).eps) / 2.0
