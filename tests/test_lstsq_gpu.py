"""Least squares (tsqr_lstsq_*) and Q^T B (tsqr_apply_qt_*) on the implicit Q
of the R-only mode, through the C ABI, against FP64 LAPACK on the host.

Bar: the solution's relative error is at most c n u kappa (1 + kappa eta),
eta = ||r|| / (||A|| ||x||) (the first-order sensitivity of least squares,
which the corrected seminormal equations attain for kappa u << 1), with
c = 8; the returned residual matches B - A X computed in FP64 from the
returned X to 8 n u (1 + ||A||_F ||X||_F / ||B||_F) (two working-precision
updates); Q^T B matches the FP64 product of the explicit Q that tsqr
returns for the same A to 8 n u kappa."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import from_device_colmajor, to_device_colmajor

pytestmark = pytest.mark.gpu


def _lstsq(A, B):
    import torch

    import tsqr_amd
    from tsqr_amd._lib import TSQR_DOUBLE, TSQR_FLOAT

    L = tsqr_amd.lib()
    m, n = A.shape
    k = B.shape[1]
    code = TSQR_FLOAT if A.dtype == np.float32 else TSQR_DOUBLE
    suf = "s" if code == TSQR_FLOAT else "d"
    bufA, pA, lda = to_device_colmajor(A)
    keepA = bufA.clone()
    bufB, pB, ldb = to_device_colmajor(B)
    bufX, pX, ldx = to_device_colmajor(np.zeros((n, k), A.dtype))
    need = L.tsqr_solve_workspace_bytes(m, n, k, code)
    ws = torch.empty(need, dtype=torch.uint8, device="cuda")
    st = getattr(L, f"tsqr_lstsq_{suf}_stream")(m, n, k, pA, lda, pB, ldb, pX, ldx, ws.data_ptr(), need, None)
    torch.cuda.synchronize()
    assert torch.equal(torch.nan_to_num(bufA, nan=3.0), torch.nan_to_num(keepA, nan=3.0)), "A was written"
    return st, from_device_colmajor(bufX, n, k, ldx), from_device_colmajor(bufB, m, k, ldb)


def _problem(m, n, k, dtype, cond, noise, seed):
    rng = np.random.default_rng(seed)
    if cond <= 10:
        A = rng.standard_normal((m, n))
    else:
        U, _ = np.linalg.qr(rng.standard_normal((m, n)))
        V, _ = np.linalg.qr(rng.standard_normal((n, n)))
        A = (U * np.geomspace(1.0, 1.0 / cond, n)) @ V.T
    A = A.astype(dtype)
    Xt = rng.standard_normal((n, k))
    B = A.astype(np.float64) @ Xt + noise * rng.standard_normal((m, k))
    return A, B.astype(dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("m,n,k,cond,noise", [
    (50000, 32, 1, 1, 0.0),
    (50000, 32, 3, 1, 1e-2),
    (20000, 20, 70, 1, 1.0),
    (120000, 16, 5, 1e3, 1e-3),
    (9000, 7, 2, 1e2, 1e-1),
])
def test_lstsq_against_lapack(m, n, k, cond, noise, dtype):
    A, B = _problem(m, n, k, dtype, cond, noise, seed=m + n + k)
    st, X, E = _lstsq(A, B)
    assert st == 0
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    Xr = np.linalg.lstsq(A64, B64, rcond=None)[0]
    kappa = np.linalg.cond(A64)
    r = B64 - A64 @ Xr
    eta = np.linalg.norm(r) / (np.linalg.norm(A64, 2) * np.linalg.norm(Xr))
    u = O.unit_roundoff(dtype)
    bound = 8 * n * u * kappa * (1 + kappa * eta)
    err = np.linalg.norm(X - Xr) / np.linalg.norm(Xr)
    # the returned residual is B - A X for the returned X
    res_err = np.linalg.norm(E - (B64 - A64 @ X.astype(np.float64))) / np.linalg.norm(B64)
    print(f"lstsq {m}x{n} k={k} {np.dtype(dtype).name} cond={kappa:.1f} eta={eta:.1e}: err={err:.3e} "
          f"bound={bound:.2e} residual={res_err:.2e}", flush=True)
    assert err <= bound
    # the residual is formed by two working-precision updates B - A X1 - A D
    assert res_err <= 8 * n * u * (1.0 + np.linalg.norm(A64) * np.linalg.norm(X) / np.linalg.norm(B64))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_apply_qt_matches_explicit_q(dtype):
    import torch

    import tsqr_amd
    from tests.gpu_util import call
    from tsqr_amd._lib import TSQR_DOUBLE, TSQR_FLOAT

    L = tsqr_amd.lib()
    m, n, k = 40000, 32, 9
    rng = np.random.default_rng(4)
    A = rng.standard_normal((m, n)).astype(dtype)
    Bm = rng.standard_normal((m, k)).astype(dtype)
    code = TSQR_FLOAT if dtype == np.float32 else TSQR_DOUBLE
    suf = "s" if code == TSQR_FLOAT else "d"
    st, Q, Rfull, _ = call("tsqr", A)
    assert st == 0
    bufA, pA, lda = to_device_colmajor(A)
    R = torch.zeros((n * n,), dtype=bufA.dtype, device="cuda")
    need_r = L.tsqr_workspace_bytes(m, n, code)
    ws_r = torch.empty(need_r, dtype=torch.uint8, device="cuda")
    assert getattr(L, f"tsqr_r_{suf}_stream")(m, n, pA, lda, R.data_ptr(), n, ws_r.data_ptr(), need_r, None) == 0
    bufB, pB, ldb = to_device_colmajor(Bm)
    bufC, pC, ldc = to_device_colmajor(np.zeros((n, k), dtype))
    need = L.tsqr_solve_workspace_bytes(m, n, k, code)
    ws = torch.empty(need, dtype=torch.uint8, device="cuda")
    st = getattr(L, f"tsqr_apply_qt_{suf}_stream")(m, n, k, pA, lda, R.data_ptr(), n, pB, ldb, pC, ldc,
                                                   ws.data_ptr(), need, None)
    torch.cuda.synchronize()
    assert st == 0
    C = from_device_colmajor(bufC, n, k, ldc)
    Cr = Q.astype(np.float64).T @ Bm.astype(np.float64)
    kappa = np.linalg.cond(A.astype(np.float64))
    err = np.linalg.norm(C - Cr) / np.linalg.norm(Cr)
    print(f"apply_qt {np.dtype(dtype).name}: err={err:.3e}", flush=True)
    assert err <= 8 * n * O.unit_roundoff(dtype) * kappa


def test_solve_rejects_bad_arguments():
    import tsqr_amd

    L = tsqr_amd.lib()
    assert L.tsqr_solve_workspace_bytes(10, 33, 1, 1) == 0
    assert L.tsqr_lstsq_s_stream(100, 33, 1, 1, 100, 1, 100, 1, 33, 1, 1 << 20, None) != 0  # n > 32
    assert L.tsqr_lstsq_d_stream(100, 8, 0, 1, 100, 1, 100, 1, 8, 1, 1 << 20, None) != 0  # nrhs < 1
    assert L.tsqr_lstsq_s_stream(100, 8, 2, 1, 100, 1, 100, 1, 8, None, 0, None) != 0  # no workspace
    assert L.tsqr_apply_qt_s_stream(100, 8, 2, 1, 100, None, 8, 1, 100, 1, 8, 1, 1 << 20, None) != 0  # no R
