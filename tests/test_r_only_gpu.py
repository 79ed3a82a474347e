"""R-only entry points (tsqr_r_s / tsqr_r_d) through the C ABI.

For every leaf path and shape: A is left bit for bit as it was, and R
matches the FP64 LAPACK oracle after sign normalisation to 8 n u kappa (the
bar of tests/test_tsqr_gpu.py).  Where the R-only sequence runs the same
kernels as the full factorisation up to the tree root (every path except
CholeskyQR2: FP64 streams the second pass's Gram matrix of A R1^-1 instead
of the stored Q1, and FP32 R only takes the flat FP64 Gram path, which needs
no second pass), R is bit-identical to tsqr_s / tsqr_d's R."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import LeafPath, call, to_device_colmajor, torch_dtype

pytestmark = pytest.mark.gpu

AUTO, HOUSEHOLDER, GRAM_TREE, CHOLQR2 = 0, 1, 2, 3


def r_only(A, lda=None, ref_shaped=False):
    import torch

    import tsqr_amd
    from tsqr_amd._lib import TSQR_FLOAT

    L = tsqr_amd.lib()
    m, n = A.shape
    code = TSQR_FLOAT if A.dtype == np.float32 else 2
    suf = "s" if code == TSQR_FLOAT else "d"
    buf, pA, lda = to_device_colmajor(A, lda)
    before = buf.clone()
    R = torch.full((n * n,), float("nan"), dtype=torch_dtype(A.dtype), device="cuda")
    if ref_shaped:
        st = getattr(L, f"tsqr_r_{suf}")(m, n, pA, lda, R.data_ptr(), n, None, 0)
    else:
        need = L.tsqr_workspace_bytes(m, n, code)
        ws = torch.empty(max(need, 1), dtype=torch.uint8, device="cuda")
        st = getattr(L, f"tsqr_r_{suf}_stream")(m, n, pA, lda, R.data_ptr(), n, ws.data_ptr(), need, None)
    torch.cuda.synchronize()
    untouched = bool(torch.equal(torch.nan_to_num(buf, nan=7.0), torch.nan_to_num(before, nan=7.0)))
    return st, R.view(n, n).cpu().numpy().T.copy(), untouched


def check_r(A, R, tag):
    n = A.shape[1]
    u = O.unit_roundoff(A.dtype)
    _, Ro = O.lapack_qr(A)
    kappa = np.linalg.cond(A.astype(np.float64))
    _, Rn = O.sign_normalise(None, R)
    _, Ron = O.sign_normalise(None, Ro)
    err = np.linalg.norm(Rn - Ron) / np.linalg.norm(Ron)
    bound = 8 * max(n, 32) * u * kappa
    print(f"{tag}: r_err={err:.3e} bound={bound:.2e}", flush=True)
    assert np.all(np.tril(R, -1) == 0), tag
    assert err <= bound, tag


SHAPES = [(100, 3), (256, 32), (257, 32), (5000, 20), (12345, 16), (70000, 32), (600000, 32)]


@pytest.mark.parametrize("path", [AUTO, HOUSEHOLDER, GRAM_TREE, CHOLQR2])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("m,n", SHAPES)
def test_r_only_matches_full_and_oracle(m, n, dtype, path):
    A = np.random.default_rng(m + 7 * n).uniform(-1, 1, (m, n)).astype(dtype)
    with LeafPath(dtype, path):
        st, R, untouched = r_only(A)
        assert st == 0
        st2, _, Rfull, _ = call("tsqr", A)
        assert st2 == 0
    tag = f"r_only {m}x{n} {np.dtype(dtype).name} path={path}"
    assert untouched, tag + ": A was written"
    check_r(A, R, tag)
    if path == CHOLQR2 and dtype == np.float64:
        assert np.linalg.norm(R - Rfull) <= 1e-12 * np.linalg.norm(Rfull), tag
    elif path == CHOLQR2:
        pass  # FP32 R only runs the flat FP64 Gram; the full call the BF16x3 mixed method
    else:
        assert np.array_equal(R, Rfull), tag + ": R differs from the full factorisation's"


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_r_only_ill_conditioned_falls_back(dtype):
    # kappa = 1e7: above both flat gates, so the Householder tree gives R
    m, n = 40000, 32
    rng = np.random.default_rng(3)
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.geomspace(1.0, 1e-7, n)
    A = ((U * s) @ V.T).astype(dtype)
    st, R, untouched = r_only(A)
    assert st == 0 and untouched
    _, Ro = O.lapack_qr(A)
    _, Rn = O.sign_normalise(None, R)
    _, Ron = O.sign_normalise(None, Ro)
    # Householder R is backward stable, so its normwise error is at most
    # about 8 n u kappa (first-order perturbation bound of the R factor)
    assert np.linalg.norm(Rn - Ron) / np.linalg.norm(Ron) <= 8 * n * O.unit_roundoff(dtype) * 1e7


def test_r_only_reference_shaped_entry_and_unaligned_lda():
    A = np.random.default_rng(9).standard_normal((30001, 32)).astype(np.float32)
    st, R, untouched = r_only(A, lda=30005, ref_shaped=True)
    assert st == 0 and untouched
    check_r(A, R, "r_only reference-shaped, lda 30005")


def test_r_only_rejects_bad_arguments():
    import tsqr_amd

    L = tsqr_amd.lib()
    assert L.tsqr_r_s(10, 20, 1, 10, 1, 20, None, 0) != 0  # m < n
    assert L.tsqr_r_d_stream(100, 33, 1, 100, 1, 33, None, 0, None) != 0  # n > 32
    assert L.tsqr_r_s_stream(1000, 32, 1, 1000, 1, 32, None, 0, None) != 0  # no workspace
