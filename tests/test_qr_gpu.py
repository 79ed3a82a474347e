"""GPU parity tests of the recursive blocked QR (qr_{s,d}) through the C ABI.

The recursion couples column halves by one pass of block classical
Gram-Schmidt (block_QR.h:88-93), so orthogonality between the halves degrades
like u*kappa(A); the inputs here are N(0,1) (kappa ~ 1.2 for m >> n).  Bars:
backward and orthogonality <= 8 n u * kappa, R vs FP64 LAPACK after sign
normalisation <= 8 n u * kappa^2, R21 exactly zero, and R12 = Q1^T A2 against
an FP64 product of the returned Q1 and the input A2.  The 27648 x 128
configuration uses its pinned N(0,1) input (oracle.gen_normal_stream).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import call

pytestmark = pytest.mark.gpu


def check_qr(A, Q, R, tag):
    u = O.unit_roundoff(A.dtype)
    m_, n = A.shape
    kappa = np.linalg.cond(A.astype(np.float64))
    Qo, Ro = O.lapack_qr(A)
    m = O.metrics(A, Q, R, Ro, Qo)
    bound = 8 * n * u * kappa
    print(f"{tag}: backward={m['backward']:.3e} orth={m['orth']:.3e} r_err={m['r_err']:.3e} "
          f"kappa={kappa:.2f} bound={bound:.2e}", flush=True)
    assert m["all_close"], tag
    assert m["backward"] <= bound, tag
    assert m["orth"] <= bound, tag
    assert m["r_err"] <= bound * kappa, tag
    assert m["r_lower_zero"], tag
    return m


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_cfg2_27648x128(dtype):
    # the configuration's pinned input: std::mt19937_64(0) + Box-Muller
    # (tests/golden/normal_stream.json)
    A = O.gen_normal_stream(27648, 128, dtype, seed=0)
    st, Q, R, _ = call("qr", A)
    assert st == 0
    check_qr(A, Q, R, f"qr 27648x128 {np.dtype(dtype).name}")
    # coupling block of the top split: R12 = Q1^T A2
    R12 = Q[:, :64].astype(np.float64).T @ A[:, 64:].astype(np.float64)
    err = np.abs(R[:64, 64:] - R12).max() / np.abs(R12).max()
    assert err <= 64 * 8 * O.unit_roundoff(dtype) * 10, err
    # the reference restatement (block_QR.h) agrees after sign normalisation
    _, Rr = O.qr_ref(A, bs=256, num_sm=108)
    _, Rn = O.sign_normalise(None, R)
    _, Rrn = O.sign_normalise(None, Rr)
    assert np.linalg.norm(Rn - Rrn) / np.linalg.norm(Rrn) <= 8 * 128 * O.unit_roundoff(dtype) * 10


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("m,n", [(5000, 64), (3000, 96), (2048, 40), (4000, 33), (20000, 256)])
def test_qr_shapes(m, n, dtype):
    A = O.gen_normal(m, n, dtype, seed=m + n)
    st, Q, R, _ = call("qr", A)
    assert st == 0
    check_qr(A, Q, R, f"qr {m}x{n} {np.dtype(dtype).name}")


def test_qr_small_n_is_tsqr():
    A = O.gen_normal(10000, 24, np.float32, seed=3)
    st1, Q1, R1, _ = call("qr", A)
    st2, Q2, R2, _ = call("tsqr", A)
    assert st1 == 0 and st2 == 0
    assert np.array_equal(Q1, Q2) and np.array_equal(R1, R2)
