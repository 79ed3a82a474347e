"""GPU tests of the mixed-precision recursive QR (qr_mixed_s*) against the
FP64 LAPACK oracle, with the condition-number sweep of the mixed-precision
configuration (A = U diag(s) V^T, s geometric from 1 to 1/cond) at a
reduced height.

Bar (kappa-independent, which is the point of the reorthogonalised
coupling and the final whole-matrix CholeskyQR pass): backward
||A-QR||_F/||A||_F <= 8 n u32 and ||I-Q^TQ||_F <= 1e-6 (the mixed-precision
configuration's target), R upper triangular, all_close(QR, A)
(utils.h:352-357).  The plain qr_s, one
coupling pass as in the reference (block_QR.h:88-93), is run beside it: its
loss of orthogonality grows with cond(A), which the sweep records.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import call

pytestmark = pytest.mark.gpu

U32 = O.unit_roundoff(np.float32)


def errors(A, Q, R):
    m = O.metrics(A, Q, R)
    return m


@pytest.mark.parametrize("n", [64, 128])
@pytest.mark.parametrize("cond", [1e2, 1e4, 1e6])
def test_qr_mixed_cond_sweep(cond, n):
    m = 262144
    A = O.gen_conditioned(m, n, cond, np.float32, seed=int(np.log10(cond)) * 10 + n)
    st, Q, R, _ = call("qr_mixed", A)
    assert st == 0
    mm = errors(A, Q, R)
    st, Qp, Rp, _ = call("qr", A)
    assert st == 0
    mp = errors(A, Qp, Rp)
    print(f"cond={cond:g} n={n}: mixed backward={mm['backward']:.3e} orth={mm['orth']:.3e} | "
          f"plain backward={mp['backward']:.3e} orth={mp['orth']:.3e}", flush=True)
    bound = 8 * n * U32
    assert mm["all_close"] and mm["r_lower_zero"]
    assert mm["backward"] <= bound
    assert mm["orth"] <= 1e-6
    assert mp["backward"] <= bound  # the plain path stays backward stable


@pytest.mark.parametrize("m,n", [(5000, 64), (3000, 96), (20000, 256), (4000, 33), (27648, 128)])
def test_qr_mixed_shapes(m, n):
    A = O.gen_normal(m, n, np.float32, seed=m + n)
    st, Q, R, _ = call("qr_mixed", A)
    assert st == 0
    mm = errors(A, Q, R)
    print(f"qr_mixed {m}x{n}: backward={mm['backward']:.3e} orth={mm['orth']:.3e}", flush=True)
    assert mm["all_close"] and mm["r_lower_zero"]
    assert mm["backward"] <= 8 * n * U32
    assert mm["orth"] <= 1e-6 * max(1, n // 32)


def test_qr_mixed_small_n_is_tsqr_mixed():
    A = O.gen_normal(20000, 32, np.float32, seed=3)
    _, Q1, R1, _ = call("qr_mixed", A)
    _, Q2, R2, _ = call("tsqr_mixed", A)
    assert np.array_equal(Q1, Q2) and np.array_equal(R1, R2)


def test_qr_mixed_reference_entry():
    A = O.gen_normal(27648, 128, np.float32, seed=5)
    _, Q1, R1, _ = call("qr_mixed", A)
    st, Q2, R2, _ = call("qr_mixed", A, ref_shaped=True, ldwork=2 * 108 * 256)
    assert st == 0
    assert np.array_equal(Q1, Q2) and np.array_equal(R1, R2)


def test_qr_mixed_r_and_coupling_against_fp64():
    """R against FP64 LAPACK after sign normalisation, and the first
    coupling block R12 against the FP64 product of the returned Q1 and the
    input A2 (the BF16x3 coupling products keep FP32 accuracy)."""
    A = O.gen_normal(50000, 128, np.float32, seed=21)
    st, Q, R, _ = call("qr_mixed", A)
    assert st == 0
    Qo, Ro = O.lapack_qr(A)
    kappa = np.linalg.cond(A.astype(np.float64))
    mm = O.metrics(A, Q, R, Ro, Qo)
    bound = 8 * 128 * U32
    print(f"qr_mixed 50000x128: r_err={mm['r_err']:.3e} kappa={kappa:.3f}", flush=True)
    assert mm["r_err"] <= bound * kappa * kappa
    Q1 = Q[:, :64].astype(np.float64)
    A2 = A[:, 64:].astype(np.float64)
    err = np.linalg.norm(R[:64, 64:] - Q1.T @ A2) / np.linalg.norm(A2)
    print(f"qr_mixed R12 vs Q1^T A2: {err:.3e}", flush=True)
    assert err <= 8 * 64 * U32
