"""GPU tests of the mixed-precision TSQR (tsqr_mixed_s*) through the C ABI.

Bar, against the FP64 LAPACK oracle on the same seeded input:
  backward error ||A-QR||_F/||A||_F  <= 8 n u32   (u32 = 2^-24)
  orthogonality  ||I-Q^TQ||_F        <= 1e-6      (the residual check's target)
  R vs FP64 LAPACK after sign normalisation <= 8 n u32 kappa
and the reference's all_close(QR, A) (utils.h:352-357).  The library's own
residual estimate must agree with the measured orthogonality.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import LAST_MIXED_STATS, call

pytestmark = pytest.mark.gpu

U32 = O.unit_roundoff(np.float32)


def check_mixed(A, Q, R, tag, orth_bound=1e-6):
    n = A.shape[1]
    Qo, Ro = O.lapack_qr(A)
    kappa = np.linalg.cond(A.astype(np.float64))
    m = O.metrics(A, Q, R, Ro, Qo)
    bound = 8 * n * U32
    print(f"{tag}: backward={m['backward']:.3e} orth={m['orth']:.3e} r_err={m['r_err']:.3e} "
          f"kappa={kappa:.2f} stats={LAST_MIXED_STATS}", flush=True)
    assert m["all_close"], tag
    assert m["r_lower_zero"], tag
    assert m["backward"] <= bound, tag
    assert m["orth"] <= orth_bound, tag
    assert m["r_err"] <= bound * kappa, tag
    return m


def tile_height(n):
    """Rows of one FP32 Householder tile: 512 bytes of registers per lane."""
    N = 8 if n <= 8 else (16 if n <= 16 else 32)
    return 64 * (512 // (N * 4))


MIXED_SHAPES = [
    (27648, 32),    # the reference harness shape, one-tile superleaves
    (100000, 32),
    (600000, 32),   # eight-tile superleaves
    (4096, 8),      # smallest batched height, n padded to 8
    (50000, 16),
    (5000, 20),     # n does not divide the tile
    (1000, 7),      # a single tile (1024 rows for n <= 8)
    (2000, 7),
    (3000, 1),
    (257, 32),      # a superleaf of one tile plus one row
]


@pytest.mark.parametrize("m,n", MIXED_SHAPES)
def test_mixed_shapes(m, n):
    A = O.gen_normal(m, n, np.float32, seed=m * 3 + n)
    st, Q, R, _ = call("tsqr_mixed", A)
    assert st == 0
    mt = check_mixed(A, Q, R, f"mixed {m}x{n}")
    if m <= tile_height(n):  # one tile: a single Householder QR
        assert LAST_MIXED_STATS["flat"] == 0
        return
    assert LAST_MIXED_STATS["flat"] == 1 and LAST_MIXED_STATS["check"] == 0
    # the library's residual estimate tracks the measured orthogonality
    assert LAST_MIXED_STATS["residual"] <= 8e-7
    assert abs(LAST_MIXED_STATS["residual"] - mt["orth"]) <= 0.5 * mt["orth"] + 2e-7
    if m >= 20 * n:
        # iid N(0,1), tall: kappa_est = ||R||_F ||R^-1||_F is within 2n, so Q1 = A R1^-1 is formed with
        # BF16x3 products on the BF16 matrix cores
        assert LAST_MIXED_STATS["bf16_q"][0] == 1


def test_mixed_single_tile_is_householder():
    A = O.gen_normal(200, 32, np.float32, seed=3)
    st, Q, R, _ = call("tsqr_mixed", A)
    assert st == 0
    check_mixed(A, Q, R, "mixed single tile")
    assert LAST_MIXED_STATS["flat"] == 0


def test_mixed_harness_input_and_reference_entry():
    A = O.gen_uniform(27648, 32, np.float32)
    st, Q, R, _ = call("tsqr_mixed", A)
    assert st == 0
    check_mixed(A, Q, R, "mixed cfg1")
    st2, Q2, R2, _ = call("tsqr_mixed", A, ref_shaped=True, ldwork=2 * 108 * 256)
    assert st2 == 0
    assert np.array_equal(Q, Q2) and np.array_equal(R, R2)


@pytest.mark.parametrize("cond", [3.0, 1e1, 1e2])
def test_mixed_conditioned_inputs_stay_on_cholesky_path(cond):
    A = O.gen_conditioned(100000, 32, cond, np.float32, seed=11)
    st, Q, R, _ = call("tsqr_mixed", A)
    assert st == 0
    check_mixed(A, Q, R, f"mixed cond={cond:g}")
    assert LAST_MIXED_STATS["flat"] == 1 and LAST_MIXED_STATS["check"] == 0
    if cond >= 1e2:
        # kappa_est above 2n: FP64 products form Q1 (the BF16x3 sums' error grows with kappa(R1^-1))
        assert LAST_MIXED_STATS["bf16_q"][0] == 0


@pytest.mark.parametrize("cond", [1e4, 1e6])
def test_mixed_ill_conditioned_escalates_to_fp64_gram(cond):
    """kappa_est above the BF16x3 gate: pass 1 is redone on the FP64 Gram
    (exact products) and Q1 is formed with FP64 products.  Q1 is then
    orthonormal to about u64 kappa^2: enough at cond 1e4, while at cond 1e6
    the second pass (BF16x3, R2^-1 close to the identity) restores it."""
    A = O.gen_conditioned(60000, 32, cond, np.float32, seed=12)
    st, Q, R, _ = call("tsqr_mixed", A)
    assert st == 0
    assert LAST_MIXED_STATS["flat"] == 1 and LAST_MIXED_STATS["bf16_q"][0] == 0
    assert LAST_MIXED_STATS["passes"] == (2 if cond >= 1e6 else 1)
    if LAST_MIXED_STATS["passes"] == 2:
        assert LAST_MIXED_STATS["bf16_q"][1] == 1
    check_mixed(A, Q, R, f"mixed cond={cond:g} FP64 Gram")


def test_mixed_escalation_tall_padded_width():
    """The escalation's FP64 Gram on 4 x 4 tiles (600000 rows) with padding
    columns in the last column group (n = 20)."""
    A = O.gen_conditioned(600000, 20, 1e6, np.float32, seed=18)
    st, Q, R, _ = call("tsqr_mixed", A)
    assert st == 0
    assert LAST_MIXED_STATS["flat"] == 1 and LAST_MIXED_STATS["bf16_q"][0] == 0
    check_mixed(A, Q, R, "mixed 600000x20 cond=1e6 FP64 Gram")


def test_mixed_beyond_fp64_gate_falls_back():
    """Numerically rank deficient in FP32 (cond 1e9): both gates reject,
    the Householder tree runs, with its kappa-independent bounds."""
    A = O.gen_conditioned(60000, 32, 1e9, np.float32, seed=12)
    st, Q, R, _ = call("tsqr_mixed", A)
    assert st == 0
    assert LAST_MIXED_STATS["flat"] == 0
    check_mixed(A, Q, R, "mixed cond=1e9 fallback", orth_bound=8 * 32 * U32)


def test_mixed_second_pass():
    """A tolerance no FP32 result can meet forces the second CholeskyQR
    pass: it must run, report the unmet check, and still deliver an
    orthogonal Q and a consistent R."""
    from tsqr_amd._lib import set_mixed

    A = O.gen_conditioned(80000, 32, 1e2, np.float32, seed=13)
    set_mixed(1024.0, 1e-12)
    try:
        st, Q, R, _ = call("tsqr_mixed", A)
    finally:
        set_mixed(1024.0, 8e-7)
    assert st == 0
    assert LAST_MIXED_STATS["flat"] == 1 and LAST_MIXED_STATS["passes"] == 2
    assert LAST_MIXED_STATS["check"] == 1
    # pass 1 on FP64 products (cond 1e2), pass 2 on BF16x3 (R2^-1 is close to the identity)
    assert LAST_MIXED_STATS["bf16_q"] == [0, 1]
    check_mixed(A, Q, R, "mixed forced second pass")


def test_mixed_second_pass_bf16_both():
    """Both Q passes on the BF16 matrix cores (iid input, forced second
    pass): the result meets the same bounds."""
    from tsqr_amd._lib import set_mixed

    A = O.gen_normal(120000, 32, np.float32, seed=17)
    set_mixed(1024.0, 1e-12)
    try:
        st, Q, R, _ = call("tsqr_mixed", A)
    finally:
        set_mixed(1024.0, 8e-7)
    assert st == 0
    assert LAST_MIXED_STATS["passes"] == 2 and LAST_MIXED_STATS["bf16_q"] == [1, 1]
    check_mixed(A, Q, R, "mixed bf16 both passes")


def test_mixed_deterministic():
    A = O.gen_normal(300000, 32, np.float32, seed=14)
    _, Q1, R1, _ = call("tsqr_mixed", A)
    _, Q2, R2, _ = call("tsqr_mixed", A)
    assert np.array_equal(Q1, Q2) and np.array_equal(R1, R2)


def test_mixed_padded_lda():
    A = O.gen_uniform(9000, 32, np.float32, seed=9)
    st, Q, R, pad_ok = call("tsqr_mixed", A, lda=9000 + 37)
    assert st == 0 and pad_ok
    check_mixed(A, Q, R, "mixed lda+37")
