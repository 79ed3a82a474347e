"""GPU parity tests of tsqr_{s,d} through the C ABI against the oracle.

Bar (per result, stated in each assertion): with u the unit roundoff of the
working precision and n the column count,
  backward error ||A-QR||_F/||A||_F        <= 8 n u
  orthogonality  ||I-Q^TQ||_F              <= 8 n u
(n is taken as at least 32 for FP64 on the Gram leaf path, whose Gram matrix
has no wider type to accumulate in: its errors grow like sqrt(leaf rows) u
rather than n u)
  R vs FP64 LAPACK after sign normalisation <= 8 n u kappa_est (random inputs
  are well conditioned, kappa_est = cond_2(A) from the FP64 SVD)
and the reference's own all_close(QR, A) (abs 1e-4, rel 1e-5, utils.h:352-357).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import LAST_LEAF_STATS, LeafPath, call

pytestmark = pytest.mark.gpu

AUTO, HOUSEHOLDER, GRAM_TREE, CHOLQR2 = 0, 1, 2, 3
PATHS = [pytest.param(AUTO, id="auto"), pytest.param(HOUSEHOLDER, id="householder"),
         pytest.param(GRAM_TREE, id="gram_tree"), pytest.param(CHOLQR2, id="cholqr2")]


def check(A, Q, R, tag, c=8.0, n_floor=1):
    u = O.unit_roundoff(A.dtype)
    n = max(A.shape[1], n_floor)
    Qo, Ro = O.lapack_qr(A)
    kappa = np.linalg.cond(A.astype(np.float64))
    m = O.metrics(A, Q, R, Ro, Qo)
    bound = c * n * u
    print(f"{tag}: backward={m['backward']:.3e} orth={m['orth']:.3e} r_err={m['r_err']:.3e} "
          f"q_err={m['q_err']:.3e} kappa={kappa:.2f} bound={bound:.2e}", flush=True)
    assert m["all_close"], tag
    assert m["r_lower_zero"], tag
    assert m["backward"] <= bound, tag
    assert m["orth"] <= bound, tag
    assert m["r_err"] <= bound * kappa, tag
    assert m["q_err"] <= bound * kappa * 4, tag
    return m


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_cfg1_harness_input(dtype, path):
    """m=27648, n=32 with the reference harness's input stream."""
    A = O.gen_uniform(27648, 32, dtype)
    with LeafPath(dtype, path):
        st, Q, R, _ = call("tsqr", A)
    assert st == 0
    check(A, Q, R, f"cfg1 {np.dtype(dtype).name} path={path}",
          n_floor=32 if (dtype == np.float64 and path != HOUSEHOLDER) else 1)
    # the reference's restatement at its A800 tuning agrees after sign normalisation
    Qr, Rr = O.tsqr_ref(A, bs=256, num_sm=108)
    _, Rn = O.sign_normalise(None, R)
    _, Rrn = O.sign_normalise(None, Rr)
    assert np.linalg.norm(Rn - Rrn) / np.linalg.norm(Rrn) <= 8 * 32 * O.unit_roundoff(dtype)


SHAPES = [
    (32, 32),      # square, single tile
    (100, 3),      # n padded to 8
    (255, 32),
    (256, 32),     # exactly one tile
    (257, 32),     # second tile of 1 row (< n)
    (1000, 7),
    (5000, 20),    # arity 12 (n does not divide the tile)
    (12345, 16),
    (3000, 1),
    (70000, 32),
    (600000, 32),  # five tree levels
]


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("m,n", SHAPES)
def test_shapes(m, n, dtype, path):
    A = O.gen_normal(m, n, dtype, seed=m * 7 + n)
    with LeafPath(dtype, path):
        st, Q, R, _ = call("tsqr", A)
    assert st == 0
    floor = 32 if (dtype == np.float64 and path != HOUSEHOLDER) else 1
    check(A, Q, R, f"{m}x{n} {np.dtype(dtype).name} path={path} leaves={LAST_LEAF_STATS}", n_floor=floor)


@pytest.mark.parametrize("path", [pytest.param(AUTO, id="auto"), pytest.param(CHOLQR2, id="cholqr2")])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("m,n", [(600000, 5), (600000, 13), (600000, 20), (524301, 32), (524301, 20)])
def test_tall_padded_widths(m, n, dtype, path):
    """From 524,288 rows the flat paths run on 4 x 4 FP64 tiles (the R pass,
    and the Q pass of FP32 AUTO and of FP64 CholeskyQR2): padded widths 8
    and 32, padding columns inside the last column group, and a short last
    chunk (524301 = 64 * 8192 + 13 rows)."""
    A = O.gen_normal(m, n, dtype, seed=n * 11 + m % 7)
    with LeafPath(dtype, path):
        st, Q, R, _ = call("tsqr", A)
    assert st == 0
    assert LAST_LEAF_STATS["flat"] == 1
    floor = 32 if dtype == np.float64 else 1
    check(A, Q, R, f"{m}x{n} {np.dtype(dtype).name} path={path} leaves={LAST_LEAF_STATS}", n_floor=floor)


@pytest.mark.parametrize("path", [pytest.param(AUTO, id="auto"), pytest.param(CHOLQR2, id="cholqr2")])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tall_unaligned_lda(dtype, path):
    """The 4 x 4-tile passes without 16-byte loads: lda = m + 37 (the
    columns start at every alignment), short last chunk, padded width."""
    m, n = 524301, 20
    A = O.gen_normal(m, n, dtype, seed=23)
    with LeafPath(dtype, path):
        st, Q, R, pad_ok = call("tsqr", A, lda=m + 37)
    assert st == 0 and pad_ok
    assert LAST_LEAF_STATS["flat"] == 1
    floor = 32 if dtype == np.float64 else 1
    check(A, Q, R, f"{m}x{n} lda+37 {np.dtype(dtype).name} path={path}", n_floor=floor)


@pytest.mark.parametrize("path", [pytest.param(AUTO, id="auto"), pytest.param(CHOLQR2, id="cholqr2")])
@pytest.mark.parametrize("m,n,dtype", [
    (65536, 32, np.float32), (65600, 32, np.float32), (65536, 13, np.float32),
    (32768, 32, np.float64), (32800, 32, np.float64), (32768, 20, np.float64)])
def test_short_r_pass_boundary(m, n, dtype, path):
    """Either side of the ticket-free R pass on 16-wave workgroups (64
    workgroups of 16 chunks: 65536 FP32 rows, 32768 FP64 rows): the short
    form and the ticketed streaming form both meet the bounds."""
    A = O.gen_uniform(m, n, dtype, seed=m + n)
    with LeafPath(dtype, path):
        st, Q, R, _ = call("tsqr", A)
    assert st == 0 and LAST_LEAF_STATS["flat"] == 1
    floor = 32 if dtype == np.float64 else 1
    check(A, Q, R, f"{m}x{n} {np.dtype(dtype).name} path={path}", n_floor=floor)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_short_calls_reuse_one_workspace(dtype):
    """Short calls set up their workspace inside the R pass's launch: on one
    workspace, calls that take the flat path and calls whose gate sends them
    to the Householder fallback alternate, and each starts from a clean
    set-up (tickets, flags, list and control words of the previous call
    cleared)."""
    import torch

    import tsqr_amd
    from tsqr_amd._lib import TSQR_DOUBLE, TSQR_FLOAT
    from tests.gpu_util import from_device_colmajor, to_device_colmajor

    m, n = 27648, 32
    code = TSQR_FLOAT if dtype == np.float32 else TSQR_DOUBLE
    entry = "tsqr_s_stream" if dtype == np.float32 else "tsqr_d_stream"
    L = tsqr_amd.lib()
    need = L.tsqr_workspace_bytes(m, n, code)
    work = torch.empty(need, dtype=torch.uint8, device="cuda")
    conds = [None, 1e6, None, 1e6, None] if dtype == np.float32 else [None, 1e12, None]
    for k, cond in enumerate(conds):
        A = (O.gen_uniform(m, n, dtype, seed=50 + k) if cond is None
             else O.gen_conditioned(m, n, cond, dtype, seed=50 + k))
        buf, pA, lda = to_device_colmajor(A)
        R = torch.full((n * n,), float("nan"), dtype=buf.dtype, device="cuda")
        assert getattr(L, entry)(m, n, pA, lda, R.data_ptr(), n, work.data_ptr(), need, None) == 0
        torch.cuda.synchronize()
        Q = from_device_colmajor(buf, m, n, lda)
        floor = 32 if dtype == np.float64 else 1
        check(A, Q, R.view(n, n).cpu().numpy().T.copy(), f"call {k} cond={cond}", n_floor=floor)


@pytest.mark.parametrize("path", [pytest.param(AUTO, id="auto"), pytest.param(CHOLQR2, id="cholqr2")])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("m,n", SHAPES)
def test_flat_paths_forced_to_fall_back(m, n, dtype, path):
    """The smallest gate (1; ||R||_F ||R^-1||_F >= sqrt(n) > 1 for n >= 2)
    sends the flat paths to their Householder fallback, whose tree keeps each
    node's explicit Q and whose superleaves form their S from the path to
    the root (no down sweep): the same bounds as the Householder path, on
    every shape (short last tiles, up to five tree levels)."""
    A = O.gen_normal(m, n, dtype, seed=m * 5 + n)
    with LeafPath(dtype, path, gate=1.0):
        st, Q, R, _ = call("tsqr", A)
    assert st == 0
    assert n == 1 or not LAST_LEAF_STATS["flat"]
    check(A, Q, R, f"{m}x{n} {np.dtype(dtype).name} path={path} forced fallback leaves={LAST_LEAF_STATS}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_padded_and_unaligned_lda(dtype):
    A = O.gen_uniform(9000, 32, dtype, seed=9)
    st, Q, R, pad_ok = call("tsqr", A, lda=9000 + 37)
    assert st == 0 and pad_ok
    check(A, Q, R, "lda+37")
    st, Q, R, pad_ok = call("tsqr", A, lda=9001, offset_elems=1)  # 4/8-byte aligned only
    assert st == 0 and pad_ok
    check(A, Q, R, "misaligned")


def test_deterministic():
    A = O.gen_normal(40000, 32, np.float32, seed=1)
    _, Q1, R1, _ = call("tsqr", A)
    _, Q2, R2, _ = call("tsqr", A)
    assert np.array_equal(Q1, Q2) and np.array_equal(R1, R2)


def test_zero_column():
    A = O.gen_normal(3000, 8, np.float64, seed=4)
    A[:, 5] = 0.0
    st, Q, R, _ = call("tsqr", A)
    assert st == 0
    assert R[5, 5] == 0.0
    m = O.metrics(A, Q, R)
    assert m["backward"] <= 8 * 8 * O.unit_roundoff(np.float64)
    assert np.isfinite(Q).all()


def test_nan_propagates():
    A = O.gen_normal(2000, 8, np.float32, seed=4)
    A[17, 2] = np.nan
    st, Q, R, _ = call("tsqr", A)
    assert st == 0
    assert np.isnan(R).any()


def test_status_codes():
    from tsqr_amd._lib import TSQR_FLOAT
    import tsqr_amd

    A = O.gen_normal(100, 8, np.float32)
    assert call("tsqr", A[:5, :])[0] == -1                 # m < n
    assert call("tsqr", O.gen_normal(200, 40, np.float32))[0] == -2  # n > 32
    need = tsqr_amd.lib().tsqr_workspace_bytes(100000, 32, TSQR_FLOAT)
    assert need > 0
    assert call("tsqr", O.gen_normal(100000, 32, np.float32), work_bytes=need - 1)[0] == -3


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_reference_shaped_entry_small_ldwork(dtype):
    """tsqr_s / tsqr_d with the reference's caller-side sizing
    ldwork = 2*NUM_SM*BLOCK_SIZE (test_tsqr.cu:42): the library falls back to
    its internal workspace when that is too small."""
    A = O.gen_uniform(27648, 32, dtype)
    st, Q, R, _ = call("tsqr", A, ref_shaped=True, ldwork=2 * 108 * 128)
    assert st == 0
    check(A, Q, R, "ref-shaped")


def test_gram_path_is_taken_for_well_conditioned_fp32():
    A = O.gen_normal(200000, 32, np.float32, seed=21)
    with LeafPath(np.float32, GRAM_TREE):
        st, Q, R, _ = call("tsqr", A)
    assert st == 0
    assert LAST_LEAF_STATS["householder"] == 0 and LAST_LEAF_STATS["gram"] > 1 and not LAST_LEAF_STATS["flat"]
    check(A, Q, R, f"gram path leaves={LAST_LEAF_STATS}")


def test_flat_path_is_taken_for_well_conditioned_fp32():
    A = O.gen_uniform(200000, 32, np.float32, seed=22)
    with LeafPath(np.float32, AUTO):
        st, Q, R, _ = call("tsqr", A)
    assert st == 0
    assert LAST_LEAF_STATS["flat"] == 1
    check(A, Q, R, f"flat path leaves={LAST_LEAF_STATS}")


@pytest.mark.parametrize("cond", [1e2, 5e3])
def test_flat_gate(cond):
    """cond 1e2 passes the gate (flat); cond 5e3 gives ||R||_F||R^-1||_F above
    1024 and must take the tree, with the same error bounds either way."""
    A = O.gen_conditioned(100000, 32, cond, np.float32, seed=6)
    with LeafPath(np.float32, AUTO):
        st, Q, R, _ = call("tsqr", A)
    assert st == 0
    assert LAST_LEAF_STATS["flat"] == (1 if cond < 1e3 else 0)
    check(A, Q, R, f"cond={cond:g} leaves={LAST_LEAF_STATS}")


@pytest.mark.parametrize("cond", [1e4, 1e6])
def test_ill_conditioned_falls_back_to_householder(cond):
    """cond(A) far above the gate: every superleaf is rejected by the Gram
    path, and the Householder result meets the kappa-independent bounds."""
    A = O.gen_conditioned(60000, 32, cond, np.float32, seed=5)
    with LeafPath(np.float32, AUTO):
        st, Q, R, _ = call("tsqr", A)
    assert st == 0
    assert LAST_LEAF_STATS["gram"] == 0
    check(A, Q, R, f"cond={cond:g} leaves={LAST_LEAF_STATS}")


@pytest.mark.parametrize("m", [40000, 4_300_000])
def test_mixed_leaf_paths_in_one_call(m):
    """One aligned block of 2048 rows made nearly rank deficient (two columns
    almost equal on its rows only) sends the superleaves covering it
    Householder, the others stay on the Gram path.  Superleaves are 1 to 8
    whole tiles (FP32, n = 32: 256 rows), chosen from m, so the block covers
    2048 / leaf_rows superleaves exactly: 8 at m = 40000, 1 at m = 4.3e6."""
    n, tile, block = 32, 256, 2048
    A = O.gen_normal(m, n, np.float32, seed=8)
    rows = slice(3 * block, 4 * block)
    A[rows, 7] = A[rows, 6] * (1 + 1e-6)
    with LeafPath(np.float32, GRAM_TREE):
        st, Q, R, _ = call("tsqr", A)
    assert st == 0
    leaves = LAST_LEAF_STATS["gram"] + LAST_LEAF_STATS["householder"]
    heights = [t * tile for t in range(1, 9) if -(-m // (t * tile)) == leaves]
    assert len(heights) == 1, (leaves, heights)
    assert LAST_LEAF_STATS["householder"] == block // heights[0], (LAST_LEAF_STATS, heights)
    check(A, Q, R, f"mixed leaves={LAST_LEAF_STATS} leaf_rows={heights[0]}")
