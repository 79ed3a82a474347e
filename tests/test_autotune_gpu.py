"""The autotuner on the GPU: it produces a table, and every candidate it may
choose (superleaf height, tree arity, streaming grid shape) meets the same
error bounds (the choice never changes results beyond tolerance)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import LeafPath, call

pytestmark = pytest.mark.gpu


def test_tune_small_table():
    from tsqr_amd import autotune

    try:
        table = autotune.tune(dtypes=("float32",), widths=(32,), log2m=(15, 18), reps=2)
    finally:
        autotune.restore_default()
    tiles = table["leaf_tiles"]["float32"]["32"]
    assert set(tiles) == {"15", "18"} and all(t in autotune.CANDIDATES for t in tiles.values())


@pytest.mark.parametrize("dtype,path", [(np.float32, 0), (np.float32, 1), (np.float32, 3), (np.float64, 3)])
@pytest.mark.parametrize("tiles", [1, 2, 4, 8])
def test_every_candidate_meets_the_bounds(tiles, dtype, path):
    import tsqr_amd
    from tsqr_amd import autotune
    from tsqr_amd._lib import TSQR_DOUBLE, TSQR_FLOAT

    m, n = 50000, 32  # floor(log2 m) = 15
    code = TSQR_FLOAT if dtype == np.float32 else TSQR_DOUBLE
    L = tsqr_amd.lib()
    A = O.gen_normal(m, n, dtype, seed=tiles)
    try:
        assert L.tsqr_set_leaf_tiles(code, n, 15, tiles) == 0
        with LeafPath(dtype, path):
            st, Q, R, _ = call("tsqr", A)
    finally:
        autotune.restore_default()
    assert st == 0
    u = O.unit_roundoff(dtype)
    mt = O.metrics(A, Q, R)
    bound = 8 * 32 * u
    print(f"tiles={tiles} {np.dtype(dtype).name} path={path}: backward={mt['backward']:.3e} "
          f"orth={mt['orth']:.3e}", flush=True)
    assert mt["all_close"] and mt["r_lower_zero"]
    assert mt["backward"] <= bound and mt["orth"] <= bound


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [8, 32])
@pytest.mark.parametrize("arity", [2, 3, 4, 0])
def test_every_tree_arity_meets_the_bounds(arity, n, dtype):
    """Householder path with one tile per superleaf, so that the tree over
    the ~200 superleaves is deep (arity 2: eight levels) and ragged."""
    import tsqr_amd
    from tsqr_amd import autotune
    from tsqr_amd._lib import TSQR_DOUBLE, TSQR_FLOAT

    m = 50000
    code = TSQR_FLOAT if dtype == np.float32 else TSQR_DOUBLE
    L = tsqr_amd.lib()
    A = O.gen_normal(m, n, dtype, seed=10 + arity)
    try:
        assert L.tsqr_set_leaf_tiles(code, n, 15, 1) == 0
        assert L.tsqr_set_tree_arity(code, n, arity) == 0
        used = L.tsqr_tree_arity(code, m, n)
        assert arity == 0 or used == arity
        with LeafPath(dtype, 1):
            st, Q, R, _ = call("tsqr", A)
    finally:
        autotune.restore_default()
    assert st == 0
    u = O.unit_roundoff(dtype)
    mt = O.metrics(A, Q, R)
    bound = 8 * 32 * u
    print(f"arity={used} n={n} {np.dtype(dtype).name}: backward={mt['backward']:.3e} orth={mt['orth']:.3e}",
          flush=True)
    assert mt["all_close"] and mt["r_lower_zero"]
    assert mt["backward"] <= bound and mt["orth"] <= bound


@pytest.mark.parametrize("groups,q_waves", [(256, 1024), (1000, 4096), (512, 2048)])
@pytest.mark.parametrize("fn", ["tsqr", "tsqr_mixed"])
def test_every_stream_shape_meets_the_bounds(groups, q_waves, fn):
    """The flat Gram path (FP32, the library default) with the streaming
    passes' grid capped and the Q pass split differently."""
    import tsqr_amd
    from tsqr_amd import autotune

    m, n = 1 << 20, 32
    L = tsqr_amd.lib()
    A = O.gen_normal(m, n, np.float32, seed=groups)
    try:
        assert L.tsqr_set_stream_shape(groups, q_waves) == 0
        st, Q, R, _ = call(fn, A)
    finally:
        autotune.restore_default()
    assert st == 0
    mt = O.metrics(A, Q, R)
    u = O.unit_roundoff(np.float32)
    print(f"groups={groups} q_waves={q_waves} {fn}: backward={mt['backward']:.3e} orth={mt['orth']:.3e}",
          flush=True)
    assert mt["all_close"] and mt["r_lower_zero"]
    assert mt["backward"] <= 8 * 32 * u and mt["orth"] <= 8 * 32 * u


def test_tune_tree_and_stream_small():
    from tsqr_amd import autotune

    try:
        t = autotune.tune_tree(dtypes=("float32",), widths=(32,), ms=(1 << 16,), reps=1, verbose=False)
        s = autotune.tune_stream(ms=(1 << 18,), reps=1, verbose=False)
    finally:
        autotune.restore_default()
    a = t["tree_arity"]["float32"]["32"]
    assert a == 0 or a in autotune.ARITY_CANDIDATES
    assert s["stream_max_groups"] in autotune.STREAM_GROUP_CANDIDATES
    assert s["gramq_waves"] in autotune.GRAMQ_WAVE_CANDIDATES
