"""The autotuner on the GPU: it produces a table, and every candidate
superleaf height it may choose meets the same error bounds (the choice never
changes results beyond tolerance)."""
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
        autotune.reset()
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
        autotune.reset()
    assert st == 0
    u = O.unit_roundoff(dtype)
    mt = O.metrics(A, Q, R)
    bound = 8 * 32 * u
    print(f"tiles={tiles} {np.dtype(dtype).name} path={path}: backward={mt['backward']:.3e} "
          f"orth={mt['orth']:.3e}", flush=True)
    assert mt["all_close"] and mt["r_lower_zero"]
    assert mt["backward"] <= bound and mt["orth"] <= bound
