"""The tuning table API and the autotuner's table format (no GPU needed:
the table lives on the host and shapes the plan only)."""
import json

import pytest

import tsqr_amd
from tsqr_amd import autotune
from tsqr_amd._lib import TSQR_DOUBLE, TSQR_FLOAT


@pytest.fixture
def lib():
    try:
        L = tsqr_amd.lib()
    except ImportError:
        pytest.skip("libtsqr_amd.so not built")
    yield L
    autotune.restore_default()


def test_leaf_tiles_table(lib):
    m, n = 3 << 19, 32  # floor(log2 m) = 20
    base = lib.tsqr_leaf_tiles(TSQR_FLOAT, m, n)
    assert 1 <= base <= 8
    for t in (1, 2, 4, 8):
        assert lib.tsqr_set_leaf_tiles(TSQR_FLOAT, n, 20, t) == 0
        assert lib.tsqr_leaf_tiles(TSQR_FLOAT, m, n) == t
        # other buckets, widths and dtypes are untouched
        assert lib.tsqr_leaf_tiles(TSQR_FLOAT, m * 2, n) == lib.tsqr_leaf_tiles(TSQR_FLOAT, m * 2, n)
    one = lib.tsqr_workspace_bytes(m, n, TSQR_FLOAT)
    lib.tsqr_set_leaf_tiles(TSQR_FLOAT, n, 20, 1)
    assert lib.tsqr_workspace_bytes(m, n, TSQR_FLOAT) >= one  # more superleaves, more workspace
    # FP64 with n = 32 holds at most 4 tiles per superleaf: clipped
    assert lib.tsqr_set_leaf_tiles(TSQR_DOUBLE, 32, 20, 8) == 0
    assert lib.tsqr_leaf_tiles(TSQR_DOUBLE, m, 32) == 4
    assert lib.tsqr_set_leaf_tiles(TSQR_FLOAT, n, 20, 0) == 0
    assert lib.tsqr_leaf_tiles(TSQR_FLOAT, m, n) == base


def test_table_argument_checks(lib):
    assert lib.tsqr_set_leaf_tiles(3, 32, 20, 1) == -1
    assert lib.tsqr_set_leaf_tiles(TSQR_FLOAT, 33, 20, 1) == -1
    assert lib.tsqr_set_leaf_tiles(TSQR_FLOAT, 32, 32, 1) == -1
    assert lib.tsqr_set_leaf_tiles(TSQR_FLOAT, 32, 20, 9) == -1
    assert lib.tsqr_set_gemm_waves(0) == -1
    assert lib.tsqr_set_gemm_waves(1024) == 0


def test_tree_arity_table(lib):
    m = 3 << 20
    full32 = lib.tsqr_tree_arity(TSQR_FLOAT, m, 32)
    full16 = lib.tsqr_tree_arity(TSQR_FLOAT, m, 16)
    full64 = lib.tsqr_tree_arity(TSQR_DOUBLE, m, 32)
    assert full32 >= 4 and full32 & (full32 - 1) == 0  # H / n with H a power of two
    assert full16 >= 2 * full32  # a narrower width fits more child Rs into a tile
    base_ws = lib.tsqr_workspace_bytes(m, 32, TSQR_FLOAT)
    for a in (2, 4):
        assert lib.tsqr_set_tree_arity(TSQR_FLOAT, 32, a) == 0
        assert lib.tsqr_tree_arity(TSQR_FLOAT, m, 32) == a
        # a narrower tree has more nodes: more blocks in the workspace
        assert lib.tsqr_workspace_bytes(m, 32, TSQR_FLOAT) > base_ws
        # other widths and the other dtype keep full blocks
        assert lib.tsqr_tree_arity(TSQR_FLOAT, m, 16) == full16
        assert lib.tsqr_tree_arity(TSQR_DOUBLE, m, 32) == full64
    # above H / n it is clipped to the full block
    assert lib.tsqr_set_tree_arity(TSQR_FLOAT, 32, 64) == 0
    assert lib.tsqr_tree_arity(TSQR_FLOAT, m, 32) == full32
    assert lib.tsqr_set_tree_arity(TSQR_FLOAT, 32, 0) == 0
    assert lib.tsqr_workspace_bytes(m, 32, TSQR_FLOAT) == base_ws


def test_stream_shape_checks(lib):
    m = 23887872
    base = lib.tsqr_workspace_bytes(m, 32, TSQR_FLOAT)
    assert lib.tsqr_set_stream_shape(256, 1024) == 0
    # fewer streaming workgroups: fewer partial sums to hold
    assert lib.tsqr_workspace_bytes(m, 32, TSQR_FLOAT) <= base
    assert lib.tsqr_set_stream_shape(0, 0) == 0
    assert lib.tsqr_workspace_bytes(m, 32, TSQR_FLOAT) == base
    assert lib.tsqr_set_stream_shape(1025, 0) == -1
    assert lib.tsqr_set_stream_shape(-1, 0) == -1
    assert lib.tsqr_set_stream_shape(0, -5) == -1
    assert lib.tsqr_set_tree_arity(TSQR_FLOAT, 32, 1) == -1
    assert lib.tsqr_set_tree_arity(3, 32, 2) == -1
    assert lib.tsqr_tree_arity(TSQR_FLOAT, 10, 32) == -1


def test_save_load_apply_roundtrip(lib, tmp_path):
    table = {"leaf_tiles": {"float32": {"32": {"20": 2, "22": 8}}, "float64": {"16": {"18": 4}}},
             "gemm_waves": 768, "tree_arity": {"float32": {"32": 4}}, "stream_max_groups": 512,
             "gramq_waves": 4096}
    path = tmp_path / "t.json"
    autotune.save(table, str(path))
    got = autotune.load(str(path))
    assert got == json.loads(path.read_text())
    autotune.apply(got)
    assert lib.tsqr_leaf_tiles(TSQR_FLOAT, 3 << 19, 32) == 2
    assert lib.tsqr_leaf_tiles(TSQR_FLOAT, 3 << 21, 32) == 8
    assert lib.tsqr_leaf_tiles(TSQR_DOUBLE, 3 << 17, 16) == 4
    assert lib.tsqr_tree_arity(TSQR_FLOAT, 1 << 20, 32) == 4
    autotune.reset()
    assert lib.tsqr_leaf_tiles(TSQR_FLOAT, 3 << 21, 32) != 0
    assert lib.tsqr_tree_arity(TSQR_FLOAT, 1 << 20, 32) > 4


def test_committed_table_is_valid():
    import os

    if not os.path.exists(autotune.DEFAULT_FILE):
        pytest.skip("no tuned table committed")
    t = autotune.load()
    for dtype, widths in t["leaf_tiles"].items():
        assert dtype in autotune.DTYPES
        for w, buckets in widths.items():
            assert int(w) in (8, 16, 32)
            for k, tiles in buckets.items():
                assert 0 <= int(k) <= 31 and tiles in autotune.CANDIDATES
    for dtype, widths in t.get("tree_arity", {}).items():
        assert dtype in autotune.DTYPES
        for w, a in widths.items():
            assert int(w) in (8, 16, 32) and (a == 0 or a in autotune.ARITY_CANDIDATES)
    if "stream_max_groups" in t:
        assert t["stream_max_groups"] in autotune.STREAM_GROUP_CANDIDATES
        assert t["gramq_waves"] in autotune.GRAMQ_WAVE_CANDIDATES
