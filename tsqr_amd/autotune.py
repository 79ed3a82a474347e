"""Autotuner for the launch shapes the reference fixes at compile time.

The reference chooses its tile height and grid with NUM_SM / BLOCK_SIZE
macros per precision (kernelQR.h:3-11).  Here the tile shape follows from
the register budget (512 bytes of tile per lane), and what remains free is
how many tiles form a level-0 superleaf (1, 2, 4 or 8): taller superleaves
mean fewer, longer streams and a shallower tree, shorter ones more
parallelism for small m.  The library picks by a rule (about m / 2048 rows
per superleaf); this module measures instead.  Three more knobs:

  tree_arity         children per node of the Householder tree above the
                     superleaves (the reference's fan-in BLOCK_SIZE / n):
                     2 .. H/n per (dtype, padded width), 0 = H/n;
  stream_max_groups  the largest grid of the streaming Gram passes of the
                     flat path (the R pass and the checker), <= 1024;
  gramq_waves        waves the flat path's Q pass aims for when superleaves
                     are few.

tune() times each leaf-tile candidate on the current GPU for every (dtype,
padded width, floor(log2 m)) bucket asked for, with the library's own
default leaf path for the dtype; tune_tree() times the arities on the
Householder path; tune_stream() the grid shapes on the flat path.  apply()
writes a table into the library, save()/load() keep it as JSON.  The knobs
move work between workgroups and tree levels only: every candidate meets
the same error bounds (tests/test_autotune_gpu.py checks that).

    python -m tsqr_amd.autotune [--quick] [--out FILE]
"""
from __future__ import annotations

import argparse
import json
import os

DTYPES = {"float32": 1, "float64": 2}  # TSQR_FLOAT, TSQR_DOUBLE
CANDIDATES = (1, 2, 4, 8)
ARITY_CANDIDATES = (2, 4, 8, 16, 32, 64)  # clipped to H / n per (dtype, width)
STREAM_GROUP_CANDIDATES = (128, 256, 512, 1024)
GRAMQ_WAVE_CANDIDATES = (1024, 2048, 4096)
LEAF_HOUSEHOLDER = 1  # TSQR_LEAF_HOUSEHOLDER
DEFAULT_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_mi355x.json")


def _lib():
    from ._lib import lib

    return lib()


def apply(table: dict) -> None:
    """table: {"leaf_tiles": {dtype: {width: {log2m: tiles}}}, "gemm_waves": W,
    "tree_arity": {dtype: {width: arity}}, "stream_max_groups": G,
    "gramq_waves": Q}; every key but leaf_tiles is optional."""
    L = _lib()
    for dtype, widths in table.get("leaf_tiles", {}).items():
        for width, buckets in widths.items():
            for k, tiles in buckets.items():
                st = L.tsqr_set_leaf_tiles(DTYPES[dtype], int(width), int(k), int(tiles))
                if st != 0:
                    raise ValueError(f"bad tuning entry {dtype} n={width} log2m={k} tiles={tiles}")
    if "gemm_waves" in table:
        if L.tsqr_set_gemm_waves(int(table["gemm_waves"])) != 0:
            raise ValueError(f"bad gemm_waves {table['gemm_waves']}")
    for dtype, widths in table.get("tree_arity", {}).items():
        for width, arity in widths.items():
            if L.tsqr_set_tree_arity(DTYPES[dtype], int(width), int(arity)) != 0:
                raise ValueError(f"bad tree_arity {dtype} n={width} arity={arity}")
    if "stream_max_groups" in table or "gramq_waves" in table:
        g, q = int(table.get("stream_max_groups", 0)), int(table.get("gramq_waves", 0))
        if L.tsqr_set_stream_shape(g, q) != 0:
            raise ValueError(f"bad stream shape {g}, {q}")


def reset() -> None:
    """Back to the built-in rule everywhere."""
    L = _lib()
    for code in DTYPES.values():
        for width in (8, 16, 32):
            for k in range(32):
                L.tsqr_set_leaf_tiles(code, width, k, 0)
            L.tsqr_set_tree_arity(code, width, 0)
    L.tsqr_set_gemm_waves(512)
    L.tsqr_set_stream_shape(0, 0)


def restore_default() -> None:
    """Back to what the library starts with: the built-in rule, then the
    committed table for this GPU unless TSQR_AMD_NO_TUNING is set."""
    reset()
    if not os.environ.get("TSQR_AMD_NO_TUNING") and os.path.exists(DEFAULT_FILE):
        apply(load(DEFAULT_FILE))


def save(table: dict, path: str = DEFAULT_FILE) -> None:
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def load(path: str = DEFAULT_FILE) -> dict:
    with open(path) as f:
        table = json.load(f)
    if not isinstance(table, dict) or "leaf_tiles" not in table:
        raise ValueError(f"{path}: not a tuning table")
    return table


def _time_call(m: int, n: int, dtype: str, reps: int) -> float:
    """Mean milliseconds of one tsqr_*_stream call (HIP events on the launch
    stream) on a seeded uniform matrix; the calls refactor the previous Q,
    which is the same work."""
    import torch

    from . import api

    tdt = torch.float32 if dtype == "float32" else torch.float64
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    A = api.empty_colmajor(m, n, tdt, dev)
    A.uniform_(-1.0, 1.0, generator=g)
    R = api.empty_colmajor(n, n, tdt, dev)
    work = torch.empty(max(_lib().tsqr_workspace_bytes(m, n, DTYPES[dtype]), 1), dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream(dev)
    api.tsqr(A, R, work, stream)  # warm-up
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        api.tsqr(A, R, work, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


class _LeafPath:
    """Forces a leaf path for one dtype while timing, then restores the
    library default (TSQR_LEAF_AUTO for FP32, CholeskyQR2 for FP64)."""

    def __init__(self, dtype: str, path: int):
        self.code, self.path = DTYPES[dtype], path

    def __enter__(self):
        from ._lib import set_leaf_path

        set_leaf_path(self.code, self.path, 1024.0)

    def __exit__(self, *exc):
        from ._lib import set_leaf_path

        if self.code == 1:
            set_leaf_path(1, 0, 1024.0)  # TSQR_LEAF_AUTO, gate 1024
        else:
            set_leaf_path(2, 3, 1e5)  # TSQR_LEAF_CHOLQR2, gate 1e5


def tune_tree(dtypes=("float32", "float64"), widths=(8, 16, 32), ms=(3 << 17, 3 << 20), reps: int = 5,
              verbose: bool = True) -> dict:
    """Times every arity the width allows on the Householder path (the path
    the tree serves; the flat Gram path has no tree) at each m, summed over
    ms, and returns {dtype: {width: best}} (0 where the full block wins)."""
    L = _lib()
    best_table, timings = {}, {}
    for dtype in dtypes:
        code = DTYPES[dtype]
        for width in widths:
            L.tsqr_set_tree_arity(code, width, 0)
            full = L.tsqr_tree_arity(code, ms[0], width)
            times = {}
            with _LeafPath(dtype, LEAF_HOUSEHOLDER):
                for a in [c for c in ARITY_CANDIDATES if c < full] + [full]:
                    L.tsqr_set_tree_arity(code, width, a)
                    assert L.tsqr_tree_arity(code, ms[0], width) == a
                    times[a] = sum(_time_call(m, width, dtype, reps) for m in ms)
            best = min(times, key=times.get)
            L.tsqr_set_tree_arity(code, width, 0 if best == full else best)
            best_table.setdefault(dtype, {})[str(width)] = 0 if best == full else best
            timings.setdefault(dtype, {})[str(width)] = {str(a): round(v, 4) for a, v in times.items()}
            if verbose:
                print(f"tree {dtype} n={width} m={list(ms)}: " + " ".join(f"{a}:{v:.4f}" for a, v in times.items())
                      + f" -> {best}", flush=True)
    return {"tree_arity": best_table, "tree_timings_ms": timings}


def tune_stream(ms=(3 << 19, 3 << 21, 23887872), reps: int = 5, verbose: bool = True) -> dict:
    """Times the grid shapes of the flat path's streaming passes (FP32,
    n = 32, the library's default leaf path) summed over ms and returns
    {"stream_max_groups": G, "gramq_waves": Q} of the fastest."""
    L = _lib()
    times = {}
    try:
        for g in STREAM_GROUP_CANDIDATES:
            for q in GRAMQ_WAVE_CANDIDATES:
                assert L.tsqr_set_stream_shape(g, q) == 0
                times[(g, q)] = sum(_time_call(m, 32, "float32", reps) for m in ms)
                if verbose:
                    print(f"stream groups={g} q_waves={q} m={list(ms)}: {times[(g, q)]:.4f}", flush=True)
    finally:
        L.tsqr_set_stream_shape(0, 0)
    g, q = min(times, key=times.get)
    L.tsqr_set_stream_shape(g, q)
    return {"stream_max_groups": g, "gramq_waves": q,
            "stream_timings_ms": {f"{a},{b}": round(v, 4) for (a, b), v in times.items()}}


def tune(dtypes=("float32", "float64"), widths=(8, 16, 32), log2m=range(12, 25), reps: int = 5,
         verbose: bool = True) -> dict:
    """Measures every candidate superleaf height per bucket (m = 1.5 * 2^k,
    the middle of the bucket) and returns the table of the fastest; the
    library's table is left as returned."""
    L = _lib()
    table = {"leaf_tiles": {}, "gemm_waves": 512, "timings_ms": {}}
    for dtype in dtypes:
        code = DTYPES[dtype]
        for width in widths:
            for k in log2m:
                m = 3 << (k - 1)
                times = {}
                for t in CANDIDATES:
                    L.tsqr_set_leaf_tiles(code, width, k, t)
                    used = L.tsqr_leaf_tiles(code, m, width)
                    if used != t:  # clipped to the tile shape's maximum
                        continue
                    times[t] = _time_call(m, width, dtype, reps)
                best = min(times, key=times.get)
                L.tsqr_set_leaf_tiles(code, width, k, best)
                table["leaf_tiles"].setdefault(dtype, {}).setdefault(str(width), {})[str(k)] = best
                table["timings_ms"].setdefault(dtype, {}).setdefault(str(width), {})[str(k)] = {
                    str(t): round(v, 4) for t, v in times.items()}
                if verbose:
                    print(f"{dtype} n={width} m={m}: " + " ".join(f"{t}:{v:.4f}" for t, v in times.items())
                          + f" -> {best}", flush=True)
    return table


def main():
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--quick", action="store_true", help="n = 32 only, every other bucket")
    p.add_argument("--out", default=DEFAULT_FILE)
    a = p.parse_args()
    if a.quick:
        table = tune(widths=(32,), log2m=range(12, 25, 2))
        table.update(tune_tree(widths=(32,)))
    else:
        table = tune()
        table.update(tune_tree())
    table.update(tune_stream())
    save(table, a.out)
    print(f"wrote {a.out}")


if __name__ == "__main__":
    main()
