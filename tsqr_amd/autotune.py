"""Autotuner for the launch shapes the reference fixes at compile time.

The reference chooses its tile height and grid with NUM_SM / BLOCK_SIZE
macros per precision (kernelQR.h:3-11).  Here the tile shape follows from
the register budget (512 bytes of tile per lane), and what remains free is
how many tiles form a level-0 superleaf (1, 2, 4 or 8): taller superleaves
mean fewer, longer streams and a shallower tree, shorter ones more
parallelism for small m.  The library picks by a rule (about m / 2048 rows
per superleaf); this module measures instead.

tune() times each candidate on the current GPU for every (dtype, padded
width, floor(log2 m)) bucket asked for, with the library's own default leaf
path for the dtype, and returns the fastest; apply() writes a table into the
library (tsqr_set_leaf_tiles), save()/load() keep it as JSON.  The choice
only moves the superleaf boundaries: every candidate meets the same error
bounds (tests/test_autotune_gpu.py checks that).

    python -m tsqr_amd.autotune [--quick] [--out FILE]
"""
from __future__ import annotations

import argparse
import json
import os

DTYPES = {"float32": 1, "float64": 2}  # TSQR_FLOAT, TSQR_DOUBLE
CANDIDATES = (1, 2, 4, 8)
DEFAULT_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_mi355x.json")


def _lib():
    from ._lib import lib

    return lib()


def apply(table: dict) -> None:
    """table: {"leaf_tiles": {dtype: {width: {log2m: tiles}}}, "gemm_waves": W}."""
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


def reset() -> None:
    """Back to the built-in rule everywhere."""
    L = _lib()
    for code in DTYPES.values():
        for width in (8, 16, 32):
            for k in range(32):
                L.tsqr_set_leaf_tiles(code, width, k, 0)
    L.tsqr_set_gemm_waves(512)


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
    else:
        table = tune()
    save(table, a.out)
    print(f"wrote {a.out}")


if __name__ == "__main__":
    main()
