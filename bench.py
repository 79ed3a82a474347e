"""Benchmark of the TSQR hot path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

Workload (default "tsqr_max_fp32"): FP32 TSQR at the maximum height the
reference supports, m = 23,887,872, n = 32 (A is 3.06 GB), iid uniform[-1,1)
entries from a seeded generator, resident in HBM before timing starts.  One
step is one full tsqr_s_stream call (A overwritten by the explicit Q, R
written).  The kernels have no data-dependent control flow, so step k > 1
factors the Q left by step k-1 with exactly the same work; A is not
re-uploaded between steps.

metric  TSQR HBM GB/s = B_alg / (seconds per step), with
        B_alg = (3mn + n^2) * s bytes when A does not fit in the aggregate
        LDS of the part (256 CUs x 160 KiB = 40 MiB): A read twice and
        written once are then the lower bound for an explicit Q; and
        B_alg = (2mn + n^2) * s (read A once, write Q and R once) when it
        does, as for the 27648 x 32 and 27648 x 128 configurations.  The
        compulsory (2mn + n^2) * s is reported beside it.
roofline  for the dominant phase of the two streaming passes (leaf R pass:
        reads A, m*n*4 algorithmic bytes; leaf Q pass: reads A and writes Q,
        2*m*n*4), from HIP events on the launch stream over a second pass of
        the same K steps (the marks are kept out of the timed pass: each costs
        a few microseconds of GPU time), against the 8 TB/s HBM peak.
cpu_baseline  the oracle's C restatement of the reference algorithm
        (OpenMP over tiles, all host cores) on a bounded sample of the same
        workload: the first m/8 rows of the very array the GPU factors,
        copied to the host before timing; plus OpenBLAS sgeqrf+sorgqr on the
        same sample; with the host CPU model and thread count.

batched  (same line, key "batched") the headline's second half: QRs/s of the
        512-matrix uneven batch, mixed method vs plain FP32 Householder, each
        as one hipGraph replay per batch, with the max backward and
        orthogonality errors of each method over the batch.

Other configurations (27648x32, 27648x128, FP64, mixed single-matrix) are
exercised by the parity tests; --workload selects them for ad-hoc runs.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Aggregate LDS of the part (256 CUs x 160 KiB): below it A could stay on chip
# between the R and Q passes, so only (2mn + n^2) * s bytes are compulsory.
AGGREGATE_LDS_BYTES = 256 * 160 * 1024
FP64_MFMA_PEAK_TFLOPS = 78.6
BF16_MFMA_PEAK_TFLOPS = 2500.0
# Measured HBM bytes per dispatch (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by
# a 1 GiB copy); the file names the command that produced it.
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "tsqr_fp32_23887872x32_pmc_traffic.md")


def measured_traffic(workload: str):
    """traffic-json line of the counter summary, if it covers this workload."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic-json:"):
                    d = json.loads(line.split(":", 1)[1])
                    return d if d.get("workload") == workload else None
    except OSError:
        return None
    return None

WORKLOADS = {
    # name: (kind, m, n, dtype)
    "tsqr_max_fp32": ("tsqr", 23887872, 32, "float32"),
    "tsqr_max_fp64": ("tsqr", 5971968, 32, "float64"),
    "tsqr_27648x32_fp32": ("tsqr", 27648, 32, "float32"),
    "qr_27648x128_fp32": ("qr", 27648, 128, "float32"),
    "qr_27648x128_fp64": ("qr", 27648, 128, "float64"),
    "tsqr_mixed_max_fp32": ("tsqr_mixed", 23887872, 32, "float32"),
    "tsqr_mixed_27648x32_fp32": ("tsqr_mixed", 27648, 32, "float32"),
}


def algorithmic_bytes(m: int, n: int, esz: int) -> int:
    """B_alg of one explicit-Q factorisation (see the module docstring)."""
    if m * n * esz <= AGGREGATE_LDS_BYTES:
        return (2 * m * n + n * n) * esz
    return (3 * m * n + n * n) * esz


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="tsqr_max_fp32",
                   choices=sorted(WORKLOADS) + ["batch_512_fp32", "qr_mixed_1m_sweep"])
    p.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    p.add_argument("--cpu-fraction", type=int, default=8, help="cpu sample = m / this")
    p.add_argument("--copy-bytes", type=int, default=0,
                   help="before timing, copy a buffer of this many bytes device-to-device (a plain copy of "
                        "known size for calibrating memory counters) and report its bandwidth")
    p.add_argument("--leaf-path", choices=("default", "auto", "householder", "gram_tree", "cholqr2"), default="default",
                   help="level-0 superleaf path (default: the library's default for the dtype)")
    p.add_argument("--no-batch", action="store_true", help="skip the batched (mixed vs plain FP32) leg")
    p.add_argument("--batch-steps", type=int, default=5)
    p.add_argument("--batch-count", type=int, default=512)
    return p.parse_args()


def gram_error(q64):
    """||I - Q^T Q||_F in FP64 for a tall m x n GPU tensor, as a batched
    product over 8192-row blocks (one rocBLAS call with a short inner
    dimension per block instead of one m-long reduction)."""
    import torch

    m, n = q64.shape
    blk = 8192
    mp = -(-m // blk) * blk
    qp = torch.zeros((mp, n), dtype=torch.float64, device=q64.device)
    qp[:m] = q64
    qb = qp.view(-1, blk, n)
    g = torch.bmm(qb.transpose(1, 2).contiguous(), qb).sum(0)
    return float(torch.linalg.norm(torch.eye(n, dtype=torch.float64, device=q64.device) - g))


def cond_sweep(steps: int, m: int = 1048576) -> dict:
    """Mixed-precision recursive QR vs the plain qr_s at m = 1,048,576,
    n in {64, 128}, A = U diag(s) V^T (U, V from FP64 QR of seeded Gaussians
    on the GPU, s geometric from 1 to 1/cond), cond in {1e2, 1e4, 1e6}.
    Each step restores A (untimed) and times one call with HIP events on the
    launch stream; errors of the last call in FP64 on the GPU against the
    FP64 A (before its rounding to FP32) and the FP32 A."""
    import torch

    import tsqr_amd

    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    g = torch.Generator(device=dev)
    out = {"m": m, "cases": []}
    for n in (64, 128):
        for cond in (1e2, 1e4, 1e6):
            g.manual_seed(n * 7 + int(np.log10(cond)))
            U, _ = torch.linalg.qr(torch.randn(m, n, dtype=torch.float64, device=dev, generator=g))
            V, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g))
            sv = cond ** (-torch.arange(n, dtype=torch.float64, device=dev) / (n - 1))
            A64 = (U * sv) @ V.T
            del U
            src = tsqr_amd.colmajor(A64.float())
            A32 = src.double()
            case = {"n": n, "cond": cond}
            for name, fn in (("mixed", tsqr_amd.qr_mixed), ("plain", tsqr_amd.qr)):
                A = src.clone(memory_format=torch.preserve_format)
                R = tsqr_amd.empty_colmajor(n, n, torch.float32, dev)
                work = torch.empty(tsqr_amd.qr_workspace_bytes(m, n, 1), dtype=torch.uint8, device=dev)
                times = []
                for k in range(steps + 1):
                    A.copy_(src)
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    fn(A, R, work, stream)
                    e1.record(stream)
                    torch.cuda.synchronize()
                    if k:
                        times.append(e0.elapsed_time(e1))
                Q64, R64 = A.double(), torch.triu(R.double())
                back = float(torch.linalg.norm(A32 - Q64 @ R64) / torch.linalg.norm(A32))
                orth = gram_error(Q64)
                case[name] = {"ms": round(sum(times) / len(times), 4), "backward": float(f"{back:.3e}"),
                              "orth": float(f"{orth:.3e}")}
                del A, work, Q64
            print(f"# sweep {case}", file=sys.stderr, flush=True)
            out["cases"].append(case)
            del A64, src, A32
            torch.cuda.empty_cache()
    return out


def batched_leg(steps: int, count: int) -> dict:
    """512 uneven FP32 matrices (tsqr_amd.workloads.batch_shapes: m
    log-uniform in [4096, 262144] rounded down to a multiple of n, n in
    {8, 16, 32, 64}; iid N(0,1) from a seeded torch generator),
    factored by one replay of the captured hipGraph per step, mixed method
    and plain (Householder) method.  Each step starts from the same input
    (an untimed device copy restores it); the step is timed with HIP events
    on the launch stream.  Errors of the last step of each method are measured
    in FP64 on the GPU against the input: max over the batch of
    ||A - QR||_F / ||A||_F and ||I - Q^T Q||_F."""
    import torch

    import tsqr_amd
    from tsqr_amd.workloads import batch_shapes

    dev = torch.device("cuda", 0)
    shapes = batch_shapes(count=count)
    g = torch.Generator(device=dev)
    g.manual_seed(1)
    src = [tsqr_amd.empty_colmajor(m, n, torch.float32, dev) for m, n in shapes]
    for a in src:
        a.normal_(generator=g)
    mats = [a.clone(memory_format=torch.preserve_format) for a in src]
    mats = [tsqr_amd.colmajor(a) if a.stride(0) != 1 else a for a in mats]
    # compulsory bytes per matrix, (2mn + n^2) * 4: read A, write Q and R
    balg = sum((2 * m * n + n * n) * 4 for m, n in shapes)
    stream = torch.cuda.current_stream(dev)
    out = {"workload": f"{count} matrices, m log-uniform in [4096, 262144] rounded down to a multiple of n, "
                       f"n in {{8,16,32,64}}, shapes from std::mt19937_64(0) (tsqr_amd.workloads.batch_shapes), "
                       f"float32 iid N(0,1); one hipGraph replay per batch",
           "matrices": count, "algorithmic_bytes_per_batch": balg}
    for method in ("mixed", "plain"):
        b = tsqr_amd.Batch(mats, method=method)
        times = []
        for k in range(steps + 1):  # the first replay (graph capture) is untimed
            for a, s_ in zip(mats, src):
                a.copy_(s_)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            b.graph_launch(stream)
            e1.record(stream)
            torch.cuda.synchronize()
            if k:
                times.append(e0.elapsed_time(e1))
        st = b.stats()
        back = orth = 0.0
        for a, q, r in zip(src, mats, b.Rs):
            a64, q64, r64 = a.double(), q.double(), r.double()
            back = max(back, float(torch.linalg.norm(a64 - q64 @ r64) / torch.linalg.norm(a64)))
            orth = max(orth, gram_error(q64))
        ms = sum(times) / len(times)
        out[method] = {"ms_per_batch": round(ms, 4), "qrs_per_s": round(count / (ms / 1e3), 1),
                       "GBs": round(balg / (ms / 1e3) / 1e9, 1), "max_backward": float(f"{back:.3e}"),
                       "max_orth": float(f"{orth:.3e}"), "panels": st}
        b.close()
        print(f"# batched {method}: {out[method]}", file=sys.stderr, flush=True)
    out["mixed_over_plain_qrs"] = round(out["mixed"]["qrs_per_s"] / out["plain"]["qrs_per_s"], 3)
    return out


def cpu_sample_rows(m: int, n: int, frac: int) -> int:
    """Rows of the bounded CPU sample: m / frac, rounded down to whole
    27648-row chunks of the reference (NUM_SM * BLOCK_SIZE), all of m when
    m is small."""
    return max(n, (m // frac) // 27648 * 27648) if m >= 27648 * frac else m


def cpu_baseline(A: np.ndarray, frac: int) -> dict:
    """Times the oracle restatement and LAPACK on a bounded host sample A
    (Fortran-ordered ms x n, the first rows of the GPU's input)."""
    from oracle import oracle as O  # the cpu_baseline leg is the only bench user of the oracle

    ms, n = A.shape
    dtype = A.dtype.name
    cores = os.cpu_count() or 1
    esz = A.dtype.itemsize
    balg = algorithmic_bytes(ms, n, esz)
    t0 = time.perf_counter()
    ref = O.tsqr_ref if n <= 32 else O.qr_ref  # tsqr<T> / qr<T> of the reference
    ref(A, bs=256 if dtype == "float32" else 128, num_sm=108)
    t_port = time.perf_counter() - t0
    out = {
        "value": round(balg / t_port / 1e9, 3),
        "unit": "GB/s",
        "cores": cores,
        "kind": "port",
        "sample": f"the first {ms} rows x {n} {dtype} of the GPU's input (1/{frac} of the rows, copied from the "
                  f"device), oracle restatement of the reference {'tsqr' if n <= 32 else 'qr'} (NUM_SM=108, "
                  f"BLOCK_SIZE={256 if dtype == 'float32' else 128}), OpenMP over tiles",
        "seconds": round(t_port, 3),
        "cpu": cpu_model(),
        "algorithmic_bytes": balg,
    }
    try:
        import scipy.linalg.lapack as la
        from threadpoolctl import threadpool_limits

        with threadpool_limits(limits=cores):
            Af = np.asfortranarray(A)
            geqrf = la.sgeqrf if dtype == "float32" else la.dgeqrf
            orgqr = la.sorgqr if dtype == "float32" else la.dorgqr
            t0 = time.perf_counter()
            qr_, tau, _, _ = geqrf(Af, overwrite_a=True)
            orgqr(qr_, tau, overwrite_a=True)
            t_lapack = time.perf_counter() - t0
        out["lapack"] = {
            "value": round(balg / t_lapack / 1e9, 3),
            "unit": "GB/s",
            "seconds": round(t_lapack, 3),
            "what": f"OpenBLAS {'s' if dtype == 'float32' else 'd'}geqrf+{'s' if dtype == 'float32' else 'd'}orgqr "
                    f"on the same sample, {cores} threads",
        }
    except Exception as e:  # LAPACK is a reported extra, not required
        out["lapack"] = {"error": str(e)}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; the row-sharded multi-GPU TSQR "
                                   "(one R all-gather per call) is designed but not built"}))
        sys.exit(2)

    import torch

    import tsqr_amd

    if args.workload == "qr_mixed_1m_sweep":
        torch.cuda.set_device(0)
        sw = cond_sweep(args.steps)
        worst = max(c["mixed"]["orth"] for c in sw["cases"])
        print(json.dumps({"metric": "mixed-precision recursive QR: max ||I-Q^TQ||_F over cond in {1e2,1e4,1e6}",
                          "value": worst, "unit": "1", "n_gpus": 1, "steps": args.steps, "warmup": 1,
                          "ms_per_step": max(c["mixed"]["ms"] for c in sw["cases"]), "higher_is_better": False,
                          "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                          "data": "synthetic: U diag(s) V^T from seeded Gaussians",
                          "config": {"workload": "qr_mixed m=1048576 n in {64,128}"}, "sweep": sw}), flush=True)
        return
    if args.workload == "batch_512_fp32":
        # the batched leg alone: value = mixed-method QRs/s
        torch.cuda.set_device(0)
        b = batched_leg(args.steps, args.batch_count)
        print(json.dumps({"metric": "batched QR/s mixed vs FP32 at equal error", "value": b["mixed"]["qrs_per_s"],
                          "unit": "QR/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
                          "ms_per_step": b["mixed"]["ms_per_batch"], "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "f32", "data": "synthetic: iid N(0,1), seeded",
                          "config": {"workload": b["workload"]}, "batched": b}), flush=True)
        return
    from tsqr_amd._lib import (TSQR_DOUBLE, TSQR_FLOAT, TSQR_LEAF_AUTO, TSQR_LEAF_CHOLQR2, TSQR_LEAF_GRAM_TREE,
                               TSQR_LEAF_HOUSEHOLDER, leaf_stats, mixed_stats, set_leaf_path, timing_read)

    kind, m, n, dtype = WORKLOADS[args.workload]
    tdt = torch.float32 if dtype == "float32" else torch.float64
    esz = 4 if dtype == "float32" else 8
    code = TSQR_FLOAT if dtype == "float32" else TSQR_DOUBLE
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)

    print(f"# bench: {args.workload} {kind} m={m} n={n} {dtype}", file=sys.stderr, flush=True)
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    A = torch.empty((n, m), dtype=tdt, device=dev)
    A.uniform_(-1.0, 1.0, generator=g)  # column-major m x n with lda = m
    # the CPU baseline's sample: the first rows of this very array
    host_sample = None
    if not args.no_cpu:
        ms = cpu_sample_rows(m, n, args.cpu_fraction)
        host_sample = A[:, :ms].cpu().numpy().T  # Fortran-ordered ms x n
    R = torch.zeros((n, n), dtype=tdt, device=dev)
    L = tsqr_amd.lib()
    need = (L.qr_workspace_bytes if kind == "qr" else L.tsqr_workspace_bytes)(m, n, code)
    work = torch.empty(need, dtype=torch.uint8, device=dev)
    fn = getattr(L, f"{kind}_{'s' if code == TSQR_FLOAT else 'd'}_stream")
    is_tsqr = kind in ("tsqr", "tsqr_mixed")
    if args.leaf_path != "default":
        path = {"auto": TSQR_LEAF_AUTO, "householder": TSQR_LEAF_HOUSEHOLDER, "gram_tree": TSQR_LEAF_GRAM_TREE,
                "cholqr2": TSQR_LEAF_CHOLQR2}[args.leaf_path]
        set_leaf_path(code, path, 1e5 if (path == TSQR_LEAF_CHOLQR2 and code == TSQR_DOUBLE) else 1024.0)
    stream = torch.cuda.current_stream(dev)

    def step():
        st = fn(m, n, A.data_ptr(), m, R.data_ptr(), n, work.data_ptr(), need, stream.cuda_stream)
        if st != 0:
            raise tsqr_amd.TsqrError(st, "bench step")

    copy_gbs = None
    if args.copy_bytes > 0:
        src = torch.empty(args.copy_bytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        copy_gbs = round(2 * src.numel() * 4 * 5 / (e0.elapsed_time(e1) / 1e3) / 1e9, 1)
        print(f"# copy {src.numel() * 4} B: {copy_gbs} GB/s (read + write)", file=sys.stderr, flush=True)
        del src, dst

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # Per-kernel durations: the same K steps again with HIP event marks that
    # the library records on its launch stream around each phase.  They are
    # kept out of the pass timed above because a mark costs a few
    # microseconds of GPU time, which is noticeable next to the small
    # workloads' kernels.
    L.tsqr_timing_enable(1)
    torch.cuda.synchronize()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    phases = timing_read()
    L.tsqr_timing_enable(0)
    print(f"# phases (ms total, count): {phases}", file=sys.stderr, flush=True)

    sec = (t1 - t0) / args.steps
    # the recursive qr is counted like a TSQR of the whole m x n (the
    # coupling products' extra traffic is not credited)
    balg = algorithmic_bytes(m, n, esz)
    compulsory = (2 * m * n + n * n) * esz
    value = balg / sec / 1e9
    print(f"{kind} Latency: {sec * 1e3:f} ms", file=sys.stderr, flush=True)  # test_tsqr.cu:90

    # Algorithmic bytes per launch of the two streaming kernels: the leaf R
    # pass reads A (m*n elements), the leaf Q pass reads A and writes Q.
    panel_n = n if is_tsqr else 32
    alg = {"leaf_r": m * panel_n * esz, "leaf_q": 2 * m * panel_n * esz}
    kern = {}
    for name, nbytes in alg.items():
        tot_ms, cnt = phases[name]
        if cnt:
            avg_s = tot_ms / cnt / 1e3
            kern[name] = {"avg_launch_ms": round(avg_s * 1e3, 4), "algorithmic_bytes_per_launch": nbytes,
                          "achieved_GBs": round(nbytes / avg_s / 1e9, 1)}
    # Matrix-core work executed per launch.  FP64 (v_mfma_f64_16x16x4f64,
    # 2048 flops each): Gram pass 12 per 16 rows; Q pass 64 per 64 rows, 48
    # when M = R^-1 is triangular (flat and mixed paths).  The FP32 flat path
    # runs both passes on v_mfma_f64_4x4x4_4b_f64 (512 flops each): one per
    # pair of 4-column groups I <= J per 16 rows, i.e. 32 * NC (NC + 1) / 2
    # flops per row with NC = padded width / 4.  BF16
    # (v_mfma_f32_16x16x32_bf16, 16384 flops each, mixed path): six per 16x16
    # block per 32 rows, three blocks for n > 16, in the Gram pass and again
    # for the residual check fused into the Q pass.
    mixed = None
    if kind == "tsqr":
        stats = leaf_stats(work.data_ptr(), m, n, code)
        gram_rows = m if stats[1] == 0 else 0
        if code == TSQR_FLOAT and stats[2]:
            nc = (8 if n <= 8 else (16 if n <= 16 else 32)) // 4
            per_row = 32 * nc * (nc + 1) // 2
            f64 = {"leaf_r": gram_rows * per_row, "leaf_q": gram_rows * per_row}
        else:
            f64 = {"leaf_r": gram_rows / 16 * 12 * 2048,
                   "leaf_q": gram_rows / 64 * (48 if stats[2] else 64) * 2048}
        bf16 = {}
    elif kind == "tsqr_mixed":
        mixed = mixed_stats(work.data_ptr(), m, n)
        stats = (0, 0, mixed["flat"])
        rows = m if mixed["flat"] else 0
        blocks = 3 if n > 16 else 1
        nt = 2 if n > 16 else 1
        # per pass: the Gram of the stored Q (BF16x3, 6 MFMAs of 16x16x32 per 32-row step and block), and Q = A X
        # either as BF16x3 products (4 row tiles x nt column tiles x 6 MFMAs per 64 rows) or on
        # v_mfma_f64_16x16x4f64 (48 per 64 rows at n = 32: X is upper triangular)
        q_f64 = q_bf16 = 0.0
        for p_ in range(mixed["passes"]):
            if mixed["bf16_q"][p_]:
                q_bf16 += rows / 64 * 4 * nt * 6 * 16384
            else:
                q_f64 += rows / 64 * 48 * 2048
        f64 = {"leaf_q": q_f64}
        bf16 = {"leaf_r": rows / 32 * 6 * blocks * 16384,
                "leaf_q": rows / 32 * 6 * blocks * 16384 * mixed["passes"] + q_bf16}
    else:
        stats, f64, bf16 = (0, 0, 0), {}, {}
    for name in kern:
        secs = kern[name]["avg_launch_ms"] / 1e3
        if f64.get(name):
            tf = f64[name] / secs / 1e12
            kern[name]["fp64_mfma_TFLOPs"] = round(tf, 1)
            kern[name]["fp64_mfma_frac"] = round(tf / FP64_MFMA_PEAK_TFLOPS, 4)
        if bf16.get(name):
            tf = bf16[name] / secs / 1e12
            kern[name]["bf16_mfma_TFLOPs"] = round(tf, 1)
            kern[name]["bf16_mfma_frac"] = round(tf / BF16_MFMA_PEAK_TFLOPS, 4)
    roofline = None
    if kern:
        dom = max(kern, key=lambda k: kern[k]["avg_launch_ms"])
        achieved = kern[dom]["achieved_GBs"]
        wl = f"{kind} m={m} n={n} {dtype}"
        tr = measured_traffic(wl)
        roofline = {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": tr.get(dom) if tr else None,
            "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT) if tr else None,
            "avg_launch_ms": kern[dom]["avg_launch_ms"],
            "algorithmic_bytes_per_launch": kern[dom]["algorithmic_bytes_per_launch"],
            "kernels": kern,
        }
    phase_ms = {k: round(v[0] / max(v[1], 1), 4) for k, v in phases.items() if v[1]}

    # Errors of one untimed factorisation of the same input, from the
    # library's device-side FP64 checker (tsqr_check_*: n x n Gram reductions
    # of A, Q and A - QR over all m rows).
    errors = None
    if is_tsqr and n <= 32:
        from tsqr_amd._lib import check

        g.manual_seed(0)
        A.uniform_(-1.0, 1.0, generator=g)
        A0 = A.clone()
        step()
        torch.cuda.synchronize()
        errors = check(m, n, code, A0.data_ptr(), m, A.data_ptr(), m, R.data_ptr(), n, stream.cuda_stream)
        errors = {k: float(f"{v:.3e}") for k, v in errors.items()}
        errors["source"] = "tsqr_check (device FP64)"
        del A0

    out = {
        "metric": "TSQR HBM GB/s (% of peak); batched QR/s mixed vs FP32 at equal error",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if dtype == "float32" else "f64",
        "data": "synthetic: iid uniform[-1,1) from torch.Generator seed 0, resident in HBM",
        "config": {
            "workload": f"{kind} m={m} n={n} {dtype} (TSQR at the reference's maximum height)"
            if args.workload == "tsqr_max_fp32" else f"{kind} m={m} n={n} {dtype}",
            "m": m,
            "n": n,
            "algorithmic_bytes_per_step": balg,
            "compulsory_bytes_per_step": compulsory,
            "compulsory_GBs": round(compulsory / sec / 1e9, 2),
            "frac_of_hbm_peak": round(value / HBM_PEAK_GBS, 4),
            "phase_ms": phase_ms,
            "device_copy_GBs": copy_gbs,
            "superleaves": dict(zip(("gram", "householder", "flat"), stats)) if kind == "tsqr" else None,
            "mixed": mixed,
            "errors": errors,
        },
        "roofline": roofline,
    }
    if not args.no_batch and args.workload == "tsqr_max_fp32":
        out["batched"] = batched_leg(args.batch_steps, args.batch_count)
    if not args.no_cpu:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(host_sample, args.cpu_fraction)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
