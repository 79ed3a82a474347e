"""The stream-ordered entry points as a caller that overlaps work or
captures graphs uses them: the reference keeps one module-global
__device__ counter (kernelQR.h:297), so two of its calls in flight corrupt
each other's barrier and no call can be captured.  Here all state lives in
the caller's workspace and a call enqueues kernels only (no memset, no host
synchronisation, no allocation):

  - a call on a non-null stream, captured into a hipGraph (torch.cuda.graph
    over that stream) and replayed three times with eager work and device
    synchronisations between replays, gives results bit-identical to the
    eager call, for tsqr_s/d, tsqr_mixed_s, qr_s/d, qr_mixed_s and tsqr_r_s,
    including an input whose gate rejects the flat path (Householder
    fallback kernels inside the graph);
  - calls on two streams at once with separate workspaces give the results
    of the same calls run one after the other;
  - the reference-shaped entry points, whose short ldwork makes them share
    the library's internal buffer, called from two host threads at once,
    give the results of the stream calls.
"""
import threading

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _setup(kind, m, n, dtype, seed, cond=None):
    import torch

    import tsqr_amd

    tdt = torch.float32 if dtype == np.float32 else torch.float64
    if cond is None:
        A = O.gen_normal(m, n, dtype, seed=seed)
    else:
        A = O.gen_conditioned(m, n, cond, dtype, seed=seed)
    src = tsqr_amd.colmajor(torch.from_numpy(np.ascontiguousarray(A)).cuda())
    code = 1 if dtype == np.float32 else 2
    ws = (tsqr_amd.qr_workspace_bytes if kind.startswith("qr") else tsqr_amd.tsqr_workspace_bytes)(m, n, code)
    return src, tdt, ws


def _fn(kind, dtype):
    import tsqr_amd

    L = tsqr_amd.lib()
    suf = "s" if dtype == np.float32 else "d"
    return getattr(L, {"tsqr": f"tsqr_{suf}_stream", "qr": f"qr_{suf}_stream", "tsqr_mixed": "tsqr_mixed_s_stream",
                       "qr_mixed": "qr_mixed_s_stream", "tsqr_r": f"tsqr_r_{suf}_stream"}[kind])


class Call:
    """One factorisation with its own A, R and workspace on a given stream."""

    def __init__(self, kind, src, tdt, ws, stream):
        import torch

        import tsqr_amd

        m, n = src.shape
        self.kind, self.m, self.n, self.src, self.stream = kind, m, n, src, stream
        self.A = src.clone(memory_format=torch.preserve_format)
        self.R = tsqr_amd.empty_colmajor(n, n, tdt, src.device)
        self.R.fill_(float("nan"))
        self.work = torch.empty(max(ws, 1), dtype=torch.uint8, device=src.device)
        self.fn = _fn(kind, np.float32 if tdt == torch.float32 else np.float64)

    def reset(self):
        import torch

        with torch.cuda.stream(self.stream):
            self.A.copy_(self.src)
            self.R.fill_(float("nan"))

    def enqueue(self):
        st = self.fn(self.m, self.n, self.A.data_ptr(), self.A.stride(1), self.R.data_ptr(), self.R.stride(1),
                     self.work.data_ptr(), self.work.numel(), self.stream.cuda_stream)
        assert st == 0, st

    def result(self):
        import torch

        torch.cuda.synchronize()
        return self.A.cpu().numpy().copy(), self.R.cpu().numpy().copy()


GRAPH_CASES = [
    ("tsqr", 100000, 32, np.float32, None),
    ("tsqr", 27648, 32, np.float32, None),
    ("tsqr", 70001, 20, np.float32, 1e6),  # flat gate rejects: Householder tree inside the graph
    ("tsqr", 100000, 32, np.float64, None),
    ("tsqr", 60000, 16, np.float64, 1e9),  # FP64 CholeskyQR2 gate rejects
    ("tsqr_mixed", 100000, 32, np.float32, None),
    ("tsqr_r", 100000, 32, np.float32, None),
    ("qr", 27648, 128, np.float32, None),
    ("qr", 27648, 128, np.float64, None),
    ("qr_mixed", 50000, 64, np.float32, None),
]


@pytest.mark.parametrize("kind,m,n,dtype,cond", GRAPH_CASES)
def test_graph_capture_replays_bit_identical(kind, m, n, dtype, cond):
    import torch

    print(f"graph {kind} {m}x{n} {np.dtype(dtype).name} cond={cond}", flush=True)
    src, tdt, ws = _setup(kind, m, n, dtype, seed=m + n, cond=cond)
    s = torch.cuda.Stream()
    eager = Call(kind, src, tdt, ws, s)
    eager.enqueue()
    Qe, Re = eager.result()
    if kind == "tsqr_r":
        assert np.array_equal(Qe, src.cpu().numpy())  # A untouched
    c = Call(kind, src, tdt, ws, s)
    c.reset()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        c.enqueue()
    other = Call(kind, src, tdt, ws, s)  # eager work between the replays
    for rep in range(3):
        c.reset()
        with torch.cuda.stream(s):
            g.replay()
        Qg, Rg = c.result()
        assert np.array_equal(Qg, Qe), f"replay {rep}: Q differs"
        assert np.array_equal(Rg, Re, equal_nan=False), f"replay {rep}: R differs"
        other.reset()
        other.enqueue()
        torch.cuda.synchronize()
    del g


STREAM_CASES = [
    ("tsqr", 200000, 32, np.float32),
    ("tsqr", 50000, 24, np.float64),
    ("qr", 27648, 128, np.float32),
    ("tsqr_mixed", 150000, 32, np.float32),
]


@pytest.mark.parametrize("kind,m,n,dtype", STREAM_CASES)
def test_two_streams_at_once(kind, m, n, dtype):
    import torch

    print(f"two streams {kind} {m}x{n} {np.dtype(dtype).name}", flush=True)
    src1, tdt, ws = _setup(kind, m, n, dtype, seed=1)
    src2, _, _ = _setup(kind, m, n, dtype, seed=2)
    s0 = torch.cuda.Stream()
    ref = []
    for src in (src1, src2):
        c = Call(kind, src, tdt, ws, s0)
        c.enqueue()
        ref.append(c.result())
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    c1, c2 = Call(kind, src1, tdt, ws, s1), Call(kind, src2, tdt, ws, s2)
    torch.cuda.synchronize()
    for _ in range(3):
        c1.reset()
        c2.reset()
        torch.cuda.synchronize()
        c1.enqueue()
        c2.enqueue()
        for c, (Qr, Rr) in zip((c1, c2), ref):
            Q, R = c.result()
            assert np.array_equal(Q, Qr) and np.array_equal(R, Rr)


def test_reference_entry_from_two_threads():
    """tsqr_s with a short ldwork (the library's internal buffer) from two
    host threads at once, each on its own matrix, repeatedly."""
    import torch

    import tsqr_amd

    L = tsqr_amd.lib()
    m, n = 300000, 32
    srcs = [_setup("tsqr", m, n, np.float32, seed=10 + k)[0] for k in range(2)]
    ref = []
    for src in srcs:
        _, tdt, ws = _setup("tsqr", m, n, np.float32, seed=0)
        c = Call("tsqr", src, tdt, ws, torch.cuda.current_stream())
        c.enqueue()
        ref.append(c.result())
    outs = [[None] * 4 for _ in srcs]
    errors = []

    def worker(k):
        try:
            torch.cuda.set_device(0)
            A = srcs[k].clone(memory_format=torch.preserve_format)
            R = tsqr_amd.empty_colmajor(n, n, torch.float32, A.device)
            for it in range(4):
                torch.cuda.synchronize()
                A.copy_(srcs[k])
                torch.cuda.synchronize()
                st = L.tsqr_s(m, n, A.data_ptr(), m, R.data_ptr(), n, None, 0)
                assert st == 0, st
                torch.cuda.synchronize()
                outs[k][it] = (A.cpu().numpy().copy(), R.cpu().numpy().copy())
        except Exception as e:  # reported in the main thread
            errors.append(e)

    th = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=100)
    assert not errors, errors
    for k in range(2):
        for it in range(4):
            Q, R = outs[k][it]
            assert np.array_equal(Q, ref[k][0]) and np.array_equal(R, ref[k][1]), (k, it)
