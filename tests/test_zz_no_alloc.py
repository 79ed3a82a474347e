"""tsqr_s_stream_noalloc, the entry point that allocates nothing (its
workspace is the library's fixed per-GPU scratch array), against
tsqr_s_stream with a caller workspace: the same kernels, so Q and R must be
bit-identical.  Two calls in a row (the second on a smaller problem, so that
it runs over what the first left in the array), an input whose gate rejects
the flat path (the Householder kernels, which use more of the workspace),
and the rejection of a problem above the array's size.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _pair(A, stream=None):
    """(Q, R) from tsqr_s_stream with its own workspace and from
    tsqr_s_stream_noalloc, on copies of A."""
    import torch

    import tsqr_amd
    from tsqr_amd._lib import lib

    L = lib()
    m, n = A.shape
    src = tsqr_amd.colmajor(torch.from_numpy(np.ascontiguousarray(A)).cuda())
    st = (stream or torch.cuda.current_stream()).cuda_stream
    out = []
    for noalloc in (False, True):
        Ad = src.clone(memory_format=torch.preserve_format)
        R = torch.full((n, n), float("nan"), dtype=torch.float32, device="cuda")
        if noalloc:
            s = L.tsqr_s_stream_noalloc(m, n, Ad.data_ptr(), m, R.data_ptr(), n, st)
        else:
            need = L.tsqr_workspace_bytes(m, n, 1)
            work = torch.empty(need, dtype=torch.uint8, device="cuda")
            s = L.tsqr_s_stream(m, n, Ad.data_ptr(), m, R.data_ptr(), n, work.data_ptr(), need, st)
        assert s == 0, s
        torch.cuda.synchronize()
        out.append((Ad.cpu().numpy(), R.cpu().numpy().T.copy()))
    return out


def test_noalloc_matches_stream_entry_twice():
    # first call, then a smaller one over what the first left in the array
    for m, n, seed in ((190000, 32, 1), (27648, 20, 2)):
        print(f"noalloc {m}x{n}", flush=True)
        A = O.gen_normal(m, n, np.float32, seed=seed)
        (Q1, R1), (Q2, R2) = _pair(A)
        assert np.array_equal(Q1, Q2) and np.array_equal(R1, R2)
        mt = O.metrics(A, Q2, R2)
        assert mt["backward"] <= 8 * n * O.unit_roundoff(np.float32)
        assert mt["orth"] <= 8 * n * O.unit_roundoff(np.float32)


def test_noalloc_householder_fallback():
    print("noalloc fallback 100000x32 cond 1e6", flush=True)
    A = O.gen_conditioned(100000, 32, 1e6, np.float32, seed=7)
    (Q1, R1), (Q2, R2) = _pair(A)
    assert np.array_equal(Q1, Q2) and np.array_equal(R1, R2)
    mt = O.metrics(A, Q2, R2)
    assert mt["orth"] <= 8 * 32 * O.unit_roundoff(np.float32)


def test_noalloc_non_null_stream():
    import torch

    A = O.gen_normal(60000, 32, np.float32, seed=3)
    (Q1, R1), (Q2, R2) = _pair(A, torch.cuda.Stream())
    assert np.array_equal(Q1, Q2) and np.array_equal(R1, R2)


def test_noalloc_rejects_problems_above_the_array():
    import torch

    from tsqr_amd._lib import TSQR_ERR_SCRATCH, lib

    L = lib()
    m, n = 1 << 20, 32
    assert L.tsqr_workspace_bytes(m, n, 1) > 16 << 20
    A = torch.zeros((n, m), dtype=torch.float32, device="cuda")
    R = torch.zeros((n, n), dtype=torch.float32, device="cuda")
    st = L.tsqr_s_stream_noalloc(m, n, A.data_ptr(), m, R.data_ptr(), n, None)
    assert st == TSQR_ERR_SCRATCH
    assert L.tsqr_status_string(st).decode().startswith("problem needs more workspace")
    torch.cuda.synchronize()
    assert float(A.abs().max()) == 0.0  # nothing ran
    # argument errors come first, as in tsqr_s_stream
    assert L.tsqr_s_stream_noalloc(10, 20, A.data_ptr(), 10, R.data_ptr(), 20, None) == -1
