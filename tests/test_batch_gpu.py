"""GPU tests of the batched TSQR (tsqr_batch_* through the C ABI, via the
tsqr_amd.Batch wrapper) against the FP64 LAPACK oracle, matrix by matrix.

Bar per matrix, u the unit roundoff of the working precision:
  plain (Householder panels): backward and orthogonality <= 8 n u
  mixed (float32): backward <= 8 n u, orthogonality <= 1e-6
  R upper triangular (R21 = 0 for two-panel matrices) and the reference's
  all_close(QR, A) (utils.h:352-357)
"""
import numpy as np
import pytest

from oracle import oracle as O
from tsqr_amd.workloads import batch_shapes

pytestmark = pytest.mark.gpu

# ragged, single-tile, one- and two-panel shapes, n = 1 .. 64
SHAPES = [(4096, 8), (4160, 16), (9001, 32), (20000, 64), (200, 8), (100, 40), (3000, 1), (12345, 33),
          (70000, 20), (257, 64), (2048, 32), (50000, 64), (6000, 7), (16384, 48), (1024, 64), (100000, 32)]


def make_mats(shapes, dtype, seed=0):
    return [O.gen_normal(m, n, dtype, seed=seed * 1000 + i) for i, (m, n) in enumerate(shapes)]


def to_gpu(mats):
    import torch

    from tsqr_amd import colmajor

    return [colmajor(torch.from_numpy(np.ascontiguousarray(a)).cuda()) for a in mats]


def run(mats, method, graph=False, replays=1):
    import torch

    import tsqr_amd

    dev = to_gpu(mats)
    b = tsqr_amd.Batch(dev, method=method)
    for _ in range(replays):
        (b.graph_launch if graph else b.run)()
    torch.cuda.synchronize()
    stats = b.stats()
    Qs = [d.cpu().numpy() for d in dev]
    Rs = [r.cpu().numpy() for r in b.Rs]
    b.close()
    return Qs, Rs, stats


def check(A, Q, R, method, tag):
    n = A.shape[1]
    u = O.unit_roundoff(A.dtype)
    m = O.metrics(A, Q, R)
    orth_bound = 1e-6 if method == "mixed" else 8 * n * u
    print(f"{tag}: backward={m['backward']:.3e} orth={m['orth']:.3e}", flush=True)
    assert m["all_close"], tag
    assert m["r_lower_zero"], tag
    assert m["backward"] <= 8 * n * u, tag
    assert m["orth"] <= orth_bound, tag


@pytest.mark.parametrize("method,dtype", [("plain", np.float32), ("mixed", np.float32), ("plain", np.float64)])
def test_batch_uneven_shapes(method, dtype):
    mats = make_mats(SHAPES, dtype, seed=1)
    Qs, Rs, stats = run(mats, method)
    print("stats", stats, flush=True)
    for i, (A, Q, R) in enumerate(zip(mats, Qs, Rs)):
        check(A, Q, R, method, f"{method} {np.dtype(dtype).name} #{i} {A.shape}")
    panels = sum(1 + (n > 32) for _, n in SHAPES)
    assert stats["cholesky"] + stats["householder"] == panels
    if method == "mixed":
        assert stats["check_failed"] == 0
        assert stats["cholesky"] == panels  # well-conditioned inputs stay on the CholeskyQR path
        # iid N(0,1) panels at least 20n tall: kappa_est <= 2n, so Q is formed on the BF16 matrix cores
        tall = sum(1 + (n > 32) for m, n in SHAPES if m >= 20 * min(n, 32))
        assert stats["bf16_q_pass1"] >= tall
    else:
        assert stats["householder"] == panels
        assert stats["bf16_q_pass1"] == 0


def test_batch_mixed_ill_conditioned_member_falls_back():
    shapes = [(30000, 32), (30000, 32), (40000, 64)]
    mats = make_mats(shapes, np.float32, seed=2)
    mats[1] = O.gen_conditioned(30000, 32, 1e6, np.float32, seed=3)
    Qs, Rs, stats = run(mats, "mixed")
    print("stats", stats, flush=True)
    assert stats["householder"] >= 1
    for i, (A, Q, R) in enumerate(zip(mats, Qs, Rs)):
        check(A, Q, R, "mixed" if i != 1 else "plain", f"mixed-fallback #{i}")


@pytest.mark.parametrize("method", ["plain", "mixed"])
def test_batch_graph_replays_match_eager(method):
    """The captured graph, replayed three times with eager copies of the
    input in between, gives bit-identical results to the eager launches."""
    import torch

    import tsqr_amd

    shapes = batch_shapes(count=24, m_lo=4096, m_hi=65536, seed=5)
    mats = make_mats(shapes, np.float32, seed=4)
    Qe, Re, _ = run(mats, method)
    dev = to_gpu(mats)
    src = [d.clone() for d in dev]
    b = tsqr_amd.Batch(dev, method=method)
    for rep in range(3):
        for d, s in zip(dev, src):
            d.copy_(s)
        torch.cuda.synchronize()
        b.graph_launch()
        torch.cuda.synchronize()
        for i, (d, r) in enumerate(zip(dev, b.Rs)):
            assert np.array_equal(d.cpu().numpy(), Qe[i]), (rep, i)
            assert np.array_equal(r.cpu().numpy(), Re[i]), (rep, i)
    b.close()


def test_batch_status_codes():
    import ctypes

    import torch

    import tsqr_amd
    from tsqr_amd import TsqrError

    with pytest.raises(TsqrError) as e:
        tsqr_amd.Batch([tsqr_amd.colmajor(torch.zeros(1000, 65, device="cuda"))])
    assert e.value.status == -2
    with pytest.raises(TsqrError) as e:
        tsqr_amd.Batch([tsqr_amd.colmajor(torch.zeros(1000, 8, dtype=torch.float64, device="cuda"))],
                       method="mixed")
    assert e.value.status == -1
    b = tsqr_amd.Batch([tsqr_amd.colmajor(torch.randn(5000, 16, device="cuda"))])
    st = tsqr_amd.lib().tsqr_batch_run(b._h, b.work.data_ptr(), b.work.numel() - 1, None)
    assert st == -3
    st = tsqr_amd.lib().tsqr_batch_run(b._h, ctypes.c_void_p(0), b.work.numel(), None)
    assert st == -1
    b.close()


def test_batch_mixed_reorthogonalises_dependent_panels():
    """Columns 32..63 nearly in the span of columns 0..31: one projection
    removes almost all of their norm, so the mixed method must take the
    second coupling pass; without it Q2 would be orthogonal to Q1 only to
    ~1e-7 * 1e3.  A well-separated matrix in the same batch skips it."""
    rng = np.random.Generator(np.random.PCG64(7))
    m = 20000
    A1 = rng.standard_normal((m, 32))
    A2 = A1 @ rng.standard_normal((32, 32)) + 1e-3 * rng.standard_normal((m, 32))
    dep = np.asfortranarray(np.hstack([A1, A2]).astype(np.float32))
    ind = O.gen_normal(m, 64, np.float32, seed=8)
    mats = [dep, ind]
    for method in ("mixed", "plain"):
        Qs, Rs, stats = run(mats, method)
        for i, (A, Q, R) in enumerate(zip(mats, Qs, Rs)):
            if method == "plain" and i == 0:
                # one block Gram-Schmidt pass (the reference's coupling) loses
                # orthogonality here; only the backward error is bounded
                mt = O.metrics(A, Q, R)
                print(f"plain dependent: backward={mt['backward']:.3e} orth={mt['orth']:.3e}", flush=True)
                assert mt["backward"] <= 8 * 64 * O.unit_roundoff(np.float32)
                continue
            check(A, Q, R, method, f"{method} reorth #{i}")
