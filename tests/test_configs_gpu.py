"""The benchmark configurations themselves, at full size, through the C ABI.

  - maximum height (FP32 23,887,872 x 32 and FP64 5,971,968 x 32, the
    reference's largest m, TallShinnyQR.h:100 and its README): the
    harness's uniform stream (test_tsqr.cu:24-28, oracle generator), checked
    by the device FP64 checker (backward and orthogonality <= 8 n u) and R
    against the Cholesky factor of an FP64 Gram matrix formed by rocBLAS
    (torch), an implementation independent of this library's kernels, after
    row-sign normalisation; the FP32 case also on the Householder path (the
    reference's method) at full height;
  - mixed-precision recursive QR at m = 1,048,576, n in {64, 128},
    A = U diag(s) V^T with s geometric from 1 to 1/cond, cond in
    {1e2, 1e4, 1e6}: backward error and orthogonality against the FP64
    definitions on the GPU, bar 1e-6 (the configuration's target), with the
    plain qr_s (one coupling pass, block_QR.h:88-93) recorded beside it;
  - the 512-matrix uneven batch (tsqr_amd.workloads.batch_shapes): every
    matrix checked in FP64 on the GPU, mixed method <= 1e-6, plain FP32
    Householder <= 8 n u, R upper triangular.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _dev_colmajor(A):
    """Fortran-ordered host array -> column-major m x n GPU tensor."""
    import torch

    return torch.from_numpy(np.ascontiguousarray(A.T)).cuda().t()


def _gram64(A, rows=1 << 21):
    import torch

    m, n = A.shape
    G = torch.zeros((n, n), dtype=torch.float64, device=A.device)
    for r0 in range(0, m, rows):
        X = A[r0:r0 + rows].double()
        G += X.t() @ X
    return G


def _signed(R):
    d = np.sign(np.diag(R))
    d[d == 0] = 1
    return R * d[:, None]


@pytest.mark.parametrize("m,dtype,path", [(23887872, np.float32, "default"), (5971968, np.float64, "default"),
                                          (23887872, np.float32, "householder")])
def test_max_height(m, dtype, path):
    import torch

    import tsqr_amd
    from tests.gpu_util import LeafPath

    n = 32
    u = O.unit_roundoff(dtype)
    print(f"max height {m}x{n} {np.dtype(dtype).name} path={path}", flush=True)
    A0 = _dev_colmajor(O.gen_uniform(m, n, dtype))
    A = A0.clone(memory_format=torch.preserve_format)
    R = tsqr_amd.empty_colmajor(n, n, A.dtype, A.device)
    if path == "householder":
        with LeafPath(dtype, 1):
            tsqr_amd.tsqr(A, R)
            torch.cuda.synchronize()
    else:
        tsqr_amd.tsqr(A, R)
    torch.cuda.synchronize()
    e = tsqr_amd.check(A0, A, R)
    print(f"  checker: {e}", flush=True)
    assert e["backward_fro"] <= 8 * n * u and e["orth_fro"] <= 8 * n * u
    assert e["backward_2"] <= 8 * n * u and e["orth_2"] <= 8 * n * u
    Rh = R.cpu().numpy().astype(np.float64)
    assert np.all(np.tril(Rh, -1) == 0)
    G = _gram64(A0).cpu().numpy()
    Rref = np.linalg.cholesky(G).T
    rerr = np.linalg.norm(_signed(Rh) - _signed(Rref)) / np.linalg.norm(Rref)
    print(f"  R vs chol(FP64 Gram): {rerr:.3e}", flush=True)
    assert rerr <= 8 * n * u


def _conditioned(m, n, cond, seed):
    """U diag(s) V^T in FP64 on the GPU: U, V from torch's FP64 QR of seeded
    Gaussians, s geometric from 1 to 1/cond."""
    import torch

    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    U, _ = torch.linalg.qr(torch.randn(m, n, dtype=torch.float64, device=dev, generator=g))
    V, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g))
    s = cond ** (-torch.arange(n, dtype=torch.float64, device=dev) / (n - 1))
    return (U * s) @ V.T


def _errors64(A32, Q, R):
    """||A - QR||_F / ||A||_F and ||I - Q^T Q||_F in FP64 on the GPU."""
    import torch

    a, q, r = A32.double(), Q.double(), torch.triu(R.double())
    back = float(torch.linalg.norm(a - q @ r) / torch.linalg.norm(a))
    gq = _gram64(Q)
    orth = float(torch.linalg.norm(gq - torch.eye(Q.shape[1], dtype=torch.float64, device=Q.device)))
    return back, orth


MIXED_BAR = 1e-6


@pytest.mark.parametrize("n", [64, 128])
@pytest.mark.parametrize("cond", [1e2, 1e4, 1e6])
def test_mixed_recursive_qr_1m(n, cond):
    import torch

    import tsqr_amd

    m = 1048576
    print(f"qr_mixed {m}x{n} cond={cond:g}", flush=True)
    src = tsqr_amd.colmajor(_conditioned(m, n, cond, seed=n * 7 + int(np.log10(cond))).float())
    out = {}
    for name, fn in (("mixed", tsqr_amd.qr_mixed), ("plain", tsqr_amd.qr)):
        A = src.clone(memory_format=torch.preserve_format)
        R = tsqr_amd.empty_colmajor(n, n, torch.float32, A.device)
        fn(A, R)
        torch.cuda.synchronize()
        assert bool(torch.all(torch.tril(R, -1) == 0))
        out[name] = _errors64(src, A, R)
        del A
    print(f"  mixed backward={out['mixed'][0]:.3e} orth={out['mixed'][1]:.3e} | "
          f"plain backward={out['plain'][0]:.3e} orth={out['plain'][1]:.3e}", flush=True)
    assert out["mixed"][0] <= MIXED_BAR
    assert out["mixed"][1] <= MIXED_BAR
    assert out["plain"][0] <= 8 * n * O.unit_roundoff(np.float32)  # the plain path stays backward stable


@pytest.mark.parametrize("method", ["mixed", "plain"])
def test_batch_512(method):
    import torch

    import tsqr_amd
    from tsqr_amd.workloads import batch_shapes

    shapes = batch_shapes()
    assert len(shapes) == 512
    print(f"batch 512 {method}", flush=True)
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev)
    g.manual_seed(1)
    src = [tsqr_amd.empty_colmajor(m, n, torch.float32, dev) for m, n in shapes]
    for a in src:
        a.normal_(generator=g)
    mats = [a.clone(memory_format=torch.preserve_format) for a in src]
    b = tsqr_amd.Batch(mats, method=method)
    b.graph_launch()
    torch.cuda.synchronize()
    stats = b.stats()
    print(f"  panels {stats}", flush=True)
    worst_b = worst_o = 0.0
    for i, (a, q, r) in enumerate(zip(src, mats, b.Rs)):
        n = a.shape[1]
        assert bool(torch.all(torch.tril(r, -1) == 0)), i
        back, orth = _errors64(a, q, r)
        worst_b, worst_o = max(worst_b, back), max(worst_o, orth)
        bar = MIXED_BAR if method == "mixed" else 8 * n * O.unit_roundoff(np.float32)
        assert back <= bar and orth <= bar, (i, shapes[i], back, orth)
    print(f"  worst backward={worst_b:.3e} orth={worst_o:.3e}", flush=True)
    assert stats["check_failed"] == 0
    b.close()
