"""Device-side FP64 checker (tsqr_check_s/d) against the oracle's host FP64
metrics (oracle.metrics, the restatement of utils.h:317-455) on the same
factorisations, through the C ABI."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _factor_and_check(A, lda=None, perturb=None):
    import torch

    import tsqr_amd
    from tests.gpu_util import from_device_colmajor, to_device_colmajor, torch_dtype
    from tsqr_amd._lib import TSQR_DOUBLE, TSQR_FLOAT, check

    L = tsqr_amd.lib()
    m, n = A.shape
    code = TSQR_FLOAT if A.dtype == np.float32 else TSQR_DOUBLE
    suf = "s" if code == TSQR_FLOAT else "d"
    keep, pK, ldk = to_device_colmajor(A, lda)
    work_buf, pQ, ldq = to_device_colmajor(A, lda)
    R = torch.zeros((n * n,), dtype=torch_dtype(A.dtype), device="cuda")
    need = L.tsqr_workspace_bytes(m, n, code)
    ws = torch.empty(max(need, 1), dtype=torch.uint8, device="cuda")
    st = getattr(L, f"tsqr_{suf}_stream")(m, n, pQ, ldq, R.data_ptr(), n, ws.data_ptr(), need, None)
    assert st == 0
    if perturb is not None:
        i, j, delta = perturb
        work_buf.view(n, ldq)[j, i] += delta
    torch.cuda.synchronize()
    dev = check(m, n, code, pK, ldk, pQ, ldq, R.data_ptr(), n)
    Q = from_device_colmajor(work_buf, m, n, ldq)
    Rh = R.view(n, n).cpu().numpy().T.copy()
    return dev, O.metrics(A, Q, Rh)


def _agree(dev, host, what, rel=1e-6, floor=1e-15):
    assert abs(dev - host) <= rel * abs(host) + floor, f"{what}: device {dev:.6e} host {host:.6e}"


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("m,n", [(27648, 32), (4100, 16), (1000, 7), (300, 32), (64, 1), (70000, 24)])
def test_check_matches_host_metrics(m, n, dtype):
    A = np.random.default_rng(m * 37 + n).standard_normal((m, n)).astype(dtype)
    dev, host = _factor_and_check(A)
    # Both sides form Q^T Q in FP64, in different orders: for an FP64 Q,
    # whose ||I - Q^T Q|| is itself a few u64, the two figures differ by
    # their own roundoff, up to about n u64.
    orth_floor = max(1e-15, n * 2.0 ** -53)
    _agree(dev["backward_fro"], host["backward"], "backward_fro")
    _agree(dev["orth_fro"], host["orth"], "orth_fro", floor=orth_floor)
    _agree(dev["backward_2"], host["spec_backward"], "backward_2", rel=1e-5)
    _agree(dev["orth_2"], host["spec_orth"] * n, "orth_2", rel=1e-5, floor=orth_floor)


def test_check_unaligned_leading_dimension():
    A = np.random.default_rng(5).standard_normal((5003, 32)).astype(np.float32)
    dev, host = _factor_and_check(A, lda=5007)
    _agree(dev["backward_fro"], host["backward"], "backward_fro")
    _agree(dev["orth_fro"], host["orth"], "orth_fro")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_check_sees_a_planted_error(dtype):
    # one entry of Q changed by delta after the factorisation: column 17 of
    # Q^T Q moves by about |q| delta ~ 1e-5 per entry (|q| ~ m^-1/2), far
    # above the roundoff level, and the device and host figures still agree
    A = np.random.default_rng(11).standard_normal((9000, 32)).astype(dtype)
    dev, host = _factor_and_check(A, perturb=(4321, 17, 1e-3))
    assert dev["orth_fro"] > 2e-5 and dev["backward_fro"] > 1e-6
    _agree(dev["backward_fro"], host["backward"], "backward_fro")
    _agree(dev["orth_fro"], host["orth"], "orth_fro")
    _agree(dev["orth_2"], host["spec_orth"] * 32, "orth_2", rel=1e-5)


def test_check_rejects_bad_arguments():
    import tsqr_amd

    L = tsqr_amd.lib()
    import ctypes

    norms = (ctypes.c_double * 4)()
    assert L.tsqr_check_s(10, 33, 1, 10, 1, 10, 1, 33, norms, None) != 0  # n > 32
    assert L.tsqr_check_s(10, 4, None, 10, 1, 10, 1, 4, norms, None) != 0  # null A
    assert L.tsqr_check_d(10, 4, 1, 9, 1, 10, 1, 4, norms, None) != 0  # lda < m


def test_torch_wrappers_r_only_and_check():
    """tsqr_amd.tsqr_r leaves A alone and matches tsqr's R; tsqr_amd.check
    on tsqr's output agrees with the host metrics."""
    import torch

    import tsqr_amd

    A_h = np.random.default_rng(17).standard_normal((40000, 32)).astype(np.float32)
    A = tsqr_amd.colmajor(torch.from_numpy(A_h).cuda())
    keep = A.clone()
    R_only = tsqr_amd.tsqr_r(A)
    torch.cuda.synchronize()
    assert torch.equal(A, keep)
    Q, R = tsqr_amd.tsqr(A)
    torch.cuda.synchronize()
    assert torch.equal(R_only, R)
    dev = tsqr_amd.check(keep, Q, R)
    host = O.metrics(A_h, Q.cpu().numpy(), R.cpu().numpy())
    _agree(dev["backward_fro"], host["backward"], "backward_fro")
    _agree(dev["orth_fro"], host["orth"], "orth_fro")
