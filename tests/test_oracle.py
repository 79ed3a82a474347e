"""Pins the CPU oracle (oracle/) before anything is checked against it.

Each check is chosen so that a plausible mistake in the restatement fails:
the generator against libstdc++'s own stream, the reflector against closed
forms, and the tile / tree / recursion against FP64 LAPACK and the QR
invariants.  A planted sign error in an oracle output must be caught.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "uniform_stream.json")


def fnv1a64(b: bytes) -> str:
    a = np.frombuffer(b, dtype=np.uint8)
    h = 1469598103934665603
    # chunked pure-Python FNV is slow; fold in numpy-friendly blocks
    for x in a.tolist():
        h ^= x
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return "%016x" % h


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.mark.parametrize(
    "name,dtype", [("float_2048x16", np.float32), ("double_1024x8", np.float64)]
)
def test_generator_matches_libstdcxx(golden, name, dtype):
    e = golden[name]
    a = O.gen_uniform_flat(e["count"], dtype)
    assert fnv1a64(a.tobytes()) == e["fnv1a64"]
    bits = [("%08x" if dtype == np.float32 else "%016x") % v for v in a[:16].view(
        np.uint32 if dtype == np.float32 else np.uint64)]
    assert bits == e["first16_hex"]


def test_generator_first_values_cfg1(golden):
    """First 16 values of the 27648 x 32 float stream (test_tsqr.cu:24-28) and
    independence of chunked generation (jump-ahead) from the start offset."""
    e = golden["float_27648x32"]
    a = O.gen_uniform_flat(16, np.float32)
    assert ["%08x" % v for v in a.view(np.uint32)] == e["first16_hex"]
    full = O.gen_uniform_flat(5000, np.float32)
    tail = O.gen_uniform_flat(1000, np.float32, first=4000)
    assert np.array_equal(full[4000:], tail)


def test_generator_full_hash_cfg1(golden):
    """The whole 27648 x 32 float stream and the 27648 x 32 double stream of
    the harness (test_tsqr.cu:24-28) hash to libstdc++'s."""
    for name, dtype in (("float_27648x32", np.float32), ("double_27648x32", np.float64)):
        e = golden[name]
        a = O.gen_uniform_flat(e["count"], dtype)
        assert O.fnv1a64(a) == e["fnv1a64"], name
    # the C hash agrees with the plain definition
    a = O.gen_uniform_flat(4096, np.float32)
    assert O.fnv1a64(a) == fnv1a64(a.tobytes())


NORMAL = os.path.join(os.path.dirname(__file__), "golden", "normal_stream.json")


@pytest.fixture(scope="module")
def normal_golden():
    with open(NORMAL) as f:
        return json.load(f)


def test_mt19937_64_matches_libstdcxx(normal_golden):
    assert ["%016x" % v for v in O.mt64_raw(8, 0)] == normal_golden["mt19937_64_first8_hex"]


def test_normal_stream_matches_fixture(normal_golden):
    """The recursive-QR configuration's 27648 x 128 N(0,1) input: first
    values and the hash of the whole FP64 and FP32 arrays (column-major)."""
    A = O.gen_normal_stream(27648, 128, np.float64)
    flat = A.ravel(order="F")
    assert ["%016x" % v for v in flat[:16].view(np.uint64)] == normal_golden["normal_first16_hex"]
    assert O.fnv1a64(flat) == normal_golden["double_27648x128"]["fnv1a64"]
    Af = O.gen_normal_stream(27648, 128, np.float32)
    assert O.fnv1a64(Af.ravel(order="F")) == normal_golden["float_27648x128"]["fnv1a64"]
    assert np.array_equal(Af, A.astype(np.float32))
    # N(0,1) to sampling accuracy
    assert abs(flat.mean()) < 5e-3 and abs(flat.std() - 1.0) < 5e-3


def test_batch_shapes_match_fixture(normal_golden):
    """The batched configuration's 512 (m, n) pairs (tsqr_amd/workloads.py)
    are libstdc++'s, and obey the stated distribution's support."""
    from tsqr_amd.workloads import batch_shapes

    shapes = batch_shapes()
    assert shapes == [tuple(x) for x in normal_golden["batch_512_shapes"]]
    assert all(n in (8, 16, 32, 64) and m % n == 0 and 4096 - n < m <= 262144 for m, n in shapes)


def test_reflector_kat_3_4():
    """x = (3, 4): R00 = -5, Q = -(0.6, 0.8) (u = (1.26491, 0.63246))."""
    Q, R = O.qr_tile_ref(np.array([[3.0], [4.0]]))
    assert R[0, 0] == pytest.approx(-5.0, abs=1e-15)
    np.testing.assert_allclose(Q[:, 0], [-0.6, -0.8], atol=1e-15)
    u0 = np.sqrt(1.0 - Q[0, 0])  # Q e0 = e0 - u u0
    assert u0 == pytest.approx(1.2649110640673518, abs=1e-14)
    assert -Q[1, 0] / u0 == pytest.approx(0.6324555320336759, abs=1e-14)


def test_diagonal_kat():
    """A = [diag(d); 0] with nonzero d: R = -diag(d), Q = -[I; 0]."""
    d = np.array([2.0, -3.0, 0.5, -7.0])
    A = np.vstack([np.diag(d), np.zeros((4, 4))])
    Q, R = O.qr_tile_ref(A)
    np.testing.assert_allclose(R, -np.diag(d), atol=1e-14)
    np.testing.assert_allclose(Q, -np.vstack([np.eye(4), np.zeros((4, 4))]), atol=1e-14)


def test_single_column_kat():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((37, 1))
    for sgn in (1.0, -1.0):
        x[0, 0] = sgn * abs(x[0, 0])
        Q, R = O.qr_tile_ref(x)
        nrm = np.linalg.norm(x)
        assert R[0, 0] == pytest.approx(-sgn * nrm, rel=1e-14)
        np.testing.assert_allclose(Q[:, 0], -sgn * x[:, 0] / nrm, atol=1e-14)


def test_zero_column_gives_identity_reflector():
    A = np.zeros((16, 3))
    A[:, 0] = np.arange(1, 17)
    A[:, 2] = np.linspace(-1, 1, 16)
    Q, R = O.qr_tile_ref(A)
    assert R[1, 1] == 0.0
    m = O.metrics(A, Q, R)
    assert m["backward"] < 1e-14


@pytest.mark.parametrize("h,n", [(256, 32), (128, 32), (64, 8), (40, 16)])
def test_tile_matches_lapack(h, n):
    """A single tile's R equals LAPACK's R including signs (same dlarfg rule)."""
    A = O.gen_normal(h, n, np.float64, seed=h + n)
    Q, R = O.qr_tile_ref(A)
    Qo, Ro = O.lapack_qr(A)
    np.testing.assert_allclose(R, Ro, atol=1e-12 * np.abs(Ro).max())
    np.testing.assert_allclose(Q, Qo, atol=1e-12)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 2e-5)])
def test_tsqr_cfg1_shape(dtype, tol):
    """Reference tsqr at the A800 float tuning (NUM_SM=108, BLOCK_SIZE=256),
    m=27648 n=32, the harness input: R vs LAPACK after sign normalisation,
    invariants, and the reference's own all_close criterion."""
    A = O.gen_uniform(27648, 32, dtype)
    Q, R = O.tsqr_ref(A, bs=256, num_sm=108)
    Qo, Ro = O.lapack_qr(A)
    m = O.metrics(A, Q, R, Ro, Qo)
    assert m["all_close"]
    assert m["backward"] < tol and m["orth"] < tol * 10
    assert m["r_err"] < tol and m["q_err"] < tol * 10
    assert m["r_lower_zero"]


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 2e-5)])
def test_tsqr_chunked_path(dtype, tol):
    """m above one chunk (NUM_SM*BS): chunk pass, TSQR of the stacked chunk
    Rs and the chunk rescale (TallShinnyQR.h:106-127)."""
    A = O.gen_uniform(2048, 16, dtype, seed=5)
    Q, R = O.tsqr_ref(A, bs=64, num_sm=4)  # 256-row chunks, 8 chunks
    Qo, Ro = O.lapack_qr(A)
    m = O.metrics(A, Q, R, Ro, Qo)
    assert m["all_close"] and m["backward"] < tol and m["orth"] < tol * 10 and m["r_err"] < tol


def test_tsqr_rejects_outside_domain():
    with pytest.raises(ValueError):
        O.tsqr_ref(O.gen_uniform(100, 32, np.float64), bs=256, num_sm=108)  # (m % BS) % n != 0
    with pytest.raises(ValueError):
        O.tsqr_ref(O.gen_uniform(64, 24, np.float64), bs=256, num_sm=108)  # BS % n != 0


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 5e-5)])
def test_recursive_qr(dtype, tol):
    """qr (block_QR.h:72-95), n = 128: R vs LAPACK, R12 = Q1^T A2, R21 = 0."""
    A = O.gen_normal(4096, 128, dtype, seed=11)
    Q, R = O.qr_ref(A, bs=256, num_sm=108)
    Qo, Ro = O.lapack_qr(A)
    m = O.metrics(A, Q, R, Ro, Qo)
    assert m["backward"] < tol and m["orth"] < tol * 10 and m["r_err"] < tol
    assert np.all(R[64:, :64] == 0)
    # the coupling block: R12 = Q1^T A2 with Q1 the first 64 columns of Q
    R12 = Q[:, :64].astype(np.float64).T @ A[:, 64:].astype(np.float64)
    np.testing.assert_allclose(R[:64, 64:], R12, atol=tol * 50 * np.abs(R12).max())


def test_planted_sign_error_is_caught():
    """A Q column with flipped sign but the R row left alone must fail both the
    backward-error check and the reference's all_close criterion."""
    A = O.gen_uniform(1024, 8, np.float64)
    Q, R = O.tsqr_ref(A, bs=128, num_sm=108)
    Qbad = Q.copy()
    Qbad[:, 3] *= -1
    m = O.metrics(A, Qbad, R)
    assert not m["all_close"] and m["backward"] > 1e-2
    Rbad = R.copy()
    Rbad[2, 5] += 1e-3  # a single wrong R entry
    assert O.metrics(A, Q, Rbad)["backward"] > 1e-6


def test_conditioned_generator():
    A = O.gen_conditioned(2000, 16, 1e4, np.float64, seed=2)
    s = np.linalg.svd(A, compute_uv=False)
    assert s[0] / s[-1] == pytest.approx(1e4, rel=1e-8)
