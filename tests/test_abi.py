"""CPU-side checks of the C ABI: the library loads, exports every symbol
include/tsqr_amd.h declares, and its pure host functions (workspace sizing,
status text) behave.  No kernel is launched here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tsqr_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([a-z_]+)\s*\(", text)) - {"if", "defined"})


@pytest.fixture(scope="module")
def lib():
    import tsqr_amd

    return tsqr_amd.lib()


def test_header_and_binding_agree():
    from tsqr_amd._lib import EXPORTS

    assert declared_functions() == sorted(EXPORTS)


def test_every_declared_symbol_is_exported(lib):
    for name in declared_functions():
        assert hasattr(lib, name), name
        assert ctypes.cast(getattr(lib, name), ctypes.c_void_p).value


def test_status_strings(lib):
    assert lib.tsqr_status_string(0) == b"ok"
    for code in (-1, -2, -3, -4, -5):
        assert lib.tsqr_status_string(code) != b"unknown status"


def test_workspace_sizing(lib):
    # invalid shapes
    assert lib.tsqr_workspace_bytes(10, 20, 1) == 0
    assert lib.tsqr_workspace_bytes(100, 33, 1) == 0
    assert lib.tsqr_workspace_bytes(100, 8, 3) == 0
    # single superleaf: no tree, only the per-superleaf flag and the FP64
    # side blocks (superleaf Gram, whole-matrix R^-1, R, residuals) and one
    # partial sum
    assert 0 < lib.tsqr_workspace_bytes(256, 32, 1) <= 65536
    # grows at most linearly in m (arity 8 for n=32 fp32); superleaves get
    # taller with m until they hold 8 tiles, so the count of superleaves (and
    # with it the workspace) can stay flat between two sizes
    a = lib.tsqr_workspace_bytes(1 << 20, 32, 1)
    b = lib.tsqr_workspace_bytes(1 << 21, 32, 1)
    assert 0 < a <= b <= 2 * a + (1 << 20)
    assert a < (1 << 20) * 32 * 4 // 2
    # the maximum-height configuration fits in a fraction of A: one 32 x 32 R
    # per 256-row tile in the level-1 blocks (1/8 of A) plus the FP64 side
    # blocks of the mixed Q pass's superleaves
    big = lib.tsqr_workspace_bytes(23887872, 32, 1)
    assert 0 < big < 23887872 * 32 * 4 // 5
    assert lib.qr_workspace_bytes(27648, 128, 2) >= lib.tsqr_workspace_bytes(27648, 32, 2)


def test_compat_header_compiles_reference_style_caller(tmp_path):
    """A caller written against the reference's templates (test_tsqr.cu:42-53
    style) compiles unchanged against include/tsqr_compat.hpp."""
    import shutil
    import subprocess

    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("no g++")
    src = tmp_path / "caller.cpp"
    src.write_text(
        '#include "tsqr_compat.hpp"\n'
        "template <typename T> void run(int m, int n, T* A, T* R, T* work) {\n"
        "    int ldwork = 2 * NUM_SM * BLOCK_SIZE;\n"
        "    tsqr<T>(m, n, A, m, R, n, work, ldwork);\n"
        "    qr<T>(m, 4 * n, A, m, R, 4 * n, work, ldwork);\n"
        "}\n"
        "template void run<float>(int, int, float*, float*, float*);\n"
        "template void run<double>(int, int, double*, double*, double*);\n")
    inc = os.path.join(ROOT, "include")
    r = subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-I", inc, "-c", str(src), "-o",
                        str(tmp_path / "caller.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_reference_ldwork_short_at_262144(lib):
    """The drop-in harness case 262144 x 32 float really takes the library's
    internal buffer: the reference-sized work array (2 NUM_SM BLOCK_SIZE rows
    of n elements, NUM_SM = BLOCK_SIZE = 256 in tsqr_compat.hpp) is smaller
    than the workspace."""
    m, n = 262144, 32
    assert lib.tsqr_workspace_bytes(m, n, 1) > 2 * 256 * 256 * n * 4


def panel_widths(n):
    """Widths of the TSQR panels the recursive qr factors: the column split
    n1 = 32 ceil(n / 64), n2 = n - n1 applied until n <= 32."""
    if n <= 32:
        return [n]
    n1 = 32 * ((n + 63) // 64)
    return panel_widths(n1) + panel_widths(n - n1)


def _qr_covers_panels(lib, ms, ns):
    bad = []
    for dtype in (1, 2):
        for m in ms:
            for n in ns:
                if n > m:
                    continue
                q = lib.qr_workspace_bytes(m, n, dtype)
                for w in set(panel_widths(n)):
                    t = lib.tsqr_workspace_bytes(m, w, dtype)
                    if q < t:
                        bad.append((dtype, m, n, w, q, t))
    return bad


MS = [1000, 4096, 27648, 100000, 262144, 300000, 1 << 20, 3000000]
NS = [33, 40, 48, 64, 72, 96, 100, 128, 200, 256]


def test_qr_workspace_covers_every_panel(lib):
    """qr_workspace_bytes holds the TSQR workspace of every panel width the
    recursion produces: narrow panels (a 16-wide last panel of n = 48) can
    need more than a 32-wide one when they get more superleaves."""
    assert not _qr_covers_panels(lib, MS, NS)


def test_qr_workspace_covers_every_panel_tuned(lib):
    """Same under the measured tuning table, which gives narrower panels
    more, shorter superleaves (262144 x 48 FP64 needed 16.4 MB for the panel
    region and had 19.3 MB worth of writes)."""
    from tsqr_amd.autotune import DEFAULT_FILE, apply, load, restore_default

    table = load(DEFAULT_FILE)
    try:
        apply(table)
        assert lib.qr_workspace_bytes(262144, 48, 2) >= lib.tsqr_workspace_bytes(262144, 16, 2)
        assert not _qr_covers_panels(lib, MS, NS)
        # narrowest trees (arity 2, deeper, more blocks) and the smallest
        # streaming grid
        for code in (1, 2):
            for w in (8, 16, 32):
                assert lib.tsqr_set_tree_arity(code, w, 2) == 0
        assert lib.tsqr_set_stream_shape(256, 1024) == 0
        assert not _qr_covers_panels(lib, MS, NS)
    finally:
        restore_default()
