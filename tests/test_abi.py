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
    # the maximum-height configuration fits in a fraction of A
    big = lib.tsqr_workspace_bytes(23887872, 32, 1)
    assert 0 < big < 23887872 * 32 * 4 // 6
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
