"""The drop-in path end to end: tests/dropin/ref_harness, a caller written
like the reference's harnesses (test_tsqr.cu:16-108, test_qr.cu:16-108)
against include/tsqr_compat.hpp and linked with -ltsqr_amd, is run on the
GPU and must print the reference's report lines (test_tsqr.cu:90,
utils.h:380-381) with errors within 8 n u and the reference's elementwise
all_close(QR, A) (utils.h:352-357).

The harness is built by `python -m tsqr_amd.build` (or __graft_entry__.build)
next to its source; it is built here only if it is missing.
"""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "tests", "dropin", "ref_harness")


def harness():
    if not os.path.exists(HARNESS):
        from tsqr_amd import build

        build.build_harness()
    return HARNESS


# (m, n, dtype code, kind, warmup, repeat): the reference's own defaults
# (13824 x 32 double), the TSQR configuration 27648 x 32 float, the recursive
# QR configuration 27648 x 128 in both precisions, and a height whose
# workspace exceeds the reference-sized ldwork = 2 NUM_SM BLOCK_SIZE rows, so
# that the library's internal buffer is used.
CASES = [
    (13824, 32, 2, "tsqr", 20, 50),
    (27648, 32, 1, "tsqr", 20, 50),
    (27648, 128, 1, "qr", 2, 5),
    (27648, 128, 2, "qr", 2, 5),
    (262144, 32, 1, "tsqr", 2, 5),
]


@pytest.mark.parametrize("m,n,dtype,kind,warmup,repeat", CASES)
def test_reference_style_caller(m, n, dtype, kind, warmup, repeat):
    exe = harness()
    print(f"ref_harness {m} {n} {dtype} {kind}", flush=True)
    r = subprocess.run([exe, str(m), str(n), str(dtype), kind, str(warmup), str(repeat)], capture_output=True,
                       text=True, timeout=100)
    print(r.stdout + r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    lat = re.search(rf"^{kind} Latency: ([0-9.]+) ms$", r.stdout, re.M)
    assert lat and float(lat.group(1)) > 0.0
    acc = re.search(r"^\|A-QR\|/\|A\| = ([0-9.eE+-]+), \|I-Q\^TQ\| / n = ([0-9.eE+-]+)$", r.stdout, re.M)
    assert acc
    u = 5.960464477539063e-08 if dtype == 1 else 1.1102230246251565e-16
    assert float(acc.group(1)) <= 8 * n * u and float(acc.group(2)) <= 8 * u
    assert "all_close = 1, R lower zero = 1" in r.stdout
    assert "status: 0" in r.stdout

