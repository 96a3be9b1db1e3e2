"""The C host library built with ASan + UBSan (host code only; the kernel
library is linked as built) on the GPU: the host driver's dposv, and the two
testers at the reference's ctest sizes with their residual checks
(tests/Testings.cmake: potrf -N 378 -t 93 -x, potrf_dtd -N 874 -K 367 -t 76).
LeakSanitizer is off here: the HIP runtime keeps allocations until exit."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "dplasma_amd", "bin")


def run(args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")
    print("asan run:", " ".join(args), flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=120, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "Sanitizer" not in p.stderr and "runtime error" not in p.stderr, p.stderr
    return p.stdout


def test_host_driver_asan_gpu():
    assert "with GPU" in run([os.path.join(BIN, "host_driver_asan")])


@pytest.mark.parametrize("args", [["-N", "378", "-t", "93", "-x"], ["-N", "1700", "-t", "320", "-x", "-u", "U"]])
def test_testing_dpotrf_asan(args):
    run([os.path.join(BIN, "testing_dpotrf_asan")] + args)


def test_testing_dpotrf_dtd_asan():
    run([os.path.join(BIN, "testing_dpotrf_dtd_asan"), "-N", "874", "-K", "367", "-t", "76", "-x"])
