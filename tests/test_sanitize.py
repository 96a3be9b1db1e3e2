"""Sanitizer builds of the CPU-side code, run without a GPU:
- the oracle (generators, tile kernels, the pthread tile-DAG executor,
  solves, checks) under ASan + UBSan and under TSan, driven by
  tests/sanitize/oracle_driver.c at the reference's ctest sizes;
- the C host library (descriptors, tile addressing, constructors' argument
  errors, DTD task classes, parsec_init without a GPU) under ASan + UBSan,
  driven by tests/sanitize/host_driver.c.
A sanitizer report makes the program exit non-zero."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "dplasma_amd", "bin")


@pytest.fixture(scope="module")
def oracle_san(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("san"))
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize", f"SANDIR={d}"], check=True,
                   stdout=subprocess.DEVNULL)
    return d


def run(cmd, env=None):
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "Sanitizer" not in p.stderr and "runtime error" not in p.stderr, p.stderr
    return p.stdout


def test_oracle_asan_ubsan(oracle_san):
    assert "ok" in run([os.path.join(oracle_san, "oracle_asan"), "4"])


def test_oracle_dag_tsan(oracle_san):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    assert "ok" in run([os.path.join(oracle_san, "oracle_tsan"), "8"], env=env)


def test_host_library_asan_ubsan():
    subprocess.run(["make", "-C", os.path.join(ROOT, "dplasma_amd"), "sanitize"], check=True,
                   stdout=subprocess.DEVNULL)
    out = run([os.path.join(BIN, "host_driver_asan")])
    assert "host driver: ok" in out
