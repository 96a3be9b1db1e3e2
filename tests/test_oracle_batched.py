"""Pins the oracle's batched factor-and-solve (oracle/batched.py) and checks
the host-side helper of the C ABI (mi_batch_sizes, no GPU needed) against
it."""
import ctypes
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import batched as ob  # noqa: E402


def test_splitmix64_published_outputs():
    # the generator seeded with 0: first outputs as published with the algorithm
    assert ob.splitmix64(0) == 0xE220A8397B1DCDAF
    assert ob.splitmix64(ob.GOLDEN) == 0x6E789E6AA1B965F4


@pytest.mark.parametrize("n", [1, 2, 5, 17, 40])
def test_potrf_matches_numpy(n):
    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n))
    A = G.T @ G / n + np.eye(n)
    L, info = ob.potrf(A)
    assert info == 0
    Ln = np.linalg.cholesky(A)
    assert np.abs(L - Ln).max() <= 1e-13 * np.abs(Ln).max()
    b = rng.standard_normal(n)
    _, x, info = ob.posv(A, b)
    assert np.abs(x - np.linalg.solve(A, b)).max() <= 1e-12 * max(1.0, np.abs(x).max())


def test_potrf_exact_integer_factor():
    L0 = np.array([[2, 0, 0, 0], [1, 3, 0, 0], [-1, 2, 1, 0], [4, 0, -2, 5]], dtype=np.float64)
    L, info = ob.potrf(L0 @ L0.T)
    assert info == 0 and np.array_equal(L, L0)


@pytest.mark.parametrize("bad", [0, 7, 19])
def test_potrf_info(bad):
    n = 20
    A = np.eye(n) * 4.0 + 0.1
    A[bad, bad] = -1.0
    L, info = ob.potrf(A)
    assert info == bad + 1
    L2, x, info2 = ob.posv(A, np.ones(n))
    assert info2 == bad + 1 and x is None


def test_mi_batch_sizes_matches_oracle(dp):
    mi, _ = dp.libs()
    n = (ctypes.c_int * 4096)()
    assert mi.mi_batch_sizes(11, 4096, 32, 1024, n) == 0
    got = list(n)
    assert got == ob.batch_sizes(11, 4096, 32, 1024)
    assert min(got) >= 32 and max(got) <= 1024
    # uniform on [32, 1024]: E[n^3] ~ 2.77e8 (the configuration's flop estimate)
    assert abs(np.mean(np.array(got, dtype=np.float64) ** 3) / 2.77e8 - 1) < 0.06
