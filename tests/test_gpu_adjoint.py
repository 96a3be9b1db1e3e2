"""GPU parity of the blocked Cholesky adjoint (mi_dpotrf_adjoint_tiles /
mi_spotrf_adjoint_tiles) through the C ABI, against the oracle's closed form
(oracle/adjoint.py, itself pinned by finite differences) on the same L and
Lbar, and against the closed form d logdet(A)/dA = A^-1.

FP64 tolerance: max |Abar_gpu - Abar_oracle| <= 1e-12 * N * max |Abar| (the
oracle takes dense solves, the kernels tile products in another order).
FP32: 2e-4 relative to max |Abar| against the FP64 oracle on the FP32 inputs.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import adjoint as oad  # noqa: E402

pytestmark = pytest.mark.gpu


def pack_lower(A, NB, fill=np.nan, dtype=np.float64):
    """Dense -> packed lower tiles; strict upper parts of diagonal tiles and
    padding rows/columns hold `fill` (never read by the kernels)."""
    N = A.shape[0]
    MT = (N + NB - 1) // NB
    T = np.full(MT * (MT + 1) // 2 * NB * NB, fill, dtype=dtype)
    t = 0
    for n in range(MT):
        for m in range(n, MT):
            rm, cn = min(NB, N - m * NB), min(NB, N - n * NB)
            blk = np.array(A[m * NB:m * NB + rm, n * NB:n * NB + cn], dtype=dtype)
            if m == n:
                blk[np.triu_indices(rm, 1)] = fill
            tile = T[t * NB * NB:(t + 1) * NB * NB].reshape(NB, NB)   # [column][row]
            tile[:cn, :rm] = blk.T
            t += 1
    return T


def unpack_lower(T, N, NB):
    """Packed lower tiles -> dense lower triangle (float64), plus the values
    found in the never-written places (strict upper of diagonal tiles, padding)."""
    MT = (N + NB - 1) // NB
    A = np.zeros((N, N))
    untouched = []
    t = 0
    for n in range(MT):
        for m in range(n, MT):
            rm, cn = min(NB, N - m * NB), min(NB, N - n * NB)
            tile = T[t * NB * NB:(t + 1) * NB * NB].reshape(NB, NB).astype(np.float64)
            blk = tile[:cn, :rm].T.copy()
            if m == n:
                untouched.append(blk[np.triu_indices(rm, 1)])
                blk = np.tril(blk)
            A[m * NB:m * NB + rm, n * NB:n * NB + cn] = blk
            untouched.append(tile[cn:, :].ravel())
            untouched.append(tile[:cn, rm:].ravel())
            t += 1
    return A, np.concatenate(untouched) if untouched else np.zeros(0)


def spd(N, seed):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((N, N))
    return B @ B.T / N + 0.5 * np.eye(N)


def run_adjoint(dev, L, Lbar, NB, dtype=np.float64):
    N = L.shape[0]
    dl = dev.upload(pack_lower(L, NB, dtype=dtype))
    Tg = pack_lower(Lbar, NB, dtype=dtype)
    dg = dev.upload(Tg)
    fn = dev.M.mi_dpotrf_adjoint_tiles if dtype == np.float64 else dev.M.mi_spotrf_adjoint_tiles
    dev.check(fn(dev.h, N, NB, dl, dg), "adjoint")
    got = dev.download(dg, Tg.shape, dtype)
    dev.free(dl)
    dev.free(dg)
    return unpack_lower(got, N, NB)


@pytest.mark.parametrize("N,NB", [(64, 64), (100, 128), (300, 64), (378, 93), (1000, 128), (1700, 320),
                                  (2048, 512), (3000, 512)])
def test_adjoint_vs_oracle(dev, N, NB):
    A = spd(N, N + NB)
    L = np.linalg.cholesky(A)
    rng = np.random.default_rng(7 + N)
    Lbar = np.tril(rng.standard_normal((N, N)))
    ref = oad.adjoint_closed(L, Lbar)
    got, untouched = run_adjoint(dev, L, Lbar, NB)
    err = np.abs(got - np.tril(ref)).max()
    scale = np.abs(ref).max()
    print(f"N={N} NB={NB} max err {err:.3e} (scale {scale:.3e})")
    assert err <= 1e-12 * N * scale, (err, scale)
    assert np.isnan(untouched).all()          # strict upper of diagonal tiles and padding never written


@pytest.mark.parametrize("N,NB", [(500, 128), (1024, 256)])
def test_adjoint_logdet_closed_form(dev, N, NB):
    """f = logdet A = 2 sum log L_ii: Lbar = diag(2 / L_ii), Abar = A^-1."""
    A = spd(N, 3)
    L = np.linalg.cholesky(A)
    got, _ = run_adjoint(dev, L, np.diag(2.0 / np.diag(L)), NB)
    Ai = np.linalg.inv(A)
    assert np.abs(got - np.tril(Ai)).max() <= 1e-12 * N * np.abs(Ai).max()


@pytest.mark.parametrize("N,NB", [(378, 93), (1000, 256), (2048, 512)])
def test_sadjoint_vs_oracle(dev, N, NB):
    A = spd(N, 11)
    L = np.linalg.cholesky(A).astype(np.float32)
    rng = np.random.default_rng(5)
    Lbar = np.tril(rng.standard_normal((N, N))).astype(np.float32)
    ref = oad.adjoint_closed(L.astype(np.float64), Lbar.astype(np.float64))
    got, untouched = run_adjoint(dev, L, Lbar, NB, np.float32)
    err = np.abs(got - np.tril(ref)).max()
    assert err <= 2e-4 * np.abs(ref).max(), err
    assert np.isnan(untouched).all()


def test_adjoint_repeat_deterministic(dev):
    """Two calls on the same input give bit-identical results (split-K
    partials are reduced in a fixed order)."""
    N, NB = 1500, 256
    A = spd(N, 2)
    L = np.linalg.cholesky(A)
    Lbar = np.tril(np.random.default_rng(1).standard_normal((N, N)))
    a, _ = run_adjoint(dev, L, Lbar, NB)
    b, _ = run_adjoint(dev, L, Lbar, NB)
    assert np.array_equal(a, b)
