"""The recursive diagonal factor (dplasma_dpotrf_rec / setrecursive, the
RECURSIVE body of src/zpotrf_L.jdf:168-192): every diagonal tile larger than
hmb is factored as a tile Cholesky of its own with hmb x hmb sub-tiles.
Against the oracle's tile-DAG factor of the same matrix (the order of the
operations inside a diagonal tile differs, so the bound is the relative 1e-12
of the other factor tests), on ragged tilings, an hmb that is not a multiple
of 32 (rounded down), hmb >= NB (no recursion), NB = 1024 (whose whole-tile
POTRF would not fit the LDS, so only the recursive factor runs it), the info
of a non-SPD matrix, and the descriptor path with the reference's residual
check, both uplo."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def factor_rec(dev, T, N, NB, hmb):
    d = dev.upload(T)
    info = ctypes.c_int(-1)
    try:
        dev.check(dev.M.mi_dpotrf_tiles_rec(dev.h, N, NB, d, hmb, ctypes.byref(info)), "dpotrf_rec")
        out = dev.download(d, T.shape)
    finally:
        dev.free(d)
    return out, info.value


@pytest.mark.parametrize("N,NB,hmb", [(4096, 512, 128), (1700, 320, 64), (2048, 512, 100), (378, 93, 32),
                                      (1000, 256, 32), (600, 512, 128), (2000, 256, 512), (3000, 1024, 512)])
def test_rec_matches_oracle(dev, po, N, NB, hmb):
    print(f"dpotrf_rec N={N} NB={NB} hmb={hmb}", flush=True)
    T = po.dplgsy_symtiles(N, NB, float(N), 3872)
    g, info = factor_rec(dev, T, N, NB, hmb)
    assert info == 0
    To = T.copy()
    assert po.dpotrf_tiles(To, N, NB, nthreads=4) == 0
    Lg = np.tril(po.symtiles_to_full(g, N, NB))
    Lo = np.tril(po.symtiles_to_full(To, N, NB))
    assert np.abs(Lg - Lo).max() <= 1e-12 * np.abs(Lo).max()
    # the strict upper part of the diagonal tiles is never written
    MT = (N + NB - 1) // NB
    for k in range(MT):
        idx = k * MT - k * (k - 1) // 2
        tg = g[idx * NB * NB:(idx + 1) * NB * NB].reshape(NB, NB)    # [col][row]
        t0 = T[idx * NB * NB:(idx + 1) * NB * NB].reshape(NB, NB)
        up = np.triu(np.ones((NB, NB), dtype=bool), 1).T              # row < col
        assert np.array_equal(tg[up], t0[up])


@pytest.mark.parametrize("N,NB", [(4096, 512), (3000, 1024), (2000, 256)])
def test_default_diagonal_factor(dev, po, N, NB):
    """mi_dpotrf_tiles factors the diagonal tiles recursively with sub-tiles
    of NB / 2 for NB >= 512 (a multiple of 64) and with one POTRF launch
    below: bitwise the factor of mi_dpotrf_tiles_rec with hmb = NB / 2, or
    with hmb >= NB (no recursion)."""
    T = po.dplgsy_symtiles(N, NB, float(N), 3872)
    d = dev.upload(T)
    info = ctypes.c_int(-1)
    try:
        dev.check(dev.M.mi_dpotrf_tiles(dev.h, N, NB, d, ctypes.byref(info)), "dpotrf")
        a = dev.download(d, T.shape)
    finally:
        dev.free(d)
    assert info.value == 0
    b, ib = factor_rec(dev, T, N, NB, NB // 2 if NB >= 512 else NB)
    assert ib == 0
    assert np.array_equal(a, b)


@pytest.mark.parametrize("r", [0, 700, 1999])
def test_rec_info(dev, po, r):
    N, NB = 2000, 256
    A0 = po.dplgsy_full(N, float(N), 3872)
    A0[r, r] = -2.0 * N
    T = po.full_to_symtiles(A0, NB)
    _, info = factor_rec(dev, T, N, NB, 64)
    assert info == r + 1


def test_rec_rejects_nonpositive_hmb(dev, dp):
    info = ctypes.c_int(-1)
    assert dev.M.mi_dpotrf_tiles_rec(dev.h, 512, 128, None, 0, ctypes.byref(info)) == -2
    assert "hmb" in dp.last_error()


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_dplasma_dpotrf_rec_descriptor(ctx, dp, po, uplo):
    N, NB, hmb = 1700, 320, 96
    u = dp.LOWER if uplo == "L" else dp.UPPER
    A = dp.SymMatrix(ctx, N, NB, u)
    A0 = dp.SymMatrix(ctx, N, NB, u)
    assert ctx.dplgsy(float(N), u, A, 3872) == 0
    assert ctx.dplgsy(float(N), u, A0, 3872) == 0
    assert ctx.dpotrf_rec(u, A, hmb) == 0
    assert ctx.check_dpotrf(0, u, A, A0) == 0
    A.free()
    A0.free()
