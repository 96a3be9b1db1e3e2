"""GPU parity of the LAPACK-layout entries (mi_dpotrf_lapack, mi_dpotrs_lapack,
mi_dposv_lapack, dplasma_mi.h): a column-major device array with lda >= N,
the route the reference's ScaLAPACK-style pdpotrf_w takes through a
PARSEC_MATRIX_LAPACK descriptor (src/scalapack_wrappers/
dplasma_wrapper_pdpotrf.c:139-232).

Checks: the factor equals, bit for bit, the tile path's factor of the same
matrix (mi_dpotrf_tiles on the oracle's tile image); it agrees with LAPACK's
Cholesky (numpy) to a relative 1e-12; the unreferenced triangle and the rows
beyond N are not touched (they hold NaN, so a read would also show in the
factor); the solve's backward error is below 1e-14 and equals the tile
solve bit for bit; info on an indefinite matrix is the failing column + 1.
The host route (dplasma_dpotrf on a descriptor with LAPACK storage) is
checked against LAPACK's Cholesky with NaN in the other triangle, which the
host staging carries into the unread halves of the diagonal tiles.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

UPPER, LOWER = 121, 122
CASES = [(378, 93, 401), (1000, 256, 1000), (2048, 512, 2100), (100, 128, 100), (64, 64, 70), (700, 96, 733)]


def colmajor(A, lda, uplo):
    """lda x N column-major buffer: A's uplo triangle, NaN elsewhere."""
    N = A.shape[0]
    H = np.full((lda, N), np.nan, order="F")
    mask = np.tril(np.ones((N, N), bool)) if uplo == LOWER else np.triu(np.ones((N, N), bool))
    H[:N, :][mask] = A[mask]
    return H


def put(dev, H):
    return dev.upload(np.asfortranarray(H).ravel(order="F"))


def get(dev, p, shape):
    return dev.download(p, shape[0] * shape[1]).reshape(shape, order="F")


def tile_factor(dev, po, A, NB):
    """The tile path on the oracle's lower tile image of A."""
    N = A.shape[0]
    T = po.full_to_symtiles(A, NB)
    pT = dev.upload(T)
    info = ctypes.c_int()
    dev.check(dev.M.mi_dpotrf_tiles(dev.h, N, NB, pT, ctypes.byref(info)), "mi_dpotrf_tiles")
    assert info.value == 0
    return pT, np.tril(po.symtiles_to_full(dev.download(pT, T.size), N, NB))


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("N,NB,lda", CASES)
def test_dpotrf_lapack(dev, po, N, NB, lda, uplo):
    A = po.dplgsy_full(N, float(N), 3872)
    H = colmajor(A, lda, uplo)
    pA = put(dev, H)
    info = ctypes.c_int(-7)
    dev.check(dev.M.mi_dpotrf_lapack(dev.h, uplo, N, pA, lda, NB, ctypes.byref(info)), "mi_dpotrf_lapack")
    assert info.value == 0
    G = get(dev, pA, (lda, N))
    pT, L = tile_factor(dev, po, A, NB)
    F = np.tril(G[:N]) if uplo == LOWER else np.triu(G[:N]).T
    assert np.array_equal(F, L)
    Lnp = np.linalg.cholesky(A)
    assert np.abs(F - Lnp).max() <= 1e-12 * np.abs(Lnp).max()
    # the other triangle and the rows beyond N are untouched
    other = np.triu(np.ones((N, N), bool), 1) if uplo == LOWER else np.tril(np.ones((N, N), bool), -1)
    assert np.isnan(G[:N][other]).all() and np.isnan(G[N:]).all()
    dev.free(pA)
    dev.free(pT)


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("N,NB,lda,nrhs,ldb", [(378, 93, 401, 5, 390), (1000, 256, 1000, 16, 1000),
                                               (2048, 512, 2100, 1, 2048)])
def test_dposv_dpotrs_lapack(dev, po, N, NB, lda, nrhs, ldb, uplo):
    A = po.dplgsy_full(N, float(N), 3872)
    B = po.dplrnt_full(N, nrhs, 2354)
    Bh = np.full((ldb, nrhs), np.nan, order="F")
    Bh[:N] = B
    # dposv = dpotrf then dpotrs
    pA, pB = put(dev, colmajor(A, lda, uplo)), put(dev, Bh)
    info = ctypes.c_int(-7)
    dev.check(dev.M.mi_dposv_lapack(dev.h, uplo, N, nrhs, pA, lda, pB, ldb, NB, ctypes.byref(info)),
              "mi_dposv_lapack")
    assert info.value == 0
    X = get(dev, pB, (ldb, nrhs))
    assert np.isnan(X[N:]).all()
    X = X[:N]
    r = np.abs(B - A @ X).max() / (np.abs(A).sum(1).max() * np.abs(X).max() + np.abs(B).max())
    assert r < 1e-14, r
    # dpotrs on the factor dposv left in A gives the same X
    pB2 = put(dev, Bh)
    dev.check(dev.M.mi_dpotrs_lapack(dev.h, uplo, N, nrhs, pA, lda, pB2, ldb, NB), "mi_dpotrs_lapack")
    assert np.array_equal(get(dev, pB2, (ldb, nrhs))[:N], X)
    # and the tile solve on the tile factor of the same matrix gives it too
    pT, _ = tile_factor(dev, po, A, NB)
    pB3 = dev.upload(np.asfortranarray(B).ravel(order="F"))
    dev.check(dev.M.mi_dpotrs_tiles(dev.h, N, NB, pT, nrhs, pB3, N), "mi_dpotrs_tiles")
    assert np.array_equal(get(dev, pB3, (N, nrhs)), X)
    for p in (pA, pB, pB2, pB3, pT):
        dev.free(p)


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("N,NB,r", [(378, 93, 150), (1024, 256, 511), (300, 64, 0)])
def test_dpotrf_lapack_info(dev, po, N, NB, r, uplo):
    A = po.dplgsy_full(N, float(N), 3872)
    A[r, r] = -2.0 * N
    pA = put(dev, colmajor(A, N, uplo))
    info = ctypes.c_int(-7)
    dev.check(dev.M.mi_dpotrf_lapack(dev.h, uplo, N, pA, N, NB, ctypes.byref(info)), "mi_dpotrf_lapack")
    assert info.value == r + 1
    # dposv factors, reports the same info and leaves B alone
    pA2 = put(dev, colmajor(A, N, uplo))
    B = np.ones((N, 2), order="F")
    pB = put(dev, B)
    dev.check(dev.M.mi_dposv_lapack(dev.h, uplo, N, 2, pA2, N, pB, N, NB, ctypes.byref(info)), "mi_dposv_lapack")
    assert info.value == r + 1
    assert np.array_equal(get(dev, pB, (N, 2)), B)
    for p in (pA, pA2, pB):
        dev.free(p)


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
def test_lapack_staging_round_trip(dev, po, uplo):
    """Staging alone: to the tile image (zeros outside the triangle) and back."""
    N, NB, lda = 333, 64, 350
    A = po.dplgsy_full(N, float(N), 77)
    pA = put(dev, colmajor(A, lda, uplo))
    nt = po.n_sym_tiles(N, NB) * NB * NB
    pT = dev.alloc(nt * 8)
    dev.check(dev.M.mi_dsym_lapack_to_tiles(dev.h, uplo, N, pA, lda, NB, pT), "to_tiles")
    T = dev.download(pT, nt)
    ref = po.full_to_symtiles(np.tril(A), NB)
    assert np.array_equal(T, ref)
    H2 = np.full((lda, N), np.nan, order="F")
    pA2 = put(dev, H2)
    dev.check(dev.M.mi_dsym_tiles_to_lapack(dev.h, uplo, N, NB, pT, pA2, lda), "to_lapack")
    G = get(dev, pA2, (lda, N))
    assert np.array_equal(np.isnan(G), np.isnan(colmajor(A, lda, uplo)))
    mask = ~np.isnan(G)
    assert np.array_equal(G[mask], colmajor(A, lda, uplo)[mask])
    for p in (pA, pT, pA2):
        dev.free(p)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("N,NB", [(378, 93), (1000, 256)])
def test_dplasma_dpotrf_on_lapack_storage_descriptor(ctx, dp, po, N, NB, uplo):
    """dplasma_dpotrf on a block-cyclic descriptor with PARSEC_MATRIX_LAPACK
    storage over a column-major host array: the call pdpotrf_w makes
    (src/scalapack_wrappers/dplasma_wrapper_pdpotrf.c:180-229).  The factor
    matches LAPACK's to a relative 1e-12; the other triangle (NaN) is neither
    read nor written."""
    A0 = po.dplgsy_full(N, float(N), 3872)
    u = dp.LOWER if uplo == "L" else dp.UPPER
    keep = np.tril(np.ones((N, N), bool)) if uplo == "L" else np.triu(np.ones((N, N), bool))
    H = np.where(keep, A0, np.nan)
    A = dp.GeMatrix(ctx, N, N, NB, NB, storage=dp.MATRIX_LAPACK)
    A.from_numpy(H)
    assert ctx.dpotrf(u, A) == 0
    G = A.to_numpy()
    L = np.linalg.cholesky(A0)
    F = np.where(keep, G, 0.0)
    F = F if uplo == "L" else F.T
    assert np.abs(F - L).max() <= 1e-12 * np.abs(L).max()
    assert np.isnan(G[~keep]).all()
    # dplasma_dpotrs on the factor, right-hand sides in LAPACK storage too
    B0 = po.dplrnt_full(N, 3, 2354)
    B = dp.GeMatrix(ctx, N, 3, NB, NB, storage=dp.MATRIX_LAPACK)
    B.from_numpy(B0)
    assert ctx.dpotrs(u, A, B) == 0
    X = B.to_numpy()
    r = np.abs(B0 - A0 @ X).max() / (np.abs(A0).sum(1).max() * np.abs(X).max() + np.abs(B0).max())
    assert r < 1e-14, r
    A.free()
    B.free()
