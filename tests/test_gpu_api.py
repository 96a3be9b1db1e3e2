"""GPU parity of the drop-in entry points (dplasma.h) against the oracle,
mirroring the reference testers (tests/testing_zpotrf.c, testing_zposv.c):
dplgsy -> dpotrf -> check_dpotrf, dplrnt -> dpotrs -> check_daxmb, on the
sizes the reference's ctest registers (tests/Testings.cmake:98-99,164-167,
220-223), ragged tiles included, both uplo, info on non-SPD input.
Tolerances: FP64 factors agree with the oracle's tile-DAG factor to a
relative 1e-12; residual ratios are below the reference threshold 60.
"""
import os
import sys
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(378, 93), (1700, 320), (3200, 320), (4096, 256), (2048, 512), (100, 128), (64, 64)]


def oracle_factor(po, N, NB, seed=3872):
    T = po.dplgsy_symtiles(N, NB, float(N), seed)
    po.load_openblas()
    assert po.dpotrf_tiles(T, N, NB, nthreads=8, use_blas=po.lib().orc_blas_loaded()) == 0
    return T


@pytest.mark.parametrize("N,NB", SIZES)
def test_dpotrf_lower_matches_oracle(ctx, dp, po, N, NB):
    A = dp.SymMatrix(ctx, N, NB, dp.LOWER)
    assert ctx.dplgsy(float(N), dp.LOWER, A, 3872) == 0
    gen = A.array.copy()
    assert np.array_equal(gen, po.dplgsy_symtiles(N, NB, float(N), 3872))
    assert ctx.dpotrf(dp.LOWER, A) == 0
    ref = oracle_factor(po, N, NB)
    Lg, Lo = po.symtiles_to_full(A.array.copy(), N, NB), po.symtiles_to_full(ref, N, NB)
    assert np.abs(np.tril(Lg) - np.tril(Lo)).max() <= 1e-12 * np.abs(np.tril(Lo)).max()
    # outside the factor nothing moved (upper part of diagonal tiles, padding)
    MT = (N + NB - 1) // NB
    for k in range(MT):
        i = k * MT - k * (k - 1) // 2
        g = A.array[i * NB * NB:(i + 1) * NB * NB].reshape((NB, NB), order="F")
        o = gen[i * NB * NB:(i + 1) * NB * NB].reshape((NB, NB), order="F")
        assert np.array_equal(np.triu(g, 1), np.triu(o, 1))
    if N <= 2048:
        A0 = po.dplgsy_full(N, float(N), 3872)
        assert po.check_dpotrf(po.LOWER, Lg, A0) < 60
    # the library's own GPU check agrees
    A0d = dp.SymMatrix(ctx, N, NB, dp.LOWER)
    ctx.dplgsy(float(N), dp.LOWER, A0d, 3872)
    assert ctx.check_dpotrf(0, dp.LOWER, A, A0d) == 0
    A.free()
    A0d.free()


@pytest.mark.parametrize("N,NB", [(378, 93), (1000, 256)])
def test_dpotrf_upper(ctx, dp, po, N, NB):
    A = dp.SymMatrix(ctx, N, NB, dp.UPPER)
    assert ctx.dplgsy(float(N), dp.UPPER, A, 3872) == 0
    assert np.array_equal(A.array, po.dplgsy_symtiles(N, NB, float(N), 3872, uplo=po.UPPER))
    assert ctx.dpotrf(dp.UPPER, A) == 0
    U = po.symtiles_to_full(A.array.copy(), N, NB, uplo=po.UPPER)
    A0 = po.dplgsy_full(N, float(N), 3872)
    assert po.check_dpotrf(po.UPPER, U, A0) < 60
    Lo = po.symtiles_to_full(oracle_factor(po, N, NB), N, NB)
    assert np.abs(np.triu(U) - np.tril(Lo).T).max() <= 1e-12 * np.abs(Lo).max()
    A.free()


@pytest.mark.parametrize("N,NB,r", [(378, 93, 0), (378, 93, 150), (1024, 256, 1023), (1024, 256, 511)])
def test_dpotrf_info(ctx, dp, po, N, NB, r):
    A0 = po.dplgsy_full(N, float(N), 3872)
    A0[r, r] = -2.0 * N
    A = dp.SymMatrix(ctx, N, NB, dp.LOWER)
    A.array[:] = po.full_to_symtiles(A0, NB)
    assert ctx.dpotrf(dp.LOWER, A) == r + 1
    T = po.full_to_symtiles(A0, NB)
    assert po.dpotrf_tiles(T, N, NB, nthreads=4) == r + 1
    A.free()


def test_bad_uplo_behaves_as_reference(ctx, dp):
    A = dp.SymMatrix(ctx, 64, 32, dp.LOWER)
    # dplasma_zpotrf returns 0 when _New rejects uplo (zpotrf_wrapper.c:207-210,351)
    assert ctx.dpotrf(dp.UPPERLOWER, A) == 0
    B = dp.GeMatrix(ctx, 64, 1, 32, 32)
    assert ctx.dpotrs(dp.UPPERLOWER, A, B) == -1
    assert ctx.dposv(999, A, B) == -1
    assert ctx.dtrsm(999, dp.LOWER, dp.NOTRANS, dp.NONUNIT, 1.0, A, B) == -1
    assert ctx.dtrsm(dp.LEFT, 5, dp.NOTRANS, dp.NONUNIT, 1.0, A, B) == -2
    assert ctx.dtrsm(dp.LEFT, dp.LOWER, 5, dp.NONUNIT, 1.0, A, B) == -3
    assert ctx.dtrsm(dp.LEFT, dp.LOWER, dp.NOTRANS, 5, 1.0, A, B) == -4
    C = dp.GeMatrix(ctx, 65, 1, 32, 32)
    assert ctx.dtrsm(dp.LEFT, dp.LOWER, dp.NOTRANS, dp.NONUNIT, 1.0, A, C) == -6
    assert ctx.dplgsy(1.0, 5, A, 1) == -3
    A.free(); B.free(); C.free()


@pytest.mark.parametrize("N,NB,nrhs", [(378, 93, 1), (378, 93, 367), (1700, 320, 16), (4096, 256, 1),
                                       (2048, 512, 16)])
def test_dpotrs_matches_oracle(ctx, dp, po, N, NB, nrhs):
    A = dp.SymMatrix(ctx, N, NB, dp.LOWER)
    ctx.dplgsy(float(N), dp.LOWER, A, 3872)
    assert ctx.dpotrf(dp.LOWER, A) == 0
    B = dp.GeMatrix(ctx, N, nrhs, NB, NB)
    assert ctx.dplrnt(0, B, 3873) == 0
    b = B.to_numpy()
    assert np.array_equal(b, po.dplrnt_full(N, nrhs, 3873))
    assert ctx.dpotrs(dp.LOWER, A, B) == 0
    x = B.to_numpy()
    # oracle solve with the GPU factor (isolates the solve kernels)
    Bt = po.full_to_tiles(b, NB, NB)
    po.lib().orc_tile_dpotrs_L(N, NB, A.array.copy(), nrhs, Bt)
    xo = po.tiles_to_full(Bt, N, nrhs, NB, NB)
    assert np.abs(x - xo).max() <= 1e-12 * np.abs(xo).max()
    A0 = dp.SymMatrix(ctx, N, NB, dp.LOWER)
    ctx.dplgsy(float(N), dp.LOWER, A0, 3872)
    Bc = dp.GeMatrix(ctx, N, nrhs, NB, NB)
    ctx.dplrnt(0, Bc, 3873)
    assert ctx.check_daxmb(0, dp.LOWER, A0, Bc, B) == 0
    if N <= 2048:
        assert po.check_daxmb(po.LOWER, po.dplgsy_full(N, float(N), 3872), b, x) < 60
    for M in (A, B, A0, Bc):
        M.free()


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_dposv_and_trsm_scenarios(ctx, dp, po, uplo):
    """testing_zposv.c scenarios 1-3 (posv, potrf+potrs, potrf+2 trsm)."""
    N, NB, nrhs = 378, 93, 7
    u = dp.LOWER if uplo == "L" else dp.UPPER
    A0full = po.dplgsy_full(N, float(N), 3872)
    b = po.dplrnt_full(N, nrhs, 2354)
    xs = []
    # 1: posv
    A = dp.SymMatrix(ctx, N, NB, u); ctx.dplgsy(float(N), u, A, 3872)
    X = dp.GeMatrix(ctx, N, nrhs, NB, NB); ctx.dplrnt(0, X, 2354)
    assert ctx.dposv(u, A, X) == 0
    xs.append(X.to_numpy()); A.free(); X.free()
    # 2: potrf + potrs
    A = dp.SymMatrix(ctx, N, NB, u); ctx.dplgsy(float(N), u, A, 3872)
    X = dp.GeMatrix(ctx, N, nrhs, NB, NB); ctx.dplrnt(0, X, 2354)
    assert ctx.dpotrf(u, A) == 0 and ctx.dpotrs(u, A, X) == 0
    xs.append(X.to_numpy())
    # 3: potrf + two trsm on the factor from 2
    X2 = dp.GeMatrix(ctx, N, nrhs, NB, NB); ctx.dplrnt(0, X2, 2354)
    if u == dp.LOWER:
        assert ctx.dtrsm(dp.LEFT, u, dp.NOTRANS, dp.NONUNIT, 1.0, A, X2) == 0
        assert ctx.dtrsm(dp.LEFT, u, dp.CONJTRANS, dp.NONUNIT, 1.0, A, X2) == 0
    else:
        assert ctx.dtrsm(dp.LEFT, u, dp.CONJTRANS, dp.NONUNIT, 1.0, A, X2) == 0
        assert ctx.dtrsm(dp.LEFT, u, dp.NOTRANS, dp.NONUNIT, 1.0, A, X2) == 0
    xs.append(X2.to_numpy())
    A.free(); X.free(); X2.free()
    xe = np.linalg.solve(A0full, b)
    for x in xs:
        assert po.check_daxmb(po.LOWER, A0full, b, x) < 60
        assert np.abs(x - xe).max() < 1e-12


def test_trsm_alpha_unit_lapack_storage(ctx, dp, po):
    N, NB, nrhs = 300, 64, 5
    A = dp.SymMatrix(ctx, N, NB, dp.LOWER)
    ctx.dplgsy(float(N), dp.LOWER, A, 11)
    Lfull = np.tril(po.symtiles_to_full(A.array.copy(), N, NB))
    rng = np.random.default_rng(0)
    b = rng.standard_normal((N, nrhs))
    for trans in (dp.NOTRANS, dp.TRANS):
        for diag in (dp.NONUNIT, dp.UNIT):
            X = dp.GeMatrix(ctx, N, nrhs, NB, NB, storage=dp.MATRIX_LAPACK)
            X.from_numpy(b)
            assert ctx.dtrsm(dp.LEFT, dp.LOWER, trans, diag, -2.5, A, X) == 0
            Le = Lfull.copy()
            if diag == dp.UNIT:
                np.fill_diagonal(Le, 1.0)
            op = Le if trans == dp.NOTRANS else Le.T
            exp = np.linalg.solve(op, -2.5 * b)
            assert np.abs(X.to_numpy() - exp).max() <= 1e-12 * np.abs(exp).max()
            X.free()
    A.free()


def test_resident_path_and_rec(ctx, dp, po):
    N, NB = 1024, 256
    A = dp.SymMatrix(ctx, N, NB, dp.LOWER)
    assert ctx.advise(dp.LOWER, A) == 0
    assert ctx.dplgsy(float(N), dp.LOWER, A, 3872) == 0      # generated in HBM
    assert ctx.dpotrf_rec(dp.LOWER, A, 64) == 0              # stays resident
    assert ctx.sync_host(A) == 0
    Lo = po.symtiles_to_full(oracle_factor(po, N, NB), N, NB)
    Lg = po.symtiles_to_full(A.array.copy(), N, NB)
    assert np.abs(np.tril(Lg) - np.tril(Lo)).max() <= 1e-12 * np.abs(Lo).max()
    t = ctx.last_timing()
    assert t["total_ms"] > 0 and t["stage_in_ms"] == 0
    ctx.release_all()
    A.free()


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("prec", [0, 1])
def test_dposv_mixed(ctx, dp, po, uplo, prec):
    """dplasma_dposv_mixed (dsposv semantics) on the dplgsy matrix: refined
    in a few steps, A and B untouched, X equal to the FP64 oracle solution
    to FP64 accuracy, and the reference's check_daxmb < 60."""
    N, NB, nrhs = 1700, 320, 16
    u = dp.LOWER if uplo == "L" else dp.UPPER
    A = dp.SymMatrix(ctx, N, NB, u); ctx.dplgsy(float(N), u, A, 3872)
    B = dp.GeMatrix(ctx, N, nrhs, NB, NB); ctx.dplrnt(0, B, 2354)
    X = dp.GeMatrix(ctx, N, nrhs, NB, NB)
    a0 = A.array.copy()
    b = B.to_numpy().copy()
    info, it = ctx.dposv_mixed(u, A, B, X, prec)
    assert info == 0 and 0 <= it <= 5, (info, it)
    assert np.array_equal(A.array, a0)
    assert np.array_equal(B.to_numpy(), b)
    x = X.to_numpy()
    A0 = po.dplgsy_full(N, float(N), 3872)
    xo = np.linalg.solve(A0, b)
    assert np.abs(x - xo).max() <= 1e-13 * np.abs(xo).max()
    assert po.check_daxmb(po.LOWER, A0, b, x) < 60
    A.free(); B.free(); X.free()


def test_dposv_mixed_not_spd(ctx, dp, po):
    """A non-SPD matrix: the FP32 factorization fails (iter = -3), the FP64
    fallback factorization reports info > 0 as dplasma_dpotrf does."""
    N, NB = 378, 93
    A = dp.SymMatrix(ctx, N, NB, dp.LOWER); ctx.dplgsy(0.0, dp.LOWER, A, 3872)
    B = dp.GeMatrix(ctx, N, 1, NB, NB); ctx.dplrnt(0, B, 2354)
    X = dp.GeMatrix(ctx, N, 1, NB, NB)
    info, it = ctx.dposv_mixed(dp.LOWER, A, B, X)
    assert info > 0 and it == -3, (info, it)
    A.free(); B.free(); X.free()


@pytest.mark.parametrize("uplo", ["L"])
def test_dpotrf_large_resident_gpu_checks(ctx, dp, uplo):
    """A size whose early panels take the 64-row TRSM kernel and whose late
    panels take the K-split one (>= 32 panel tiles at step 0): the
    reference's check_dpotrf and check_daxmb, computed on the GPU."""
    N, NB = 16896, 512
    u = dp.LOWER
    A = dp.SymMatrix(ctx, N, NB, u)
    A0 = dp.SymMatrix(ctx, N, NB, u)
    for M in (A, A0):
        assert ctx.advise(u, M) == 0
        assert ctx.dplgsy(float(N), u, M, 3872) == 0
    assert ctx.dpotrf(u, A) == 0
    assert ctx.check_dpotrf(0, u, A, A0) == 0
    B = dp.GeMatrix(ctx, N, 3, NB, NB); ctx.dplrnt(0, B, 3873)
    X = dp.GeMatrix(ctx, N, 3, NB, NB); ctx.dplrnt(0, X, 3873)
    assert ctx.dpotrs(u, A, X) == 0
    assert ctx.check_daxmb(0, u, A0, B, X) == 0
    for M in (A, A0, B, X):
        M.free()


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_dpotrf_adjoint_descriptors(ctx, dp, po, uplo):
    """dplasma_dpotrf_adjoint on descriptors: factor with dplasma_dpotrf, then
    the symmetric gradient of f = sum(W * L) against the oracle's closed form
    (Upper descriptors hold U = L^T and df/dU = W^T)."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    import adjoint as oad
    N, NB = 1000, 192
    u = dp.LOWER if uplo == "L" else dp.UPPER
    pu = po.LOWER if uplo == "L" else po.UPPER
    A = dp.SymMatrix(ctx, N, NB, u); ctx.dplgsy(float(N), u, A, 3872)
    assert ctx.dpotrf(u, A) == 0
    S = po.symtiles_to_full(A.array, N, NB, pu)
    L = np.tril(S) if uplo == "L" else np.triu(S).T
    W = np.tril(np.random.default_rng(4).standard_normal((N, N)))
    G = dp.SymMatrix(ctx, N, NB, u)
    G.array[:] = po.full_to_symtiles(W if uplo == "L" else W.T, NB, pu)
    assert ctx.dpotrf_adjoint(u, A, G) == 0
    R = po.symtiles_to_full(G.array, N, NB, pu)
    got = np.tril(R) if uplo == "L" else np.triu(R).T
    want = np.tril(oad.adjoint_closed(L, W))
    assert np.abs(got - want).max() <= 1e-12 * N * np.abs(want).max()
    A.free(); G.free()


def test_dpotrf_adjoint_bad_args(ctx, dp):
    A = dp.SymMatrix(ctx, 100, 32, dp.LOWER)
    G = dp.SymMatrix(ctx, 120, 32, dp.LOWER)
    assert ctx.dpotrf_adjoint(dp.UPPERLOWER, A, A) == -1
    assert ctx.dpotrf_adjoint(dp.LOWER, A, G) == -2
    A.free(); G.free()


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("N,NB", [(378, 93), (1000, 192), (1700, 320)])
def test_dpoinv(ctx, dp, po, uplo, N, NB):
    """dplasma_dpoinv (factor + inverse) on dplgsy: against numpy's inverse of
    the same matrix (relative 1e-12), the reference's check_dpoinv < 10 on the
    GPU and in the oracle, and dplasma_dpotri after dplasma_dpotrf giving the
    same inverse bit for bit."""
    u = dp.LOWER if uplo == "L" else dp.UPPER
    pu = po.LOWER if uplo == "L" else po.UPPER
    A = dp.SymMatrix(ctx, N, NB, u); ctx.dplgsy(float(N), u, A, 3872)
    A0 = dp.SymMatrix(ctx, N, NB, u); ctx.dplgsy(float(N), u, A0, 3872)
    assert ctx.dpoinv(u, A) == 0
    full = po.dplgsy_full(N, float(N), 3872)
    S = po.symtiles_to_full(A.array, N, NB, pu)
    got = np.tril(S) + np.tril(S, -1).T if uplo == "L" else np.triu(S) + np.triu(S, 1).T
    want = np.linalg.inv(full)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    assert ctx.check_dpoinv(0, u, A0, A) == 0
    assert po.check_dpoinv(pu, full, got) < 10
    B = dp.SymMatrix(ctx, N, NB, u); ctx.dplgsy(float(N), u, B, 3872)
    assert ctx.dpotrf(u, B) == 0
    assert ctx.dpotri(u, B) == 0
    assert np.array_equal(B.array, A.array)
    A.free(); A0.free(); B.free()


def test_dpoinv_args_and_singular(ctx, dp, po):
    A = dp.SymMatrix(ctx, 200, 64, dp.LOWER); ctx.dplgsy(200.0, dp.LOWER, A, 3872)
    assert ctx.dpotri(dp.UPPERLOWER, A) == -1
    assert ctx.dpoinv(123, A) == -1
    # a factor with an exact zero at L(70,70): dtrtri's info = 71, A unchanged
    assert ctx.dpotrf(dp.LOWER, A) == 0
    S = po.symtiles_to_full(A.array, 200, 64, po.LOWER)
    S[70, 70] = 0.0
    A.array[:] = po.full_to_symtiles(S, 64, po.LOWER)
    before = A.array.copy()
    assert ctx.dpotri(dp.LOWER, A) == 71
    assert np.array_equal(A.array, before)
    # not SPD: dplasma_dpoinv returns dplasma_dpotrf's info
    M = dp.SymMatrix(ctx, 200, 64, dp.LOWER); ctx.dplgsy(0.0, dp.LOWER, M, 3872)
    assert ctx.dpoinv(dp.LOWER, M) > 0
    A.free(); M.free()
