"""Parity of the GPU-side checks (check_dpotrf, check_daxmb, check_dpoinv,
src/dplasma_zcheck.c:63-133,179-224,264-314) with the oracle's restatement
(oracle/check.c) and with exact arithmetic.

The checks are compared value by value, not only on pass/fail: the norms
behind each ratio come back through dplasma_mi_last_check.  Cases:
  * integer factors L0 (entries -2..2, diagonal 1..3): A0 = L0 L0^T is exact
    in FP64, so ||L L^T - A0|| is exactly 0 for the true factor;
  * a planted error of 2^-10 in one entry (factor, solution or inverse):
    every quantity the checks form is then still exact in FP64, so the GPU
    norms must equal the closed form to the last bit, and the check must
    report failure;
  * a transposed operand (the lower factor given as an Upper descriptor)
    must fail with the oracle's residual;
  * the dplgsy matrices of the reference testers: ||A0|| agrees with the
    oracle to 1e-13 relative, and both residual ratios are below 60 and
    within 2 of each other (they are rounding noise of two summation
    orders).
Sizes follow the reference's ctest cases (tests/Testings.cmake:98,164-166),
both uplo.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [(378, 93), (1700, 320), (4096, 256)]
DELTA = 2.0 ** -10
EPS = 2.0 ** -53


def int_factor(N, seed):
    rng = np.random.default_rng(seed)
    L = np.tril(rng.integers(-2, 3, size=(N, N))).astype(np.float64)
    np.fill_diagonal(L, rng.integers(1, 4, size=N).astype(np.float64))
    return L


def inf_norm(M):
    return np.abs(M).sum(axis=1).max()


def sym_desc(dp, po, ctx, F, NB, uplo):
    """Symmetric descriptor holding the uplo triangle of the full array F."""
    N = F.shape[0]
    S = dp.SymMatrix(ctx, N, NB, uplo)
    S.array[:] = po.full_to_symtiles(F, NB, uplo=uplo)
    return S


def stored(L, uplo, dp):
    return L if uplo == dp.LOWER else np.ascontiguousarray(L.T)


@pytest.mark.parametrize("N,NB", CASES)
@pytest.mark.parametrize("u", ["L", "U"])
def test_check_dpotrf_exact_and_planted(ctx, dp, po, N, NB, u):
    uplo = dp.LOWER if u == "L" else dp.UPPER
    L = int_factor(N, N)
    A0 = L @ L.T                                  # exact: integers below 2^53
    A0d = sym_desc(dp, po, ctx, A0, NB, uplo)
    # the true factor: residual exactly zero
    F = sym_desc(dp, po, ctx, stored(L, uplo, dp), NB, uplo)
    assert ctx.check_dpotrf(0, uplo, F, A0d) == 0
    c = ctx.last_check()
    assert c["Rnorm"] == 0.0 and c["result"] == 0.0
    assert c["Anorm"] == inf_norm(A0)
    F.free()
    # planted error: L(i, j) += 2^-10 in the stored triangle
    i, j = (2 * N) // 3, N // 5
    Lp = L.copy()
    Lp[i, j] += DELTA
    R = Lp @ Lp.T - A0                            # exact (at most 34 significant bits)
    F = sym_desc(dp, po, ctx, stored(Lp, uplo, dp), NB, uplo)
    assert ctx.check_dpotrf(0, uplo, F, A0d) == 1
    c = ctx.last_check()
    assert c["Rnorm"] == inf_norm(R)
    assert c["result"] == pytest.approx(inf_norm(R) / (inf_norm(A0) * N * EPS), rel=1e-15)
    if N <= 1700:
        ro = po.check_dpotrf(po.LOWER if u == "L" else po.UPPER, np.asfortranarray(stored(Lp, uplo, dp)), A0)
        assert c["result"] == pytest.approx(ro, rel=1e-14)
    F.free()
    A0d.free()


@pytest.mark.parametrize("N,NB", CASES[:2])
def test_check_dpotrf_transposed_operand_fails(ctx, dp, po, N, NB):
    """The lower factor passed as an Upper descriptor (only its diagonal is
    then stored): U^T U = diag(L)^2, far from A0, as the oracle finds."""
    L = int_factor(N, 7 * N)
    A0 = L @ L.T
    A0d = sym_desc(dp, po, ctx, A0, NB, dp.UPPER)
    F = sym_desc(dp, po, ctx, L, NB, dp.UPPER)
    assert ctx.check_dpotrf(0, dp.UPPER, F, A0d) == 1
    c = ctx.last_check()
    D = np.diag(np.diag(L))
    R = D.T @ D - A0
    assert c["Rnorm"] == inf_norm(np.triu(R) + np.triu(R, 1).T)
    ro = po.check_dpotrf(po.UPPER, np.asfortranarray(L), A0)
    assert ro > 60 and c["result"] == pytest.approx(ro, rel=1e-14)
    F.free()
    A0d.free()


@pytest.mark.parametrize("N,NB", CASES)
@pytest.mark.parametrize("u", ["L", "U"])
def test_check_dpotrf_on_dplgsy_factor(ctx, dp, po, N, NB, u):
    uplo = dp.LOWER if u == "L" else dp.UPPER
    A = dp.SymMatrix(ctx, N, NB, uplo)
    assert ctx.dplgsy(float(N), uplo, A, 3872) == 0
    A0d = dp.SymMatrix(ctx, N, NB, uplo)
    assert ctx.dplgsy(float(N), uplo, A0d, 3872) == 0
    assert ctx.dpotrf(uplo, A) == 0
    assert ctx.check_dpotrf(0, uplo, A, A0d) == 0
    c = ctx.last_check()
    A0 = po.dplgsy_full(N, float(N), 3872)
    assert c["Anorm"] == pytest.approx(inf_norm(A0), rel=1e-13)
    assert c["result"] < 60
    if N <= 1700:
        pu = po.LOWER if u == "L" else po.UPPER
        F = po.symtiles_to_full(A.array.copy(), N, NB, uplo=pu)
        ro = po.check_dpotrf(pu, F, A0)
        assert ro < 60 and abs(c["result"] - ro) < 2.0
    A.free()
    A0d.free()


@pytest.mark.parametrize("N,NB,nrhs", [(378, 93, 7), (1700, 320, 16), (4096, 256, 1)])
@pytest.mark.parametrize("u", ["L", "U"])
def test_check_daxmb_exact_and_planted(ctx, dp, po, N, NB, nrhs, u):
    uplo = dp.LOWER if u == "L" else dp.UPPER
    L = int_factor(N, N + 1)
    A0 = L @ L.T
    rng = np.random.default_rng(N)
    x = rng.integers(-3, 4, size=(N, nrhs)).astype(np.float64)
    b = A0 @ x                                    # exact
    A0d = sym_desc(dp, po, ctx, A0, NB, uplo)
    bd = dp.GeMatrix(ctx, N, nrhs, NB, NB)
    bd.from_numpy(b)
    xd = dp.GeMatrix(ctx, N, nrhs, NB, NB)
    xd.from_numpy(x)
    assert ctx.check_daxmb(0, uplo, A0d, bd, xd) == 0
    c = ctx.last_check()
    assert c["Rnorm"] == 0.0
    assert (c["Anorm"], c["Bnorm"], c["Xnorm"]) == (inf_norm(A0), inf_norm(b), inf_norm(x))
    # planted error in x: r = -2^-10 A0(:, i) in column j
    i, j = N // 3, nrhs - 1
    xp = x.copy()
    xp[i, j] += DELTA
    xd.from_numpy(xp)
    assert ctx.check_daxmb(0, uplo, A0d, bd, xd) == 1
    c = ctx.last_check()
    assert c["Rnorm"] == DELTA * np.abs(A0[:, i]).max()
    exp = c["Rnorm"] / ((inf_norm(A0) * inf_norm(xp) + inf_norm(b)) * N * EPS)
    assert c["result"] == pytest.approx(exp, rel=1e-15)
    if N <= 1700:
        ro = po.check_daxmb(po.LOWER, A0, b, xp)
        assert c["result"] == pytest.approx(ro, rel=1e-14)
    for M in (A0d, bd, xd):
        M.free()


@pytest.mark.parametrize("N,NB", CASES[:2])
@pytest.mark.parametrize("u", ["L", "U"])
def test_check_dpoinv_exact_and_planted(ctx, dp, po, N, NB, u):
    """A = L0 L0^T with L0 = I - (subdiagonal ones): A^-1(i, j) = N - max(i, j),
    exact integers; then a planted 2^-10 in A^-1(i, j) and A^-1(j, i)."""
    uplo = dp.LOWER if u == "L" else dp.UPPER
    L0 = np.eye(N) - np.eye(N, k=-1)
    A = L0 @ L0.T
    idx = np.arange(N)
    Ainv = (N - np.maximum.outer(idx, idx)).astype(np.float64)
    assert np.array_equal(Ainv @ A, np.eye(N))
    Ad = sym_desc(dp, po, ctx, A, NB, uplo)
    Id = sym_desc(dp, po, ctx, Ainv, NB, uplo)
    assert ctx.check_dpoinv(0, uplo, Ad, Id) == 0
    c = ctx.last_check()
    assert c["Rnorm"] == 0.0
    assert (c["Anorm"], c["Ainvnorm"]) == (inf_norm(A), inf_norm(Ainv))
    i, j = (3 * N) // 4, N // 4
    Ap = Ainv.copy()
    Ap[i, j] += DELTA
    Ap[j, i] += DELTA
    Id.free()
    Id = sym_desc(dp, po, ctx, Ap, NB, uplo)
    assert ctx.check_dpoinv(0, uplo, Ad, Id) == 1
    c = ctx.last_check()
    R = np.eye(N) - Ap @ A                        # exact
    assert c["Rnorm"] == np.abs(R).sum(axis=0).max()
    ro = po.check_dpoinv(po.LOWER if u == "L" else po.UPPER, A, Ap)
    assert c["result"] == pytest.approx(ro, rel=1e-14)
    Ad.free()
    Id.free()
