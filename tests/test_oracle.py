"""Pins the oracle (CPU restatement of DPLASMA's Cholesky path) before it is
trusted as the checker of the GPU kernels.  Each test names what pins it:
the reference's own header code compiled from its sources (oracle/_ref), the
independent fixtures of tests/golden/, closed forms, brute force on tiny
inputs or special matrices whose factor is known exactly.
"""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "generator_fixtures.json")
A_LCG, MASK = 6364136223846793005, (1 << 64) - 1


def test_rnd64_jump_matches_brute_force(po):
    L = po.lib()
    for seed in (0, 1, 3872, 3873, 2354, 2**63 + 5):
        x = seed
        for t in range(300):
            assert L.orc_rnd64_jump(t, seed) == x, (seed, t)
            x = (A_LCG * x + 1) & MASK


def test_rnd64_jump_matches_reference_header(po):
    ref = po.ref()
    if ref is None:
        pytest.skip("reference sources absent: oracle/_ref not built")
    L = po.lib()
    rng = np.random.default_rng(0)
    for _ in range(2000):
        n = int(rng.integers(0, 2**62))
        seed = int(rng.integers(0, 2**63))
        assert L.orc_rnd64_jump(n, seed) == ref.ref_rnd64_jump(n, seed)


def test_flops_match_reference_header(po):
    ref = po.ref()
    if ref is None:
        pytest.skip("reference sources absent: oracle/_ref not built")
    L = po.lib()
    for n in (1, 93, 378, 4096, 32768):
        assert L.orc_flops_dpotrf(n) == ref.ref_flops_dpotrf(n)
        for nrhs in (1, 16):
            assert L.orc_flops_dpotrs(n, nrhs) == ref.ref_flops_dpotrs(n, nrhs)
        assert L.orc_flops_dgemm(n, 7, 5) == ref.ref_flops_dgemm(n, 7, 5)
        assert L.orc_flops_dsyrk(n, 9) == ref.ref_flops_dsyrk(n, 9)
        for side in (141, 142):
            assert L.orc_flops_dtrsm(side, n, 11) == ref.ref_flops_dtrsm(side, n, 11)


def test_flops_closed_forms(po):
    L = po.lib()
    # the counts quoted for the benchmark configurations
    assert L.orc_flops_dpotrf(4096) == 22914881536.0
    assert L.orc_flops_dpotrf(32768) == 11728660905984.0
    assert L.orc_flops_dpotrs(32768, 1) == 2 * 32768**2
    # sum of the tile task counts equals the LAWN-41 total for NB | N
    N, NB = 1024, 128
    MT = N // NB
    tot = MT * L.orc_flops_dpotrf(NB)
    tot += MT * (MT - 1) // 2 * (L.orc_flops_dtrsm(142, NB, NB) + L.orc_flops_dsyrk(NB, NB))
    tot += sum((MT - k - 1) * (MT - k - 2) // 2 for k in range(MT)) * L.orc_flops_dgemm(NB, NB, NB)
    assert abs(tot - L.orc_flops_dpotrf(N)) / L.orc_flops_dpotrf(N) < 2e-3


def _gold():
    with open(GOLD) as f:
        return json.load(f)


def test_generators_match_golden_fixtures(po):
    for case in _gold()["cases"]:
        if case["kind"] == "dplgsy_full":
            N = case["N"]
            A = po.dplgsy_full(N, case["bump"], case["seed"])
            exp = np.array([[float.fromhex(v) for v in row] for row in case["values"]])
            assert np.array_equal(A, exp), case["seed"]
        elif case["kind"] == "dplgsy_entries":
            N, NB = case["N"], (93 if case["N"] == 378 else 256)
            T = po.dplgsy_symtiles(N, NB, case["bump"], case["seed"])
            A = po.symtiles_to_full(T, N, NB)
            for i, j, v in case["entries"]:
                lo, hi = (i, j) if i >= j else (j, i)
                assert A[lo, hi] == float.fromhex(v), (N, i, j)
        elif case["kind"] == "dplrnt":
            M, nr = case["M"], case["N"]
            B = po.dplrnt_full(M, nr, case["seed"])
            exp = np.array([[float.fromhex(v) for v in row] for row in case["values"]])
            assert np.array_equal(B, exp)


def test_generator_invariants(po):
    N, NB, seed = 378, 93, 3872
    T = po.dplgsy_symtiles(N, NB, float(N), seed, uplo=po.LOWER)
    A = po.symtiles_to_full(T, N, NB)
    A = np.tril(A) + np.tril(A, -1).T
    off = A - np.diag(np.diag(A))
    assert np.all(np.abs(off) <= 0.5)
    d = np.diag(A)
    assert np.all((d > N - 0.5) & (d <= N + 0.5))
    # off-diagonal values are binary32 numbers
    assert np.array_equal(off.astype(np.float32).astype(np.float64), off)
    # the upper descriptor stores the mirrored tiles
    TU = po.dplgsy_symtiles(N, NB, float(N), seed, uplo=po.UPPER)
    AU = po.symtiles_to_full(TU, N, NB, uplo=po.UPPER)
    assert np.array_equal(np.triu(AU), np.triu(A))
    # tile generation equals whole-matrix generation
    Af = po.dplgsy_full(N, float(N), seed)
    assert np.array_equal(Af, A)


def test_potrf_kernel_exact_factor(po):
    rng = np.random.default_rng(1)
    for n in (1, 2, 5, 17, 40):
        L0 = np.tril(rng.integers(-3, 4, size=(n, n))).astype(np.float64)
        np.fill_diagonal(L0, rng.integers(1, 5, size=n))
        A = np.asfortranarray(L0 @ L0.T)
        info = po.lib().orc_dpotrf(po.LOWER, n, A, n)
        assert info == 0
        assert np.array_equal(np.tril(A), L0)
        U = np.asfortranarray(L0 @ L0.T)
        assert po.lib().orc_dpotrf(po.UPPER, n, U, n) == 0
        assert np.array_equal(np.triu(U), L0.T)


def test_potrf_kernel_info(po):
    A = np.asfortranarray(np.diag([4.0, 9.0, -1.0, 16.0]))
    assert po.lib().orc_dpotrf(po.LOWER, 4, A, 4) == 3
    assert A[0, 0] == 2.0 and A[1, 1] == 3.0


def test_tile_kernels_against_numpy(po):
    L = po.lib()
    rng = np.random.default_rng(2)
    m, n, k = 7, 5, 6
    A = np.asfortranarray(rng.standard_normal((m, k)))
    B = np.asfortranarray(rng.standard_normal((n, k)))
    C = np.asfortranarray(rng.standard_normal((m, n)))
    C1 = C.copy(order="F")
    L.orc_dgemm(po.NOTRANS, po.TRANS, m, n, k, -1.0, A, m, B, n, 1.0, C1, m)
    assert np.allclose(C1, C - A @ B.T, atol=1e-13)
    S = np.asfortranarray(rng.standard_normal((n, n)))
    S1 = S.copy(order="F")
    Bk = np.asfortranarray(rng.standard_normal((n, k)))
    L.orc_dsyrk_ln(n, k, -1.0, Bk, n, 1.0, S1, n)
    ref = S - Bk @ Bk.T
    assert np.allclose(np.tril(S1), np.tril(ref), atol=1e-13)
    assert np.array_equal(np.triu(S1, 1), np.triu(S, 1))
    # all Left/Right x Lower/Upper x N/T x diag variants against a dense solve
    T = np.asfortranarray(np.tril(rng.standard_normal((n, n))) + 4 * np.eye(n))
    for side in (po.LEFT, po.RIGHT):
        for uplo in (po.LOWER, po.UPPER):
            Tm = T if uplo == po.LOWER else np.asfortranarray(T.T)
            for tr in (po.NOTRANS, po.TRANS):
                for diag in (po.NONUNIT, po.UNIT):
                    Teff = np.tril(Tm) if uplo == po.LOWER else np.triu(Tm)
                    if diag == po.UNIT:
                        Teff = Teff - np.diag(np.diag(Teff)) + np.eye(n)
                    op = Teff if tr == po.NOTRANS else Teff.T
                    X = np.asfortranarray(rng.standard_normal((n, 3) if side == po.LEFT else (3, n)))
                    X1 = X.copy(order="F")
                    mm, nn = X.shape
                    L.orc_dtrsm(side, uplo, tr, diag, mm, nn, 2.0, Tm, n, X1, mm)
                    exp = np.linalg.solve(op, 2.0 * X) if side == po.LEFT else np.linalg.solve(op.T, 2.0 * X.T).T
                    assert np.allclose(X1, exp, atol=1e-12), (side, uplo, tr, diag)


@pytest.mark.parametrize("N,NB", [(378, 93), (300, 64), (256, 64), (100, 128), (1, 8)])
def test_dag_equals_sequential_bitwise(po, N, NB):
    T = po.dplgsy_symtiles(N, NB, float(N), 3872)
    T1, T2 = T.copy(), T.copy()
    assert po.dpotrf_tiles(T1, N, NB, nthreads=4) == 0
    assert po.lib().orc_seq_dpotrf_L(N, NB, T2, 0) == 0
    assert np.array_equal(T1, T2)


@pytest.mark.parametrize("N,NB", [(378, 93), (512, 128), (97, 32)])
def test_dag_factor_against_numpy(po, N, NB):
    T = po.dplgsy_symtiles(N, NB, float(N), 3872)
    A0 = po.dplgsy_full(N, float(N), 3872)
    assert po.dpotrf_tiles(T, N, NB, nthreads=3) == 0
    Lf = po.symtiles_to_full(T, N, NB)
    assert np.abs(np.tril(Lf) - np.linalg.cholesky(A0)).max() < 1e-12
    assert po.check_dpotrf(po.LOWER, Lf, A0) < 60
    # untouched strict upper part of the diagonal tiles keeps A's values
    for k in range((N + NB - 1) // NB):
        s = slice(k * NB, min(N, (k + 1) * NB))
        assert np.array_equal(np.triu(Lf[s, s], 1), np.triu(A0[s, s], 1))


def test_dag_info_positions(po):
    N, NB = 256, 64
    for r in (0, 10, 63, 64, 130, 255):
        A0 = po.dplgsy_full(N, float(N), 3872)
        A0[r, r] = -1.0 - abs(A0[r, :r] @ A0[r, :r]) if r else -1.0
        T = po.full_to_symtiles(A0, NB)
        assert po.dpotrf_tiles(T, N, NB, nthreads=4) == r + 1


def test_check_functions_detect_planted_errors(po):
    N = 64
    A0 = po.dplgsy_full(N, float(N), 5)
    L = np.asfortranarray(np.linalg.cholesky(A0))
    assert po.check_dpotrf(po.LOWER, L, A0) < 10
    Lbad = L.copy(order="F")
    Lbad[40, 3] += 1e-9
    assert po.check_dpotrf(po.LOWER, Lbad, A0) > 60
    # transposed operand: a U^T U check of the lower factor must fail
    assert po.check_dpotrf(po.UPPER, np.asfortranarray(L), A0) > 60
    # upper variant on the correct factor
    assert po.check_dpotrf(po.UPPER, np.asfortranarray(L.T), A0) < 10
    b = po.dplrnt_full(N, 2, 7)
    x = np.asfortranarray(np.linalg.solve(A0, b))
    assert po.check_daxmb(po.LOWER, A0, b, x) < 10
    x[5, 1] *= 1 + 1e-10
    assert po.check_daxmb(po.LOWER, A0, b, x) > 60
    # hand-computed residual ratio on a 2x2 case
    A2 = np.asfortranarray([[4.0, 1.0], [1.0, 3.0]])
    b2 = np.asfortranarray([[1.0], [2.0]])
    x2 = np.asfortranarray([[0.0], [0.0]])
    eps = po.lib().orc_dlamch_eps()
    r = po.check_daxmb(po.LOWER, A2, b2, x2)
    assert r == pytest.approx(2.0 / ((5.0 * 0.0 + 2.0) * 2 * eps))


def test_tile_solve_against_numpy(po):
    N, NB, nrhs = 378, 93, 3
    T = po.dplgsy_symtiles(N, NB, float(N), 3872)
    A0 = po.dplgsy_full(N, float(N), 3872)
    assert po.dpotrf_tiles(T, N, NB) == 0
    b = po.dplrnt_full(N, nrhs, 3873)
    Bt = po.full_to_tiles(b, NB, NB)
    po.lib().orc_tile_dpotrs_L(N, NB, T, nrhs, Bt)
    x = po.tiles_to_full(Bt, N, nrhs, NB, NB)
    assert np.abs(x - np.linalg.solve(A0, b)).max() < 1e-13
    assert po.check_daxmb(po.LOWER, A0, b, x) < 60


def test_optimized_blas_path_agrees(po):
    if po.load_openblas() is None:
        pytest.skip("no optimized BLAS library found")
    N, NB = 512, 128
    T = po.dplgsy_symtiles(N, NB, float(N), 3872)
    T1 = T.copy()
    assert po.dpotrf_tiles(T, N, NB, nthreads=4, use_blas=True) == 0
    assert po.dpotrf_tiles(T1, N, NB, nthreads=4, use_blas=False) == 0
    assert np.abs(T - T1).max() < 1e-12


def test_check_dpoinv_pinned(po):
    """orc_check_dpoinv (src/dplasma_zcheck.c:264-314): exact inverses give ~0
    for both uplo; a hand-computed 2x2 case; a planted 1e-9 error and a
    transposed non-symmetric 'inverse' fail the < 10 threshold."""
    I = np.eye(7)
    assert po.check_dpoinv(po.LOWER, I, I) == 0.0
    A = np.array([[4.0, 2.0], [2.0, 3.0]])
    Ai = np.array([[3.0, -2.0], [-2.0, 4.0]]) / 8.0
    assert po.check_dpoinv(po.LOWER, A, Ai) < 1.0 and po.check_dpoinv(po.UPPER, A, Ai) < 1.0
    rng = np.random.default_rng(3)
    B = rng.standard_normal((40, 40))
    A = B @ B.T / 40 + np.eye(40)
    Ai = np.linalg.inv(A)
    assert po.check_dpoinv(po.LOWER, A, Ai) < 10
    bad = Ai.copy()
    bad[5, 2] += 1e-9
    assert po.check_dpoinv(po.LOWER, A, bad) > 10
    # only the uplo triangle is referenced: garbage above the diagonal is ignored
    junk = np.tril(Ai) + np.triu(np.full_like(Ai, 7.0), 1)
    assert po.check_dpoinv(po.LOWER, A, junk) < 10
    assert po.check_dpoinv(po.UPPER, A, junk) > 10
