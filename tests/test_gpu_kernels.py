"""GPU parity of the individual gfx950 kernels through the C ABI
(dplasma_mi.h), against the oracle on the same inputs.  Generators are
bit-exact; FP64 factor/solve kernels agree with the oracle to a relative
1e-12 (different summation orders), and every factor passes the reference's
||A - L L^T|| / (||A|| N eps) < 60 test.
"""
import ctypes
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "generator_fixtures.json")


def n_sym(N, NB):
    MT = (N + NB - 1) // NB
    return MT * (MT + 1) // 2 * NB * NB


@pytest.mark.parametrize("N,NB", [(378, 93), (1000, 256), (64, 64), (1, 16)])
def test_dplgsy_bit_exact(dev, po, N, NB):
    ref = po.dplgsy_symtiles(N, NB, float(N), 3872)
    d = dev.alloc(ref.nbytes)
    dev.check(dev.M.mi_memset(dev.h, d, 0, ref.nbytes), "memset")
    dev.check(dev.M.mi_dplgsy_tiles(dev.h, float(N), N, NB, d, 3872), "dplgsy")
    got = dev.download(d, ref.shape)
    dev.free(d)
    assert np.array_equal(got, ref)


def test_dplrnt_bit_exact_and_golden(dev, po):
    gold = json.load(open(GOLD))
    for case in gold["cases"]:
        if case["kind"] != "dplrnt":
            continue
        M, nr, seed = case["M"], case["N"], case["seed"]
        d = dev.alloc(M * nr * 8)
        dev.check(dev.M.mi_dplrnt(dev.h, M, nr, d, M, M, 0, 0, seed), "dplrnt")
        got = dev.download(d, (nr, M)).T
        dev.free(d)
        exp = np.array([[float.fromhex(v) for v in row] for row in case["values"]])
        assert np.array_equal(got, exp)
        assert np.array_equal(got, po.dplrnt_full(M, nr, seed))


@pytest.fixture(params=[0, 1], ids=["panel_valu", "panel_mfma"])
def panel(request, dev):
    dev.check(dev.M.mi_set_panel_engine(request.param), "panel engine")
    yield request.param
    dev.M.mi_set_panel_engine(1)


@pytest.mark.parametrize("n", [1, 31, 32, 33, 93, 256, 320, 512])
def test_potrf_diag_tile(dev, po, n, panel):
    A0 = po.dplgsy_full(n, float(n), 3872)
    T = np.asfortranarray(A0).reshape(-1, order="F").copy()
    d = dev.upload(T)
    info = ctypes.c_int(-1)
    dev.check(dev.M.mi_dpotrf_step_potrf(dev.h, n, n, d, 0, ctypes.byref(info)), "potrf")
    got = dev.download(d, (n * n,)).reshape((n, n), order="F")
    dev.free(d)
    assert info.value == 0
    ref = np.asfortranarray(A0.copy())
    assert po.lib().orc_dpotrf(po.LOWER, n, ref, n) == 0
    assert np.abs(np.tril(got) - np.tril(ref)).max() <= 1e-12 * np.abs(ref).max()
    # the strict upper triangle is never written
    assert np.array_equal(np.triu(got, 1), np.triu(A0, 1))


@pytest.mark.parametrize("n,bad", [(64, 0), (64, 40), (256, 100), (93, 92)])
def test_potrf_diag_info(dev, po, n, bad, panel):
    A0 = po.dplgsy_full(n, float(n), 3872)
    A0[bad, bad] = -5.0 * n
    d = dev.upload(np.asfortranarray(A0).reshape(-1, order="F"))
    info = ctypes.c_int(-1)
    dev.check(dev.M.mi_dpotrf_step_potrf(dev.h, n, n, d, 0, ctypes.byref(info)), "potrf")
    dev.free(d)
    assert info.value == bad + 1


@pytest.mark.parametrize("N,NB", [(378, 93), (640, 128), (1024, 512), (700, 320)])
def test_panel_trsm(dev, po, N, NB, panel):
    # after POTRF(0) on GPU, TRSM(1..MT-1, 0) against the oracle tile TRSM
    T = po.dplgsy_symtiles(N, NB, float(N), 3872)
    d = dev.upload(T)
    info = ctypes.c_int(-1)
    dev.check(dev.M.mi_dpotrf_step_potrf(dev.h, N, NB, d, 0, ctypes.byref(info)), "potrf")
    dev.check(dev.M.mi_dpotrf_step_trsm(dev.h, N, NB, d, 0), "trsm")
    got = dev.download(d, T.shape)
    dev.free(d)
    MT = (N + NB - 1) // NB
    L = po.lib()
    ref = T.copy()
    L00 = np.asfortranarray(ref[:NB * NB].reshape((NB, NB), order="F").copy())
    n0 = min(NB, N)
    assert L.orc_dpotrf(po.LOWER, n0, L00, NB) == 0
    scale = np.abs(L00).max()
    for m in range(1, MT):
        tm = min(NB, N - m * NB)
        C = np.asfortranarray(ref[m * NB * NB:(m + 1) * NB * NB].reshape((NB, NB), order="F").copy())
        L.orc_dtrsm(po.RIGHT, po.LOWER, po.TRANS, po.NONUNIT, tm, NB, 1.0, L00, NB, C, NB)
        g = got[m * NB * NB:(m + 1) * NB * NB].reshape((NB, NB), order="F")
        assert np.abs(g[:tm] - C[:tm]).max() <= 1e-12 * max(1.0, np.abs(C[:tm]).max()) * scale
        # padding rows of the ragged tile untouched
        assert np.array_equal(g[tm:], C[tm:])


@pytest.mark.parametrize("engine", [0, 1], ids=["valu", "mfma"])
@pytest.mark.parametrize("N,NB", [(378, 93), (1024, 256), (1536, 512), (700, 128), (1000, 320), (5120, 512)])
def test_trailing_update(dev, po, N, NB, engine):
    """One step's trailing update; small launches run 64x64 blocks, the
    5120/512 case (720 blocks of 128x128) the 128x128 kernel."""
    rng = np.random.default_rng(N)
    MT = (N + NB - 1) // NB
    T = rng.standard_normal(n_sym(N, NB))
    d = dev.upload(T)
    # whole trailing matrix of step 0
    dev.check(dev.M.mi_set_update_engine(engine), "engine")
    try:
        dev.check(dev.M.mi_dpotrf_step_update(dev.h, N, NB, d, 0, 1, MT), "update")
    finally:
        dev.M.mi_set_update_engine(1)
    got = dev.download(d, T.shape)
    dev.free(d)

    def tile(X, m, n):
        i = n * MT - n * (n - 1) // 2 + (m - n)
        return X[i * NB * NB:(i + 1) * NB * NB].reshape((NB, NB), order="F")

    for n in range(1, MT):
        for m in range(n, MT):
            tm, tn = min(NB, N - m * NB), min(NB, N - n * NB)
            A, B, C = tile(T, m, 0), tile(T, n, 0), tile(T, m, n)
            exp = C.copy()
            exp[:tm, :tn] = C[:tm, :tn] - A[:tm, :] @ B[:tn, :].T
            g = tile(got, m, n)
            if m == n:
                low = np.tril(np.ones((tm, tm), bool))
                assert np.abs(g[:tm, :tm][low] - exp[:tm, :tm][low]).max() < 1e-11
                assert np.array_equal(g[:tm, :tm][~low], C[:tm, :tm][~low])
            else:
                assert np.abs(g[:tm, :tn] - exp[:tm, :tn]).max() < 1e-11
            # padding never written
            assert np.array_equal(g[tm:, :], C[tm:, :])


def test_update_column_range_only(dev, po):
    N, NB = 1024, 256
    MT = 4
    rng = np.random.default_rng(3)
    T = rng.standard_normal(n_sym(N, NB))
    d = dev.upload(T)
    dev.check(dev.M.mi_dpotrf_step_update(dev.h, N, NB, d, 0, 2, 3), "update")
    got = dev.download(d, T.shape)
    dev.free(d)
    for n in range(MT):
        for m in range(n, MT):
            i = n * MT - n * (n - 1) // 2 + (m - n)
            s = slice(i * NB * NB, (i + 1) * NB * NB)
            if n != 2:
                assert np.array_equal(got[s], T[s]), (m, n)


def _tri_tiles(po, N, NB, seed):
    """Lower triangular L (well conditioned: |off-diagonal| <= 1/N, diagonal in
    [1, 2]) in packed lower tiles; the strict upper triangle of the diagonal
    tiles and all padding hold NaN, so any read of them shows in the result."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1.0, 1.0, (N, N)) / N
    A[np.diag_indices(N)] = 1.0 + rng.uniform(0.0, 1.0, N)
    L = np.tril(A)
    T = po.full_to_symtiles(L, NB)
    MT = (N + NB - 1) // NB
    for n in range(MT):
        for m in range(n, MT):
            i = n * MT - n * (n - 1) // 2 + (m - n)
            t = T[i * NB * NB:(i + 1) * NB * NB].reshape((NB, NB), order="F")
            tm, tn = min(NB, N - m * NB), min(NB, N - n * NB)
            t[tm:, :] = np.nan
            t[:, tn:] = np.nan
            if m == n:
                t[np.triu_indices(tm, 1)] = np.nan
    return L, T


@pytest.mark.parametrize("N,NB,nrhs,trans,unit,alpha", [
    (700, 128, 1, 111, False, 1.0), (700, 128, 1, 112, False, 1.0),
    (1000, 320, 5, 111, True, -0.5), (1000, 320, 5, 112, True, -0.5),
    (378, 93, 16, 111, False, 1.0), (378, 93, 17, 112, False, 2.0),
    (378, 93, 1, 112, False, 1.0), (64, 64, 1, 111, False, 1.0), (50, 64, 3, 112, False, 1.0),
    (2048, 512, 1, 111, False, 1.0), (2048, 512, 1, 112, False, 1.0),
    (1300, 600, 1, 111, False, 2.0), (1300, 600, 2, 112, False, 1.0),
    (2000, 960, 1, 112, False, 1.0),
])
def test_dtrsm_LL_kernel(dev, po, N, NB, nrhs, trans, unit, alpha):
    """Left/Lower tile solve (ztrsm_LLN.jdf / ztrsm_LLT.jdf) against the oracle
    tile TRSM, column-major right-hand sides with ldb > N (padding rows of B
    hold NaN and must stay untouched)."""
    L, T = _tri_tiles(po, N, NB, N + nrhs)
    rng = np.random.default_rng(7)
    Bh = rng.standard_normal((N, nrhs))
    ldb = N + 3
    Bpad = np.full((ldb, nrhs), np.nan)
    Bpad[:N] = Bh
    dT = dev.upload(T)
    dB = dev.upload(np.asfortranarray(Bpad).reshape(-1, order="F"))
    diag = 132 if unit else 131
    dev.check(dev.M.mi_dtrsm_LL(dev.h, trans, diag, alpha, N, NB, dT, nrhs, dB, ldb), "dtrsm")
    got = dev.download(dB, (ldb * nrhs,)).reshape((ldb, nrhs), order="F")
    dev.free(dT)
    dev.free(dB)
    assert np.isnan(got[N:]).all()
    Tc = np.nan_to_num(T, nan=0.0)
    Bt = po.full_to_tiles(Bh, NB, NB)
    po.lib().orc_tile_dtrsm_LL(po.TRANS if trans == 112 else po.NOTRANS,
                               po.UNIT if unit else po.NONUNIT, alpha, N, NB, Tc, nrhs, Bt)
    xo = po.tiles_to_full(Bt, N, nrhs, NB, NB)
    assert np.isfinite(got[:N]).all()
    assert np.abs(got[:N] - xo).max() <= 1e-12 * max(1.0, np.abs(xo).max())
    # and against the mathematics: op(L) X = alpha B
    Lu = L.copy()
    if unit:
        Lu[np.diag_indices(N)] = 1.0
    op = Lu.T if trans == 112 else Lu
    assert np.abs(op @ got[:N] - alpha * Bh).max() <= 1e-12 * max(1.0, abs(alpha)) * np.abs(Bh).max() * N


# ---------------------------------------------------------------- FP32 ("s")
EPS32 = 2.0 ** -24


@pytest.mark.parametrize("N,NB", [(378, 93), (1000, 256), (1536, 512), (700, 320), (64, 64)])
def test_spotrf_tiles(dev, po, N, NB):
    """FP32 factorization (FP32 MFMA update, FP32 diagonal factor and panel
    solve) of the dplgsy matrix rounded to FP32: the reference's
    ||A - L L^T|| / (||A|| N eps) < 60 test with the FP32 eps, and agreement
    with the FP64 oracle factor of the same FP32 input to FP32 accuracy."""
    T64 = po.dplgsy_symtiles(N, NB, float(N), 3872)
    T32 = T64.astype(np.float32)
    d = dev.upload(T32)
    info = ctypes.c_int(-1)
    dev.check(dev.M.mi_spotrf_tiles(dev.h, N, NB, d, ctypes.byref(info)), "spotrf")
    got = dev.download(d, T32.shape, np.float32)
    dev.free(d)
    assert info.value == 0
    S = po.symtiles_to_full(T32.astype(np.float64), N, NB)   # stored (lower) tiles only
    A = np.tril(S) + np.tril(S, -1).T
    L = np.tril(po.symtiles_to_full(got.astype(np.float64), N, NB))
    anorm = np.abs(A).sum(axis=1).max()
    res = np.abs(A - L @ L.T).sum(axis=1).max() / (anorm * N * EPS32)
    assert res < 60, res
    ref = T32.astype(np.float64)
    assert po.dpotrf_tiles(ref, N, NB) == 0
    Lo = np.tril(po.symtiles_to_full(ref, N, NB))
    assert np.abs(L - Lo).max() <= 256 * EPS32 * np.abs(Lo).max()


def test_spotrf_info(dev, po):
    N, NB = 700, 128
    A = po.dplgsy_full(N, float(N), 3872)
    A[300, 300] = -1.0
    T = po.full_to_symtiles(A, NB).astype(np.float32)
    d = dev.upload(T)
    info = ctypes.c_int(-1)
    dev.check(dev.M.mi_spotrf_tiles(dev.h, N, NB, d, ctypes.byref(info)), "spotrf")
    dev.free(d)
    assert info.value == 301


@pytest.mark.parametrize("N,NB,nrhs,trans", [(700, 128, 1, 111), (700, 128, 1, 112), (1000, 320, 16, 111),
                                             (1000, 320, 16, 112), (378, 93, 5, 112),
                                             (3000, 256, 17, 111), (3000, 256, 17, 112),
                                             (1000, 64, 16, 112), (1000, 96, 3, 111), (1000, 96, 3, 112),
                                             (5000, 256, 16, 111), (5000, 256, 16, 112),
                                             (4999, 512, 2, 111), (4999, 512, 2, 112)])
def test_strsm_LL_kernel(dev, po, N, NB, nrhs, trans):
    """FP32 tile solve against the FP64 oracle on the same FP32 data.  With
    NB a multiple of 32 the wavefront kernel runs; at N = 5000 and 4999 its
    workgroups multiply up to 38 earlier blocks of 128 rows two at a time
    (an even and an odd count, the last block 8 or 7 rows)."""
    L, T = _tri_tiles(po, N, NB, N + nrhs)
    T32 = T.astype(np.float32)
    rng = np.random.default_rng(9)
    B = rng.standard_normal((N, nrhs)).astype(np.float32)
    dT = dev.upload(T32)
    dB = dev.upload(np.asfortranarray(B).reshape(-1, order="F"))
    dev.check(dev.M.mi_strsm_LL(dev.h, trans, 131, 1.0, N, NB, dT, nrhs, dB, N), "strsm")
    got = dev.download(dB, (N * nrhs,), np.float32).reshape((N, nrhs), order="F").astype(np.float64)
    dev.free(dT)
    dev.free(dB)
    Tc = np.nan_to_num(T32.astype(np.float64), nan=0.0)
    Bt = po.full_to_tiles(B.astype(np.float64), NB, NB)
    po.lib().orc_tile_dtrsm_LL(trans, po.NONUNIT, 1.0, N, NB, Tc, nrhs, Bt)
    xo = po.tiles_to_full(Bt, N, nrhs, NB, NB)
    assert np.abs(got - xo).max() <= 64 * EPS32 * max(1.0, np.abs(xo).max())


@pytest.mark.parametrize("n", [31, 64, 93, 256, 512])
def test_s_potrf_diag_tile(dev, po, n):
    A0 = po.dplgsy_full(n, float(n), 3872).astype(np.float32)
    d = dev.upload(np.asfortranarray(A0).reshape(-1, order="F"))
    info = ctypes.c_int(-1)
    dev.check(dev.M.mi_spotrf_step_potrf(dev.h, n, n, d, 0, ctypes.byref(info)), "spotrf step")
    got = dev.download(d, (n * n,), np.float32).reshape((n, n), order="F").astype(np.float64)
    dev.free(d)
    assert info.value == 0
    ref = np.asfortranarray(A0.astype(np.float64))
    assert po.lib().orc_dpotrf(po.LOWER, n, ref, n) == 0
    assert np.abs(np.tril(got) - np.tril(ref)).max() <= 64 * EPS32 * np.abs(ref).max()
    assert np.array_equal(np.triu(got, 1), np.triu(A0, 1).astype(np.float64))


@pytest.mark.parametrize("N,NB", [(378, 93), (640, 128), (1024, 512)])
def test_s_panel_trsm(dev, po, N, NB):
    """TRSM(1..MT-1, 0) in FP32 against the FP64 oracle TRSM applied with the
    GPU's own L_00 (isolates the panel solve)."""
    T = po.dplgsy_symtiles(N, NB, float(N), 3872).astype(np.float32)
    d = dev.upload(T)
    info = ctypes.c_int(-1)
    dev.check(dev.M.mi_spotrf_step_potrf(dev.h, N, NB, d, 0, ctypes.byref(info)), "spotrf step")
    dev.check(dev.M.mi_spotrf_step_trsm(dev.h, N, NB, d, 0), "strsm step")
    got = dev.download(d, T.shape, np.float32).astype(np.float64)
    dev.free(d)
    MT = (N + NB - 1) // NB
    L00 = np.asfortranarray(got[:NB * NB].reshape((NB, NB), order="F").copy())
    for m in range(1, MT):
        tm = min(NB, N - m * NB)
        C = np.asfortranarray(T[m * NB * NB:(m + 1) * NB * NB].astype(np.float64).reshape((NB, NB), order="F"))
        po.lib().orc_dtrsm(po.RIGHT, po.LOWER, po.TRANS, po.NONUNIT, tm, NB, 1.0, L00, NB, C, NB)
        g = got[m * NB * NB:(m + 1) * NB * NB].reshape((NB, NB), order="F")
        assert np.abs(g[:tm] - C[:tm]).max() <= 64 * EPS32 * max(1.0, np.abs(C[:tm]).max()), m


@pytest.mark.parametrize("N,NB", [(378, 93), (1024, 256), (1536, 512)])
def test_s_trailing_update(dev, po, N, NB):
    rng = np.random.default_rng(N + 1)
    MT = (N + NB - 1) // NB
    T = rng.standard_normal(n_sym(N, NB)).astype(np.float32)
    d = dev.upload(T)
    dev.check(dev.M.mi_spotrf_step_update(dev.h, N, NB, d, 0, 1, MT), "supdate")
    got = dev.download(d, T.shape, np.float32).astype(np.float64)
    dev.free(d)
    T64 = T.astype(np.float64)

    def tile(X, m, n):
        i = n * MT - n * (n - 1) // 2 + (m - n)
        return X[i * NB * NB:(i + 1) * NB * NB].reshape((NB, NB), order="F")

    for n in range(1, MT):
        for m in range(n, MT):
            tm, tn = min(NB, N - m * NB), min(NB, N - n * NB)
            A, B, C = tile(T64, m, 0), tile(T64, n, 0), tile(T64, m, n)
            exp = C[:tm, :tn] - A[:tm, :] @ B[:tn, :].T
            g = tile(got, m, n)[:tm, :tn]
            tol = 8 * NB * EPS32 * (1.0 + np.abs(A).max() * np.abs(B).max())
            if m == n:
                low = np.tril(np.ones((tm, tm), bool))
                assert np.abs(g[low] - exp[low]).max() < tol, (m, n)
                assert np.array_equal(g[~low], C[:tm, :tm][~low]), (m, n)
            else:
                assert np.abs(g - exp).max() < tol, (m, n)
            assert np.array_equal(tile(got, m, n)[tm:, :], C[tm:, :])


@pytest.mark.parametrize("N,NB", [(378, 93), (1024, 256), (1536, 512)])
def test_s_trailing_update_bf16x3(dev, po, N, NB):
    """FP32 trailing update on BF16 MFMA with the two-term split: each
    product carries ~16 significant bits per operand, so the error against
    the FP64 update of the same FP32 data is bounded by ~NB 2^-16 |a||b|
    (and is far below what a single BF16 product would give)."""
    rng = np.random.default_rng(N + 2)
    MT = (N + NB - 1) // NB
    T = rng.standard_normal(n_sym(N, NB)).astype(np.float32)
    d = dev.upload(T)
    dev.check(dev.M.mi_set_spotrf_update(1), "bf16x3")
    try:
        dev.check(dev.M.mi_spotrf_step_update(dev.h, N, NB, d, 0, 1, MT), "supdate")
    finally:
        dev.M.mi_set_spotrf_update(0)
    got = dev.download(d, T.shape, np.float32).astype(np.float64)
    dev.free(d)
    T64 = T.astype(np.float64)

    def tile(X, m, n):
        i = n * MT - n * (n - 1) // 2 + (m - n)
        return X[i * NB * NB:(i + 1) * NB * NB].reshape((NB, NB), order="F")

    worst = 0.0
    for n in range(1, MT):
        for m in range(n, MT):
            tm, tn = min(NB, N - m * NB), min(NB, N - n * NB)
            A, B, C = tile(T64, m, 0), tile(T64, n, 0), tile(T64, m, n)
            exp = C[:tm, :tn] - A[:tm, :] @ B[:tn, :].T
            g = tile(got, m, n)[:tm, :tn]
            if m == n:
                low = np.tril(np.ones((tm, tm), bool))
                err = np.abs(g[low] - exp[low]).max()
                assert np.array_equal(g[~low], C[:tm, :tm][~low]), (m, n)
            else:
                err = np.abs(g - exp).max()
            worst = max(worst, err / (np.abs(A).max() * np.abs(B).max()))
            assert np.array_equal(tile(got, m, n)[tm:, :], C[tm:, :])
    assert worst < 4 * NB * 2.0 ** -16, worst
    assert worst > 0.0


@pytest.mark.parametrize("N,NB", [(1536, 512), (1700, 320), (4096, 256), (2000, 128), (378, 93),
                                  (9512, 512), (12000, 256), (11000, 1024)])
def test_bf16_panel_split_same_factor(dev, po, N, NB):
    """The BF16-split FP32 factor whose main-stream updates load each panel
    column split once per step into hi/lo BF16 images, against the factor
    whose update workgroups split the FP32 panel tiles themselves: the same
    bit for bit (the split of -a is minus the split of a, and rounding to
    nearest is symmetric in sign).  NB = 93 is not a multiple of 32 and
    takes the per-workgroup split both times.  At (9512, 512), (12000, 256)
    and (11000, 1024) the early main-stream launches cover enough tiles for
    the 256 x 256 block kernel (ragged last tiles of 296, 224 and 760 rows;
    four blocks a side at NB = 1024)."""
    T = po.dplgsy_symtiles(N, NB, float(N), 3872).astype(np.float32)
    out = []
    dev.check(dev.M.mi_set_spotrf_update(1), "bf16x3")
    try:
        for which in (0, 1):
            dev.check(dev.M.mi_set_bf16_panel_split(which), "panel split")
            d = dev.upload(T)
            info = ctypes.c_int(-1)
            try:
                dev.check(dev.M.mi_spotrf_tiles(dev.h, N, NB, d, ctypes.byref(info)), "spotrf")
                out.append(dev.download(d, T.shape, np.float32))
            finally:
                dev.free(d)
            assert info.value == 0
    finally:
        dev.M.mi_set_spotrf_update(0)
        dev.M.mi_set_bf16_panel_split(1)
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("N,NB", [(2048, 512), (4096, 512), (3000, 1024), (1700, 320)])
def test_spotrf_bf16x3_factor(dev, po, N, NB):
    """The BF16-split FP32 factor (whose diagonal tiles from step MT / 2 on
    are factored recursively with sub-tiles of NB / 2 when NB >= 512)
    against the FP64 oracle factor of the same FP32 input: the split keeps
    ~16 bits of each operand and each of the MT steps adds one update, so
    the factor is good to MT 2^-16 relative (a wrong sub-tile or image
    mapping gives an O(1) error)."""
    T = po.dplgsy_symtiles(N, NB, float(N), 3872).astype(np.float32)
    dev.check(dev.M.mi_set_spotrf_update(1), "bf16x3")
    d = dev.upload(T)
    info = ctypes.c_int(-1)
    try:
        dev.check(dev.M.mi_spotrf_tiles(dev.h, N, NB, d, ctypes.byref(info)), "spotrf")
        got = dev.download(d, T.shape, np.float32)
    finally:
        dev.free(d)
        dev.M.mi_set_spotrf_update(0)
    assert info.value == 0
    ref = T.astype(np.float64)
    assert po.dpotrf_tiles(ref, N, NB) == 0
    L = np.tril(po.symtiles_to_full(got.astype(np.float64), N, NB))
    Lo = np.tril(po.symtiles_to_full(ref, N, NB))
    rel = np.abs(L - Lo).max() / np.abs(Lo).max()
    print(f"bf16x3 factor N={N} NB={NB}: rel {rel:.3e}", flush=True)
    MT = (N + NB - 1) // NB
    assert rel <= MT * 2.0 ** -16, rel


def test_bf16_panel_split_rejects_unknown(dev, dp):
    assert dev.M.mi_set_bf16_panel_split(2) == -2
    assert "panel split" in dp.last_error()
