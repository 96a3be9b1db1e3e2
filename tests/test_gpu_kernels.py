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
@pytest.mark.parametrize("N,NB", [(378, 93), (1024, 256), (1536, 512), (700, 128), (1000, 320)])
def test_trailing_update(dev, po, N, NB, engine):
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
