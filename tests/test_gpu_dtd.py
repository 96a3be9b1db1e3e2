"""The insert-task (DTD) front-end of the tile Cholesky
(src/dtd_wrappers/z{potrf,trsm,herk,gemm}.c, tests/testing_zpotrf_dtd.c).

  * the C tester on the reference's ctest case, potrf_dtd -N 874 -K 367
    -t 76 (tests/Testings.cmake:148), both uplo: check_dpotrf and
    check_daxmb pass;
  * the engine (mi_dtd_*) with the tasks inserted in the reference's
    right-looking order gives the oracle's tile-DAG factor to 1e-12, and a
    left-looking insertion order (every update of a column, then its POTRF
    and TRSMs) gives the same factor bit for bit (each tile receives its
    updates in the same order);
  * info of a non-SPD matrix is that of dplasma_dpotrf;
  * the PaRSEC-shaped insert call rejects tasks outside the tile-Cholesky
    pattern (-2) and enqueues nothing for them.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTER = os.path.join(ROOT, "dplasma_amd", "bin", "testing_dpotrf_dtd")


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_dtd_tester_reference_case(uplo):
    r = subprocess.run([TESTER, "-N", "874", "-K", "367", "-t", "76", "-x", "-u", uplo],
                       capture_output=True, text=True, timeout=120)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert "Factorization is CORRECT" in out and "Solution is CORRECT" in out, out
    assert "dpotrf_dtd" in out


def right_looking(MT):
    for k in range(MT):
        yield ("potrf", k)
        for m in range(k + 1, MT):
            yield ("trsm", m, k)
        for m in range(k + 1, MT):
            yield ("syrk", k, m)
            for n in range(m + 1, MT):
                yield ("gemm", n, m, k)


def left_looking(MT):
    for j in range(MT):
        for k in range(j):
            yield ("syrk", k, j)
            for m in range(j + 1, MT):
                yield ("gemm", m, j, k)
        yield ("potrf", j)
        for m in range(j + 1, MT):
            yield ("trsm", m, j)


def run_engine(dev, T, N, NB, order):
    M = dev.M
    d = dev.upload(T)
    eng = ctypes.c_void_p()
    dev.check(M.mi_dtd_create(dev.h, N, NB, d, ctypes.byref(eng)), "dtd_create")
    try:
        for t in order:
            f = {"potrf": M.mi_dtd_potrf, "trsm": M.mi_dtd_trsm, "syrk": M.mi_dtd_syrk, "gemm": M.mi_dtd_gemm}[t[0]]
            dev.check(f(eng, *t[1:]), t[0])
        info, nt = ctypes.c_int(-1), ctypes.c_long(0)
        dev.check(M.mi_dtd_wait(eng, ctypes.byref(info), ctypes.byref(nt)), "dtd_wait")
    finally:
        M.mi_dtd_destroy(eng)
    out = dev.download(d, T.shape)
    dev.free(d)
    return out, info.value, nt.value


@pytest.mark.parametrize("N,NB", [(874, 76), (1700, 320), (2048, 256)])
def test_dtd_engine_matches_oracle_and_orders(dev, po, N, NB):
    MT = (N + NB - 1) // NB
    T = po.dplgsy_symtiles(N, NB, float(N), 3872)
    a, ia, na = run_engine(dev, T, N, NB, right_looking(MT))
    b, ib, nb = run_engine(dev, T, N, NB, left_looking(MT))
    assert ia == ib == 0
    assert na == nb == MT + MT * (MT - 1) + MT * (MT - 1) * (MT - 2) // 6
    assert np.array_equal(a, b)
    To = T.copy()
    assert po.dpotrf_tiles(To, N, NB, nthreads=4) == 0
    Lg = np.tril(po.symtiles_to_full(a, N, NB))
    Lo = np.tril(po.symtiles_to_full(To, N, NB))
    assert np.abs(Lg - Lo).max() <= 1e-12 * np.abs(Lo).max()


def test_dtd_info(dev, po):
    N, NB, r = 874, 76, 400
    A0 = po.dplgsy_full(N, float(N), 3872)
    A0[r, r] = -2.0 * N
    T = po.full_to_symtiles(A0, NB)
    _, info, _ = run_engine(dev, T, N, NB, right_looking((N + NB - 1) // NB))
    assert info == r + 1


def test_dtd_insert_rejects_outside_pattern(ctx, dp, po):
    mi, D = dp.libs()
    N, NB = 378, 93
    A = dp.SymMatrix(ctx, N, NB, dp.LOWER)
    assert ctx.dplgsy(float(N), dp.LOWER, A, 3872) == 0
    a0 = A.array.copy()
    tp = D.parsec_dtd_taskpool_new()
    assert D.parsec_context_add_taskpool(ctx.ptr, tp) == 0
    assert D.parsec_dtd_data_collection_init(A.ptr) == 0
    gemm = D.parsec_dtd_create_dgemm_task_class(tp, 0, 0x7f)
    tile = lambda m, n: ctypes.c_void_p(D.parsec_dtd_tile_of(A.ptr, D.parsec_shim_data_key(A.ptr, m, n)))
    I = lambda v: ctypes.byref(ctypes.c_int(v))
    Dv = lambda v: ctypes.byref(ctypes.c_double(v))
    z = ctypes.c_int(0)
    args = [z, I(dp.NOTRANS), z, I(dp.TRANS), z, I(NB), z, I(NB), z, I(NB), z, Dv(+1.0),   # alpha +1: not a Cholesky update
            z, tile(3, 0), z, I(NB), z, tile(2, 0), z, I(NB), z, Dv(1.0), z, tile(3, 2), z, I(NB), z]
    rc = D.parsec_dtd_insert_task_with_task_class(ctypes.c_void_p(tp), ctypes.c_void_p(gemm), ctypes.c_int(1),
                                                  ctypes.c_int(0x7f), *args)
    assert rc == -2
    args[11] = Dv(-1.0)
    args[13] = tile(1, 0)                                 # A's row is not C's row
    rc = D.parsec_dtd_insert_task_with_task_class(ctypes.c_void_p(tp), ctypes.c_void_p(gemm), ctypes.c_int(1),
                                                  ctypes.c_int(0x7f), *args)
    assert rc == -2
    assert D.parsec_shim_dtd_ntasks(tp) == 0
    D.parsec_dtd_data_flush_all(tp, A.ptr)
    D.parsec_taskpool_wait(tp)
    D.parsec_context_wait(ctx.ptr)
    assert np.array_equal(A.array, a0)
    D.parsec_dtd_task_class_release(tp, gemm)
    D.parsec_taskpool_free(tp)
    D.parsec_dtd_data_collection_fini(A.ptr)
    A.free()
