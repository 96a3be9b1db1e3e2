"""The persistent diagonal-factor server (mi_set_potrf_server) against
per-step POTRF launches: the same kernel body on the same tiles in the same
order, so the factors are bitwise equal, in FP64, FP32 and with the BF16-split
trailing update; the same info on a non-SPD matrix; and the server's bounded
waits forced to give up from the host (-5), after which the next
factorization is correct again."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def factor(dev, T, N, NB, server, fp32=False, bf16=False):
    M = dev.M
    assert M.mi_set_potrf_server(3 if server else 0) == 0
    info = ctypes.c_int(-1)
    try:
        if fp32:
            d = dev.upload(T.astype(np.float32))
            M.mi_set_spotrf_update(1 if bf16 else 0)
            rc = M.mi_spotrf_tiles(dev.h, N, NB, d, ctypes.byref(info))
            M.mi_set_spotrf_update(0)
            dev.check(rc, "spotrf")
            out = dev.download(d, T.shape, np.float32)
        else:
            d = dev.upload(T)
            dev.check(M.mi_dpotrf_tiles(dev.h, N, NB, d, ctypes.byref(info)), "dpotrf")
            out = dev.download(d, T.shape)
    finally:
        M.mi_set_potrf_server(0)
    dev.free(d)
    return out, info.value


@pytest.mark.parametrize("N,NB", [(4096, 512), (1700, 320), (378, 93), (2048, 256), (500, 512)])
@pytest.mark.parametrize("kind", ["fp64", "fp32", "bf16x3"])
def test_server_matches_launches(dev, po, N, NB, kind):
    T = po.dplgsy_symtiles(N, NB, float(N), 3872)
    kw = dict(fp32=kind != "fp64", bf16=kind == "bf16x3")
    a, ia = factor(dev, T, N, NB, server=False, **kw)
    b, ib = factor(dev, T, N, NB, server=True, **kw)
    assert ia == ib == 0
    assert np.array_equal(a, b)
    if kind == "fp64" and N <= 2048:
        To = T.copy()
        assert po.dpotrf_tiles(To, N, NB, nthreads=4) == 0
        Lg = np.tril(po.symtiles_to_full(b, N, NB))
        Lo = np.tril(po.symtiles_to_full(To, N, NB))
        assert np.abs(Lg - Lo).max() <= 1e-12 * np.abs(Lo).max()


@pytest.mark.parametrize("r", [0, 700, 1999])
def test_server_info(dev, po, r):
    N, NB = 2000, 256
    A0 = po.dplgsy_full(N, float(N), 3872)
    A0[r, r] = -2.0 * N
    T = po.full_to_symtiles(A0, NB)
    a, ia = factor(dev, T, N, NB, server=False)
    b, ib = factor(dev, T, N, NB, server=True)
    assert ia == ib == r + 1
    # tile columns before the failing one are final and equal; what follows
    # the failing pivot is undefined (as after LAPACK dpotrf with info > 0)
    MT = (N + NB - 1) // NB
    kf = r // NB
    ndone = kf * MT - kf * (kf - 1) // 2        # tiles of columns 0 .. kf-1
    assert np.array_equal(a[:ndone * NB * NB], b[:ndone * NB * NB])


def test_server_wait_limit(dev, po, dp):
    mi, _ = dp.libs()
    N, NB = 2048, 256
    T = po.dplgsy_symtiles(N, NB, float(N), 3872)
    d = dev.upload(T)
    info = ctypes.c_int(-1)
    assert mi.mi_set_potrf_server(3) == 0
    assert mi.mi_set_solve_wait_limit(0.0) == 0
    try:
        rc = dev.M.mi_dpotrf_tiles(dev.h, N, NB, d, ctypes.byref(info))
        msg = dp.last_error()
    finally:
        assert mi.mi_set_solve_wait_limit(10.0) == 0
        assert mi.mi_set_potrf_server(0) == 0
    assert rc == -5, (rc, msg)
    assert "server" in msg
    dev.free(d)
    b, ib = factor(dev, T, N, NB, server=True)
    a, ia = factor(dev, T, N, NB, server=False)
    assert ia == ib == 0 and np.array_equal(a, b)
