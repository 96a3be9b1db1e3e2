"""mi_dpotrf_tiles_noalloc (the factorization that allocates nothing, its
scratch in a fixed per-GPU device array) against mi_dpotrf_tiles on the same
tiles: bitwise the same factor and info, on calls in a row (the first at the
largest NB its scratch holds, whose recursive diagonal factor fills the whole
sub-tile image; then smaller, the last with a ragged last tile and the
single-launch diagonal factor), plus the rejection of NB beyond the fixed
scratch and a non-SPD matrix's info."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run(dev, entry, T, N, NB):
    d = dev.upload(T)
    info = ctypes.c_int(-1)
    try:
        dev.check(getattr(dev.M, entry)(dev.h, N, NB, d, ctypes.byref(info)), entry)
        out = dev.download(d, T.shape)
    finally:
        dev.free(d)
    return out, info.value


def test_noalloc_matches_dpotrf_tiles_twice(dev, po):
    for N, NB in [(4096, 1024), (4096, 512), (1700, 320)]:   # smaller each call, the last ragged
        print(f"noalloc: N={N} NB={NB}", flush=True)
        T = po.dplgsy_symtiles(N, NB, float(N), 3872)
        a, ia = run(dev, "mi_dpotrf_tiles", T, N, NB)
        b, ib = run(dev, "mi_dpotrf_tiles_noalloc", T, N, NB)
        assert ia == ib == 0
        assert np.array_equal(a, b)
        To = T.copy()
        assert po.dpotrf_tiles(To, N, NB, nthreads=4) == 0
        Lg = np.tril(po.symtiles_to_full(b, N, NB))
        Lo = np.tril(po.symtiles_to_full(To, N, NB))
        assert np.abs(Lg - Lo).max() <= 1e-12 * np.abs(Lo).max()


def test_noalloc_info_matches(dev, po):
    N, NB = 2000, 256
    A0 = po.dplgsy_full(N, float(N), 3872)
    A0[700, 700] = -2.0 * N
    T = po.full_to_symtiles(A0, NB)
    a, ia = run(dev, "mi_dpotrf_tiles", T, N, NB)
    b, ib = run(dev, "mi_dpotrf_tiles_noalloc", T, N, NB)
    assert ia == ib == 701
    assert np.array_equal(a, b)


def test_noalloc_rejects_large_nb(dev, dp):
    info = ctypes.c_int(-1)
    rc = dev.M.mi_dpotrf_tiles_noalloc(dev.h, 4096, 2048, None, ctypes.byref(info))
    assert rc == -2
    assert "NB=2048" in dp.last_error()
