"""The FP64 headline configuration at full size: dplasma_dplgsy (bump N,
seed 3872), dpotrf Lower, N = 32768, NB = 512, then dpotrs with one
right-hand side b = dplrnt(seed 3873) (tests/testing_zpotrf.c:57-127).

  * the GPU factor equals the oracle's tile-DAG factor of the same matrix
    (the task classes of src/zpotrf_L.jdf on the host cores) to 1e-12
    relative over every referenced element;
  * the GPU solve equals the oracle's tile solve (ztrsm LLN + LLT) applied to
    the GPU factor to 1e-12 relative;
  * the reference's checks pass on the GPU: ||A - L L^T|| ratio and
    ||b - A x|| ratio below 60.
The oracle run takes about 10 s on 16 host cores with OpenBLAS tile kernels.
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, NB = 32768, 512
EPS = 2.0 ** -53


def lower_tile_max_diff(G, C, MT):
    g = G.reshape(-1, NB, NB)
    c = C.reshape(-1, NB, NB)
    diag = set(n * MT - n * (n - 1) // 2 for n in range(MT))
    low = np.triu(np.ones((NB, NB), dtype=bool))      # [col][row]: row >= col
    d = s = 0.0
    for t in range(g.shape[0]):
        if t in diag:
            d = max(d, np.abs(g[t] - c[t])[low].max())
            s = max(s, np.abs(c[t])[low].max())
        else:
            d = max(d, np.abs(g[t] - c[t]).max())
            s = max(s, np.abs(c[t]).max())
    return d, s


def test_dpotrf_dpotrs_full_size_vs_oracle(dev, po, dp):
    M = dev.M
    MT = N // NB
    nt = MT * (MT + 1) // 2 * NB * NB
    print(f"[fullsize] N={N} NB={NB}: GPU dplgsy + dpotrf", flush=True)
    dA = dev.alloc(nt * 8)
    dA0 = dev.alloc(nt * 8)
    dev.check(M.mi_dplgsy_tiles(dev.h, float(N), N, NB, dA, 3872), "dplgsy")
    dev.check(M.mi_dplgsy_tiles(dev.h, float(N), N, NB, dA0, 3872), "dplgsy")
    info = ctypes.c_int(-1)
    dev.check(M.mi_dpotrf_tiles(dev.h, N, NB, dA, ctypes.byref(info)), "dpotrf")
    assert info.value == 0
    an, rn = ctypes.c_double(), ctypes.c_double()
    dev.check(M.mi_dcheck_potrf(dev.h, N, NB, dA, dA0, ctypes.byref(an), ctypes.byref(rn)), "check")
    ratio = rn.value / (an.value * N * EPS)
    assert ratio < 60, ratio
    G = dev.download(dA, (nt,))

    print("[fullsize] oracle tile-DAG factor on the host cores", flush=True)
    blas = po.load_openblas()
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    T = po.dplgsy_symtiles(N, NB, float(N), 3872)
    assert po.dpotrf_tiles(T, N, NB, nthreads=max(1, min(cores, 32)), use_blas=blas is not None) == 0
    d, s = lower_tile_max_diff(G, T, MT)
    print(f"[fullsize] factor: max |L_gpu - L_oracle| = {d:.3e}, max |L| = {s:.3e}", flush=True)
    assert d <= 1e-12 * s

    print("[fullsize] dpotrs, one right-hand side", flush=True)
    b = po.dplrnt_full(N, 1, 3873)[:, 0].copy()
    dB = dev.upload(b)
    dev.check(M.mi_dpotrs_tiles(dev.h, N, NB, dA, 1, dB, N), "dpotrs")
    x = dev.download(dB, (N,))
    Bt = po.full_to_tiles(b.reshape(N, 1), NB, NB)
    po.lib().orc_tile_dpotrs_L(N, NB, G, 1, Bt)
    xo = po.tiles_to_full(Bt, N, 1, NB, NB)[:, 0]
    assert np.abs(x - xo).max() <= 1e-12 * np.abs(xo).max()
    db = dev.upload(b)
    bn, xn = ctypes.c_double(), ctypes.c_double()
    dev.check(M.mi_dcheck_axmb(dev.h, N, NB, dA0, 1, db, N, dB, N, ctypes.byref(an), ctypes.byref(bn),
                               ctypes.byref(xn), ctypes.byref(rn)), "check_axmb")
    r = rn.value / ((an.value * xn.value + bn.value) * N * EPS)
    print(f"[fullsize] check_dpotrf {ratio:.3e}, check_daxmb {r:.3e}", flush=True)
    assert r < 60, r
    for p in (dA, dA0, dB, db):
        dev.free(p)
