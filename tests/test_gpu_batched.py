"""GPU parity of the batched factor-and-solve (mi_dposv_batched /
mi_sposv_batched) through the C ABI against the oracle (oracle/batched.py)
on the same matrices: FP64 factors within 1e-12 and solutions within 1e-11
(relative to the largest entry; the matrices are G^T G / n + I, condition
below ~10), FP32 within 2e-5 of the FP64 oracle on the FP32 inputs.  info of
a non-SPD matrix in the batch, graph replays against the eager run, the
launch loop and every tile shape of the autotuner give bit-identical
results (same summation order)."""
import ctypes
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import batched as ob  # noqa: E402

pytestmark = pytest.mark.gpu

EDGE = [1, 2, 3, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129, 200, 255, 256, 257, 300, 511, 512, 513, 700,
        1000, 1023, 1024]


def make_batch(ns, seed, dtype=np.float64):
    rng = np.random.default_rng(seed)
    As, bs = [], []
    for n in ns:
        G = rng.standard_normal((n, n))
        A = G.T @ G / n + np.eye(n)
        As.append(A.astype(dtype))
        bs.append(rng.uniform(-0.5, 0.5, n).astype(dtype))
    return As, bs


def flat(As, bs, dtype, upper_fill=np.nan):
    """Concatenated column-major storage; strict upper triangles set to
    `upper_fill` (the kernels read only the lower triangle)."""
    a = []
    for A in As:
        M = np.array(A, dtype=dtype)
        M[np.triu_indices(M.shape[0], 1)] = upper_fill
        a.append(M.ravel(order="F"))
    return np.concatenate(a), np.concatenate(bs).astype(dtype)


class Batch:
    def __init__(self, dev, ns):
        self.dev, self.ns = dev, list(ns)
        self.p = ctypes.c_void_p()
        arr = (ctypes.c_int * len(ns))(*ns)
        dev.check(dev.M.mi_batch_create(dev.h, len(ns), arr, ctypes.byref(self.p)), "batch_create")

    def run(self, a, b, dtype=np.float64, graph=0, loop=False):
        dev = self.dev
        da, db = dev.upload(a), dev.upload(b)
        if loop:
            fn = dev.M.mi_dposv_batched_loop if dtype == np.float64 else dev.M.mi_sposv_batched_loop
            dev.check(fn(self.p, da, db), "posv_batched_loop")
        else:
            fn = dev.M.mi_dposv_batched if dtype == np.float64 else dev.M.mi_sposv_batched
            dev.check(fn(self.p, da, db, graph), "posv_batched")
        ra, rb = dev.download(da, a.shape, dtype), dev.download(db, b.shape, dtype)
        dev.free(da)
        dev.free(db)
        return ra, rb, self.info()

    def info(self):
        out = (ctypes.c_int * len(self.ns))()
        self.dev.check(self.dev.M.mi_batch_info(self.p, out), "info")
        return list(out)

    def close(self):
        self.dev.M.mi_batch_destroy(self.p)


def split(ns, a, b):
    ao, bo, _, _ = ob.offsets(ns)
    return ([a[o:o + n * n].reshape(n, n, order="F") for o, n in zip(ao, ns)],
            [b[o:o + n] for o, n in zip(bo, ns)])


def check(ns, As, bs, ra, rb, tol_l, tol_x):
    Ls, xs = split(ns, ra, rb)
    for n, A, b, L, x in zip(ns, As, bs, Ls, xs):
        Lo, xo, info = ob.posv(A.astype(np.float64), b.astype(np.float64))
        assert info == 0
        got = np.tril(L.astype(np.float64))
        assert np.abs(got - Lo).max() <= tol_l * np.abs(Lo).max(), n
        assert np.abs(x.astype(np.float64) - xo).max() <= tol_x * max(1.0, np.abs(xo).max()), n
        # the strict upper triangle is not written
        assert np.isnan(L[np.triu_indices(n, 1)]).all()


@pytest.mark.parametrize("graph", [0, 1])
def test_batched_fp64_edges(dev, graph):
    rng = np.random.default_rng(3)
    ns = EDGE + list(rng.integers(1, 400, 40))
    As, bs = make_batch(ns, 1)
    a, b = flat(As, bs, np.float64)
    bt = Batch(dev, ns)
    ra, rb, info = bt.run(a, b, graph=graph)
    bt.close()
    assert info == [0] * len(ns)
    check(ns, As, bs, ra, rb, 1e-12, 1e-11)


def test_batched_fp32(dev):
    ns = EDGE
    As, bs = make_batch(ns, 2, np.float32)
    a, b = flat(As, bs, np.float32)
    bt = Batch(dev, ns)
    ra, rb, info = bt.run(a, b, np.float32)
    bt.close()
    assert info == [0] * len(ns)
    check(ns, As, bs, ra, rb, 2e-5, 2e-5)


def test_batched_info_non_spd(dev):
    ns = [50, 300, 77, 1000, 33]
    As, bs = make_batch(ns, 4)
    As[1][140, 140] = -5.0            # first failure at column 141 of matrix 1
    As[4][0, 0] = 0.0                 # and at column 1 of matrix 4
    a, b = flat(As, bs, np.float64)
    bt = Batch(dev, ns)
    ra, rb, info = bt.run(a, b)
    bt.close()
    assert info == [0, ob.potrf(As[1])[1], 0, 0, 1]
    assert info[1] == 141
    keep = [0, 2, 3]
    Ls, xs = split(ns, ra, rb)
    for i in keep:
        Lo, xo, _ = ob.posv(As[i], bs[i])
        assert np.abs(xs[i] - xo).max() <= 1e-11 * max(1.0, np.abs(xo).max())
    # no solve for a failed matrix: its right-hand side is untouched
    assert np.array_equal(xs[1], bs[1]) and np.array_equal(xs[4], bs[4])


def test_graph_replays_match_eager(dev):
    """Capture once, replay three times on fresh copies of the input with
    eager work and synchronizations between the replays: every replay equals
    the eager result bit for bit."""
    ns = list(np.random.default_rng(9).integers(1, 600, 60))
    As, bs = make_batch(ns, 5)
    a, b = flat(As, bs, np.float64)
    bt = Batch(dev, ns)
    ea, eb, _ = bt.run(a, b, graph=0)
    da, db = dev.upload(a), dev.upload(b)
    dscr = dev.alloc(b.nbytes)
    for rep in range(3):
        dev.check(dev.M.mi_memcpy_h2d(dev.h, da, a.ctypes.data, a.nbytes), "h2d")
        dev.check(dev.M.mi_memcpy_h2d(dev.h, db, b.ctypes.data, b.nbytes), "h2d")
        dev.check(dev.M.mi_dposv_batched(bt.p, da, db, 1), "graph run")
        ga, gb = dev.download(da, a.shape), dev.download(db, b.shape)
        assert np.array_equal(ga, ea, equal_nan=True) and np.array_equal(gb, eb), rep
        # eager work of the program between replays
        dev.check(dev.M.mi_dcopy(dev.h, b.size, db, dscr), "copy")
        dev.check(dev.M.mi_synchronize(dev.h), "sync")
    dev.free(da)
    dev.free(db)
    dev.free(dscr)
    bt.close()


def test_loop_and_shapes_bit_identical(dev):
    ns = [1000, 3, 64, 513, 250, 129, 31, 700]
    As, bs = make_batch(ns, 6)
    a, b = flat(As, bs, np.float64)
    bt = Batch(dev, ns)
    ra, rb, _ = bt.run(a, b)
    la, lb, info = bt.run(a, b, loop=True)
    assert info == [0] * len(ns)
    assert np.array_equal(la, ra, equal_nan=True) and np.array_equal(lb, rb)
    for nsplit, wg in ((0, 1), (256, 2), (1024, 1)):
        dev.check(dev.M.mi_batch_set_shape(bt.p, nsplit, wg), "shape")
        sa, sb, _ = bt.run(a, b)
        assert np.array_equal(sa, ra, equal_nan=True) and np.array_equal(sb, rb), (nsplit, wg)
    # the autotuner picks one of the shapes; results stay identical
    da0, db0 = dev.upload(a), dev.upload(b)
    da, db = dev.upload(a), dev.upload(b)
    choice = ctypes.c_int(-1)
    dev.check(dev.M.mi_dbatch_autotune(bt.p, da0, db0, da, db, ctypes.byref(choice)), "autotune")
    assert 0 <= choice.value < 8
    ta, tb, _ = bt.run(a, b)
    assert np.array_equal(ta, ra, equal_nan=True) and np.array_equal(tb, rb)
    for p in (da0, db0, da, db):
        dev.free(p)
    bt.close()


def test_batch_generator(dev):
    """mi_dbatch_spd: A_i = G_i^T G_i / n + I from consecutive elements of the
    mi_dgauss stream (checked against mi_dgauss itself), lower triangles."""
    ns = [5, 40, 3]
    ao, _, tot, _ = ob.offsets(ns)
    dA = dev.alloc(tot * 8)
    dev.check(dev.M.mi_memset(dev.h, dA, 0, tot * 8), "memset")
    arr = (ctypes.c_int * 3)(*ns)
    dev.check(dev.M.mi_dbatch_spd(dev.h, 3, arr, 11, dA), "gen")
    got = dev.download(dA, (tot,))
    dG = dev.alloc(tot * 8)
    dev.check(dev.M.mi_dgauss(dev.h, tot, dG, 11), "gauss")
    g = dev.download(dG, (tot,))
    for o, n in zip(ao, ns):
        G = g[o:o + n * n].reshape(n, n, order="F")
        want = np.tril(G.T @ G / n + np.eye(n))
        A = np.tril(got[o:o + n * n].reshape(n, n, order="F"))
        assert np.abs(A - want).max() <= 1e-13 * np.abs(want).max()
    dev.free(dA)
    dev.free(dG)
