"""Error paths and isolation of the kernel library, on the GPU:
  * the mixed solve's FP32 overflow rules (LAPACK dsposv / dlag2s / dlat2s:
    an overflow in B gives ITER = -2 and the FP64 answer; values in the
    unreferenced triangle are not converted);
  * dplasma_dpotrs / dplasma_dposv reject a right-hand side whose row count
    is not the order of A, leaving A and B untouched;
  * contexts own their workspaces: calls on two contexts interleaved on one
    thread, and run at the same time from two threads, give the answers each
    gives alone;
  * the wavefront solve's bounded hand-off wait, forced to give up at once
    from the host (dpotrs and the mixed solve): the call returns -5 and the
    next call is correct again.
Every kernel in these tests exits normally."""
import ctypes
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def spd(N, seed, kappa=1e2):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((N, N)))
    A = (Q * np.geomspace(1.0, 1.0 / kappa, N)) @ Q.T
    return 0.5 * (A + A.T)


def colmajor(X):
    return np.asfortranarray(X).reshape(-1, order="F")


def run_dsposv(dev, T, N, NB, B, prec=0):
    nrhs = B.shape[1]
    dT, dB = dev.upload(T), dev.upload(colmajor(B))
    dX = dev.alloc(N * nrhs * 8)
    it, info = ctypes.c_int(-99), ctypes.c_int(-99)
    dev.check(dev.M.mi_dsposv_tiles(dev.h, N, NB, dT, nrhs, dB, N, dX, N, prec, ctypes.byref(it),
                                    ctypes.byref(info)), "dsposv")
    X = dev.download(dX, (N * nrhs,)).reshape((N, nrhs), order="F")
    for p in (dT, dB, dX):
        dev.free(p)
    return X, it.value, info.value


def test_dsposv_overflow_in_rhs_gives_iter_minus2(dev, po):
    N, NB = 64, 32
    d = np.linspace(2.0, 5.0, N)
    A = np.diag(d)
    B = np.ones((N, 2))
    B[5, 0] = 1e39                       # above FLT_MAX: dlag2s of B overflows
    X, it, info = run_dsposv(dev, po.full_to_symtiles(A, NB), N, NB, B)
    assert info == 0 and it == -2
    assert np.abs(X - B / d[:, None]).max() <= 1e-15 * np.abs(B / d[:, None]).max()


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("NB", [64, 93])
def test_dsposv_ignores_unreferenced_triangle(dev, po, prec, NB):
    """NB = 64 converts runs of four elements, NB = 93 one at a time."""
    N, nrhs = 300, 3
    A = spd(N, 3)
    T = po.full_to_symtiles(A, NB)
    MT = (N + NB - 1) // NB
    for k in range(MT):                  # 1e300 above the diagonal and in padding
        i = k * MT - k * (k - 1) // 2
        t = T[i * NB * NB:(i + 1) * NB * NB].reshape((NB, NB), order="F")
        t[np.triu_indices(NB, 1)] = 1e300
        t[min(NB, N - k * NB):, :] = 1e300
    B = np.random.default_rng(1).uniform(-0.5, 0.5, (N, nrhs))
    X, it, info = run_dsposv(dev, T, N, NB, B, prec)
    assert info == 0 and 0 <= it <= 30, it
    assert np.abs(X - np.linalg.solve(A, B)).max() < 1e-10


@pytest.mark.parametrize("NB", [64, 93])
def test_dsposv_overflow_in_matrix_gives_iter_minus2(dev, po, NB):
    """A diagonal entry above FLT_MAX in the referenced triangle (the matrix
    stays SPD): the rounding to FP32 overflows, ITER = -2, and the FP64
    factor and solve give the answer."""
    N, nrhs = 300, 2
    A = spd(N, 5)
    A[200, 200] = 1e39
    T = po.full_to_symtiles(A, NB)
    B = np.random.default_rng(2).uniform(-0.5, 0.5, (N, nrhs))
    X, it, info = run_dsposv(dev, T, N, NB, B, 1)
    assert it == -2 and info == 0
    assert np.abs(X - np.linalg.solve(A, B)).max() < 1e-10


def test_dpotrs_and_dposv_reject_wrong_rhs_rows(ctx, dp, po):
    N, NB = 378, 93
    A = dp.SymMatrix(ctx, N, NB, dp.LOWER)
    assert ctx.dplgsy(float(N), dp.LOWER, A, 3872) == 0
    a0 = A.array.copy()
    B = dp.GeMatrix(ctx, N + 22, 2, NB, NB)
    assert ctx.dplrnt(0, B, 3873) == 0
    b0 = B.array.copy()
    assert ctx.dposv(dp.LOWER, A, B) == -6
    assert np.array_equal(A.array, a0) and np.array_equal(B.array, b0)
    assert ctx.dpotrf(dp.LOWER, A) == 0
    assert ctx.dpotrs(dp.LOWER, A, B) == -6
    assert np.array_equal(B.array, b0)
    A.free()
    B.free()


def factor_solve(dev, po, N, NB, seed, nrhs):
    """GPU factor + solve of the dplgsy matrix of order N through one
    context; returns (x, reference x)."""
    T = po.dplgsy_symtiles(N, NB, float(N), seed)
    b = po.dplrnt_full(N, nrhs, seed + 1)
    dT, dB = dev.upload(T), dev.upload(colmajor(b))
    info = ctypes.c_int(-1)
    dev.check(dev.M.mi_dpotrf_tiles(dev.h, N, NB, dT, ctypes.byref(info)), "dpotrf")
    assert info.value == 0
    dev.check(dev.M.mi_dpotrs_tiles(dev.h, N, NB, dT, nrhs, dB, N), "dpotrs")
    x = dev.download(dB, (N * nrhs,)).reshape((N, nrhs), order="F")
    dev.free(dT)
    dev.free(dB)
    return x, np.linalg.solve(po.dplgsy_full(N, float(N), seed), b)


def test_two_contexts_interleaved(dp, po):
    """Different orders and tile sizes, so the two contexts' workspaces
    differ in size; each call's answer is what it is alone."""
    with dp.Device(0) as d1, dp.Device(0) as d2:
        for rep in range(2):
            x1, e1 = factor_solve(d1, po, 1100, 256, 11 + rep, 16)
            x2, e2 = factor_solve(d2, po, 600, 128, 21 + rep, 3)
            x3, e3 = factor_solve(d1, po, 1500, 512, 31 + rep, 1)
            for x, e in ((x1, e1), (x2, e2), (x3, e3)):
                assert np.abs(x - e).max() <= 1e-12 * np.abs(e).max()


def test_two_contexts_from_two_threads(dp, po):
    res, errs = {}, []

    def work(name, N, NB, seed, nrhs):
        try:
            with dp.Device(0) as d:
                out = []
                for rep in range(3):
                    out.append(factor_solve(d, po, N, NB, seed + rep, nrhs))
                res[name] = out
        except Exception as e:            # reported by the main thread
            errs.append(repr(e))

    th = [threading.Thread(target=work, args=("a", 2048, 256, 5, 16)),
          threading.Thread(target=work, args=("b", 1000, 128, 9, 2))]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=100)
    assert not errs, errs
    assert sorted(res) == ["a", "b"]
    for out in res.values():
        for x, e in out:
            assert np.abs(x - e).max() <= 1e-12 * np.abs(e).max()


def test_solve_wait_limit_gives_up_and_recovers(dev, po, dp):
    mi, _ = dp.libs()
    N, NB, nrhs = 1024, 256, 4
    T = po.dplgsy_symtiles(N, NB, float(N), 3872)
    dT = dev.upload(T)
    info = ctypes.c_int(-1)
    dev.check(dev.M.mi_dpotrf_tiles(dev.h, N, NB, dT, ctypes.byref(info)), "dpotrf")
    b = po.dplrnt_full(N, nrhs, 3873)
    dB = dev.upload(colmajor(b))
    assert mi.mi_set_solve_wait_limit(0.0) == 0
    try:
        rc = dev.M.mi_dpotrs_tiles(dev.h, N, NB, dT, nrhs, dB, N)
        msg = dp.last_error()
    finally:
        assert mi.mi_set_solve_wait_limit(10.0) == 0
    assert rc == -5, (rc, msg)
    assert "time limit" in msg
    # the next call clears the abort word and solves correctly
    bc = colmajor(b)
    dev.check(dev.M.mi_memcpy_h2d(dev.h, dB, bc.ctypes.data, bc.nbytes), "h2d")
    dev.check(dev.M.mi_dpotrs_tiles(dev.h, N, NB, dT, nrhs, dB, N), "dpotrs")
    x = dev.download(dB, (N * nrhs,)).reshape((N, nrhs), order="F")
    e = np.linalg.solve(po.dplgsy_full(N, float(N), 3872), b)
    assert np.abs(x - e).max() <= 1e-12 * np.abs(e).max()
    assert mi.mi_set_solve_wait_limit(-1.0) == -2
    dev.free(dT)
    dev.free(dB)


def test_dsposv_wait_limit_gives_up_and_recovers(dev, po, dp):
    """The mixed solve reads the abort word with its first norms: a zero wait
    limit makes its FP32 solves give up, the call returns -5, and the next
    call (limit restored) converges."""
    mi, _ = dp.libs()
    N, NB, nrhs = 1024, 256, 20          # two chunks of right-hand sides
    A = spd(N, 7)
    T = po.full_to_symtiles(A, NB)
    B = np.random.default_rng(4).uniform(-0.5, 0.5, (N, nrhs))
    dT, dB = dev.upload(T), dev.upload(colmajor(B))
    dX = dev.alloc(N * nrhs * 8)
    it, info = ctypes.c_int(-99), ctypes.c_int(-99)
    assert mi.mi_set_solve_wait_limit(0.0) == 0
    try:
        rc = dev.M.mi_dsposv_tiles(dev.h, N, NB, dT, nrhs, dB, N, dX, N, 1, ctypes.byref(it), ctypes.byref(info))
        msg = dp.last_error()
    finally:
        assert mi.mi_set_solve_wait_limit(10.0) == 0
    assert rc == -5, (rc, msg)
    assert "time limit" in msg
    for p in (dT, dB, dX):
        dev.free(p)
    X, it2, info2 = run_dsposv(dev, T, N, NB, B, 1)
    assert info2 == 0 and 0 <= it2 <= 30, it2
    assert np.abs(X - np.linalg.solve(A, B)).max() < 1e-10


def test_dpotrf_image_larger_than_hbm_returns_negative(ctx, dp):
    """A symmetric descriptor whose packed tile image (N = 300000, NB = 512:
    ~361 GB) exceeds the 288 GB of HBM: dplasma_dpotrf returns a negative
    code (out of device memory) instead of aborting, and the context stays
    usable.  The device allocation fails before any host tile is read, so
    the descriptor's host storage is a one-element stand-in."""
    _, D = dp.libs()
    N, NB = 300000, 512
    desc = dp.SymBlockCyclic()
    D.parsec_matrix_sym_block_cyclic_init(ctypes.byref(desc), dp.MATRIX_DOUBLE, 0, NB, NB, N, N,
                                          0, 0, N, N, 1, 1, dp.LOWER)
    host = np.zeros(1)
    desc.mat = host.ctypes.data
    mt = (N + NB - 1) // NB
    assert mt * (mt + 1) // 2 * NB * NB * 8 > 300e9
    print("out-of-HBM dpotrf: N=300000 NB=512", flush=True)
    rc = D.dplasma_dpotrf(ctx.ptr, dp.LOWER, ctypes.byref(desc))
    assert rc < 0
    assert "memory" in dp.last_error().lower()
    D.parsec_tiled_matrix_destroy(ctypes.byref(desc))
    # the context still factors a small matrix
    A = dp.SymMatrix(ctx, 378, 93, dp.LOWER)
    assert ctx.dplgsy(378.0, dp.LOWER, A, 3872) == 0
    assert ctx.dpotrf(dp.LOWER, A) == 0
    A.free()
