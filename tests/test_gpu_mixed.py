"""Mixed-precision solve (FP32 factor + FP64 refinement, LAPACK dsposv
semantics) and its FP64 building blocks, through the C ABI (dplasma_mi.h)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -53


def spd_cond(N, kappa, seed=7):
    """A = Q diag(lambda) Q^T: Q from the Householder QR of a Gaussian matrix,
    lambda geometric from 1 down to 1/kappa (the test-matrix recipe of the
    mixed-precision configuration, at test sizes)."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((N, N)))
    lam = np.geomspace(1.0, 1.0 / kappa, N)
    A = (Q * lam) @ Q.T
    return 0.5 * (A + A.T)


def nan_unreferenced(po, T, N, NB):
    """NaN in the strict upper triangle of the diagonal tiles and in all
    padding: a kernel that reads them produces NaN."""
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
    return T


@pytest.mark.parametrize("kernel", [1, 0], ids=["lds_chunks", "registers"])
@pytest.mark.parametrize("N,NB,nrhs", [(378, 93, 16), (1000, 256, 5), (700, 64, 1), (1100, 512, 17),
                                       (1024, 512, 16), (300, 32, 16), (1500, 640, 3),
                                       (777, 96, 16), (1000, 250, 7), (1000, 128, 16), (1300, 384, 16),
                                       (2100, 512, 3), (130, 128, 2)])
def test_dsymm_resid(dev, po, N, NB, nrhs, kernel):
    """R = B - A X on the packed lower tiles (NaN in every unreferenced
    element); NB = 128, 256, 384, 512 run the LDS-chunk kernel when it is
    selected, the others the register kernel both times."""
    dev.check(dev.M.mi_set_symm_kernel(kernel), "symm kernel")
    rng = np.random.default_rng(N)
    A = rng.standard_normal((N, N))
    A = A + A.T
    T = nan_unreferenced(po, po.full_to_symtiles(A, NB), N, NB)
    X = rng.standard_normal((N, nrhs))
    B = rng.standard_normal((N, nrhs))
    dT, dX, dB = dev.upload(T), dev.upload(np.asfortranarray(X).reshape(-1, order="F")), \
        dev.upload(np.asfortranarray(B).reshape(-1, order="F"))
    dR = dev.alloc(N * nrhs * 8)
    try:
        dev.check(dev.M.mi_dsymm_resid(dev.h, N, NB, dT, nrhs, dX, N, dB, N, dR, N), "resid")
    finally:
        dev.M.mi_set_symm_kernel(1)
    R = dev.download(dR, (N * nrhs,)).reshape((N, nrhs), order="F")
    for p in (dT, dX, dB, dR):
        dev.free(p)
    exp = B - A @ X
    assert np.abs(R - exp).max() <= 1e-13 * np.abs(A).sum(axis=1).max() * np.abs(X).max()


def test_dlag2s_overflow(dev):
    v = np.array([1.0, -2.5, 3.5e38, 1e300, 0.0])
    d = dev.upload(v)
    f = dev.alloc(v.size * 4)
    ovf = ctypes.c_int(-1)
    dev.check(dev.M.mi_dlag2s_tiles(dev.h, v.size, d, f, ctypes.byref(ovf)), "dlag2s")
    got = dev.download(f, (v.size,), np.float32)
    dev.free(d)
    dev.free(f)
    assert ovf.value == 1
    assert np.array_equal(got[[0, 1, 4]], np.float32([1.0, -2.5, 0.0]))
    d = dev.upload(v[:2])
    f = dev.alloc(8)
    dev.check(dev.M.mi_dlag2s_tiles(dev.h, 2, d, f, ctypes.byref(ovf)), "dlag2s")
    dev.free(d)
    dev.free(f)
    assert ovf.value == 0


@pytest.mark.parametrize("N,NB,nrhs,kappa,prec,expect", [
    (1000, 128, 16, 1e2, 0, "refined"), (1000, 128, 16, 1e5, 0, "refined"), (1000, 128, 16, 1e8, 0, "fallback"),
    (700, 93, 3, 1e3, 0, "refined"), (512, 512, 1, 1e2, 0, "refined"),
    (1000, 128, 16, 1e2, 1, "refined"), (1000, 128, 16, 1e5, 1, "refined"), (1000, 128, 16, 1e8, 1, "fallback"),
])
def test_dsposv(dev, po, N, NB, nrhs, kappa, prec, expect):
    """FP32 factor (prec 0: FP32 MFMA, 1: BF16 split updates) + FP64
    refinement reaches FP64-level backward error
    (LAPACK dsposv test: ||b - A x|| <= ||x|| ||A|| eps sqrt(N) per column),
    or falls back to the FP64 factorization when FP32 cannot converge
    (kappa * eps32 >~ 1), with the FP64-quality answer either way."""
    A = spd_cond(N, kappa)
    T = po.full_to_symtiles(A, NB)
    rng = np.random.default_rng(11)
    B = rng.uniform(-0.5, 0.5, (N, nrhs))
    dT = dev.upload(T)
    dB = dev.upload(np.asfortranarray(B).reshape(-1, order="F"))
    dX = dev.alloc(N * nrhs * 8)
    it, info = ctypes.c_int(-99), ctypes.c_int(-99)
    dev.check(dev.M.mi_dsposv_tiles(dev.h, N, NB, dT, nrhs, dB, N, dX, N, prec, ctypes.byref(it),
                                    ctypes.byref(info)), "dsposv")
    X = dev.download(dX, (N * nrhs,)).reshape((N, nrhs), order="F")
    Tafter = dev.download(dT, T.shape)
    for p in (dT, dB, dX):
        dev.free(p)
    assert info.value == 0
    if expect == "refined":
        assert 0 <= it.value <= 30, it.value
        assert np.array_equal(Tafter, T)          # A untouched on success
    else:
        assert it.value < 0, it.value
    anrm = np.abs(A).sum(axis=1).max()
    # refined: the dsposv stopping test (slack 4 for numpy's own rounding of
    # A x); fallback: the reference's check_daxmb scale, N eps
    scale = 4 * np.sqrt(N) if expect == "refined" else 60 * N
    for j in range(nrhs):
        r = B[:, j] - A @ X[:, j]
        assert np.abs(r).max() <= np.abs(X[:, j]).max() * anrm * EPS * scale, j
    # against the FP64 oracle solve
    xo = np.linalg.solve(A, B)
    assert np.abs(X - xo).max() <= 64 * kappa * EPS * np.sqrt(N) * np.abs(xo).max()


# ------------------------------------------------ SPD generator of given kappa
GOLD = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1


def splitmix64(x):
    """splitmix64 output for state increments already applied (numpy uint64)."""
    z = (x + np.uint64(GOLD)) & np.uint64(M64)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def gauss_ref(n, seed):
    """Restatement of the generator's Gaussian stream: words 2t, 2t+1 of the
    splitmix64 sequence from seed, Box-Muller."""
    t = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        a = splitmix64(np.uint64(seed) + (np.uint64(2) * t) * np.uint64(GOLD))
        b = splitmix64(np.uint64(seed) + (np.uint64(2) * t + np.uint64(1)) * np.uint64(GOLD))
    u1 = ((a >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (b >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def test_splitmix64_known_answer():
    # first outputs of splitmix64 seeded with 0 (published reference values)
    with np.errstate(over="ignore"):
        got = [int(splitmix64(np.uint64(w * GOLD & M64))) for w in range(3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_gauss_stream(dev):
    n = 20000
    d = dev.alloc(n * 8)
    dev.check(dev.M.mi_dgauss(dev.h, n, d, 7), "gauss")
    g = dev.download(d, (n,))
    dev.free(d)
    ref = gauss_ref(n, 7)
    assert np.abs(g - ref).max() <= 1e-14 * np.abs(ref).max()
    assert abs(g.mean()) < 0.03 and abs(g.std() - 1.0) < 0.03


@pytest.mark.parametrize("N,NB,kappa", [(300, 64, 1e5), (512, 128, 1e2), (257, 93, 1e8)])
def test_spd_generator(dev, po, N, NB, kappa):
    dQ = dev.alloc(N * N * 8)
    dev.check(dev.M.mi_dorth_haar(dev.h, N, 7, dQ), "orth")
    Q = dev.download(dQ, (N * N,)).reshape((N, N), order="F")
    # orthogonal, and the Q of numpy's Householder QR of the same Gaussian matrix
    assert np.abs(Q.T @ Q - np.eye(N)).max() < 1e-13
    Qn, _ = np.linalg.qr(gauss_ref(N * N, 7).reshape((N, N), order="F"))
    assert np.abs(Q - Qn).max() < 1e-10
    MT = (N + NB - 1) // NB
    dT = dev.alloc(MT * (MT + 1) // 2 * NB * NB * 8)
    dev.check(dev.M.mi_dspd_from_orth_tiles(dev.h, N, NB, dQ, kappa, dT), "spd")
    T = dev.download(dT, (MT * (MT + 1) // 2 * NB * NB,))
    dev.free(dQ)
    dev.free(dT)
    A = po.symtiles_to_full(T, N, NB)
    lam = np.sort(np.linalg.eigvalsh(A))[::-1]
    exp = np.geomspace(1.0, 1.0 / kappa, N)
    assert np.abs(lam - exp).max() <= 1e-12
