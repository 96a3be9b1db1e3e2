"""ctypes view of liboracle.so -- the CPU restatement of DPLASMA's tile Cholesky
path (TEST INFRASTRUCTURE).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker or the timed CPU
baseline; the product library never loads it.  Each wrapper names the C
function it calls; the C sources cite the reference file:line they restate.
"""
import ctypes
import glob
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libdplasma_ref.so")

LOWER, UPPER, UPPERLOWER = 122, 121, 123
NOTRANS, TRANS = 111, 112
NONUNIT, UNIT = 131, 132
LEFT, RIGHT = 141, 142

_lib = None
_ref = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_dp1 = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="F_CONTIGUOUS")
_u64 = ctypes.c_uint64
_i = ctypes.c_int
_d = ctypes.c_double


def lib():
    """Load liboracle.so (built by oracle/Makefile)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("liboracle.so missing: run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "orc_rnd64_jump": (_u64, [_u64, _u64]),
        "orc_rnd_to_entry": (ctypes.c_float, [_u64]),
        "orc_dplgsy_tile": (None, [_d, _i, _i, _dp, _i, _i, _i, _i, _u64]),
        "orc_splgsy_tile": (None, [ctypes.c_float, _i, _i, _fp, _i, _i, _i, _i, _u64]),
        "orc_dplrnt_tile": (None, [_i, _i, _dp, _i, _i, _i, _i, _u64]),
        "orc_dplgsy_full": (None, [_d, _i, _dp, _i, _u64]),
        "orc_dplrnt_full": (None, [_i, _i, _dp, _i, _u64]),
        "orc_sym_tile_index": (ctypes.c_int64, [_i, _i, _i, _i]),
        "orc_dplgsy_symtiles": (None, [_d, _i, _i, _i, _dp1, _u64]),
        "orc_dplrnt_tiles": (None, [_i, _i, _i, _i, _dp1, _u64]),
        "orc_symtiles_to_full": (None, [_i, _i, _i, _dp1, _dp, _i]),
        "orc_full_to_symtiles": (None, [_i, _i, _i, _dp, _i, _dp1]),
        "orc_tiles_to_full": (None, [_i, _i, _i, _i, _dp1, _dp, _i]),
        "orc_full_to_tiles": (None, [_i, _i, _i, _i, _dp, _i, _dp1]),
        "orc_flops_dpotrf": (_d, [_d]),
        "orc_flops_dpotrs": (_d, [_d, _d]),
        "orc_flops_dgemm": (_d, [_d, _d, _d]),
        "orc_flops_dsyrk": (_d, [_d, _d]),
        "orc_flops_dtrsm": (_d, [_i, _d, _d]),
        "orc_dpotrf": (_i, [_i, _i, _dp, _i]),
        "orc_dtrsm": (None, [_i, _i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i]),
        "orc_dsyrk_ln": (None, [_i, _i, _d, _dp, _i, _d, _dp, _i]),
        "orc_dgemm": (None, [_i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i, _d, _dp, _i]),
        "orc_blas_load": (_i, [ctypes.c_char_p, ctypes.c_char_p]),
        "orc_blas_loaded": (_i, []),
        "orc_blas_set_threads": (None, [_i]),
        "orc_dag_dpotrf_L": (_i, [_i, _i, _dp1, _i, _i, _i]),
        "orc_seq_dpotrf_L": (_i, [_i, _i, _dp1, _i]),
        "orc_tile_dpotrs_L": (None, [_i, _i, _dp1, _i, _dp1]),
        "orc_tile_dtrsm_LL": (None, [_i, _i, _d, _i, _i, _dp1, _i, _dp1]),
        "orc_dlamch_eps": (_d, []),
        "orc_dlansy_inf": (_d, [_i, _i, _dp, _i]),
        "orc_dlange_inf": (_d, [_i, _i, _dp, _i]),
        "orc_check_dpotrf": (_d, [_i, _i, _dp, _i, _dp, _i, ctypes.POINTER(_d), ctypes.POINTER(_d)]),
        "orc_check_daxmb": (_d, [_i, _i, _i, _dp, _i, _dp, _i, _dp, _i]),
        "orc_check_dpoinv": (_d, [_i, _i, _dp, _i, _dp, _i, ctypes.POINTER(_d), ctypes.POINTER(_d),
                                  ctypes.POINTER(_d)]),
        "orc_wtime": (_d, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def ref():
    """The reference's header-only arithmetic built from its sources, or None."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB_PATH):
        R = ctypes.CDLL(REF_LIB_PATH)
        R.ref_rnd64_jump.restype = _u64
        R.ref_rnd64_jump.argtypes = [_u64, _u64]
        for n in ("ref_flops_dpotrf", "ref_flops_spotrf"):
            getattr(R, n).restype = _d
            getattr(R, n).argtypes = [_d]
        R.ref_flops_dpotrs.restype = _d
        R.ref_flops_dpotrs.argtypes = [_d, _d]
        R.ref_flops_dgemm.restype = _d
        R.ref_flops_dgemm.argtypes = [_d, _d, _d]
        R.ref_flops_dsyrk.restype = _d
        R.ref_flops_dsyrk.argtypes = [_d, _d]
        R.ref_flops_dtrsm.restype = _d
        R.ref_flops_dtrsm.argtypes = [_i, _d, _d]
        _ref = R
    return _ref


# --------------------------------------------------------------- helpers
def mt(n, nb):
    return (n + nb - 1) // nb


def n_sym_tiles(n, nb):
    t = mt(n, nb)
    return t * (t + 1) // 2


def dplgsy_full(N, bump, seed):
    A = np.zeros((N, N), dtype=np.float64, order="F")
    lib().orc_dplgsy_full(float(bump), N, A, N, seed)
    return A


def dplrnt_full(M, N, seed):
    A = np.zeros((M, N), dtype=np.float64, order="F")
    lib().orc_dplrnt_full(M, N, A, M, seed)
    return A


def dplgsy_symtiles(N, NB, bump, seed, uplo=LOWER):
    T = np.zeros(n_sym_tiles(N, NB) * NB * NB, dtype=np.float64)
    lib().orc_dplgsy_symtiles(float(bump), uplo, N, NB, T, seed)
    return T


def symtiles_to_full(T, N, NB, uplo=LOWER):
    A = np.zeros((N, N), dtype=np.float64, order="F")
    lib().orc_symtiles_to_full(uplo, N, NB, T, A, N)
    return A


def full_to_symtiles(A, NB, uplo=LOWER):
    N = A.shape[0]
    A = np.asfortranarray(A, dtype=np.float64)
    T = np.zeros(n_sym_tiles(N, NB) * NB * NB, dtype=np.float64)
    lib().orc_full_to_symtiles(uplo, N, NB, A, N, T)
    return T


def tiles_to_full(T, M, N, MB, NB):
    A = np.zeros((M, N), dtype=np.float64, order="F")
    lib().orc_tiles_to_full(M, N, MB, NB, T, A, M)
    return A


def full_to_tiles(A, MB, NB):
    M, N = A.shape
    A = np.asfortranarray(A, dtype=np.float64)
    T = np.zeros(mt(M, MB) * mt(N, NB) * MB * NB, dtype=np.float64)
    lib().orc_full_to_tiles(M, N, MB, NB, A, M, T)
    return T


def dpotrf_tiles(T, N, NB, nthreads=1, use_blas=False, pri_change=-1):
    """Tile-DAG dpotrf (lower) in place on T; returns info."""
    return lib().orc_dag_dpotrf_L(N, NB, T, nthreads, int(use_blas), pri_change)


def check_dpotrf(uplo, A, A0):
    N = A.shape[0]
    an, rn = _d(), _d()
    r = lib().orc_check_dpotrf(uplo, N, np.asfortranarray(A), N, np.asfortranarray(A0), N,
                               ctypes.byref(an), ctypes.byref(rn))
    return r


def check_dpoinv(uplo, A, Ainv):
    """orc_check_dpoinv: ||I - Ainv A||_1 / (||A||_1 ||Ainv||_1 N eps)."""
    N = A.shape[0]
    an, ain, rn = _d(), _d(), _d()
    return lib().orc_check_dpoinv(uplo, N, np.asfortranarray(A), N, np.asfortranarray(Ainv), N,
                                  ctypes.byref(an), ctypes.byref(ain), ctypes.byref(rn))


def check_daxmb(uplo, A0, b, x):
    N = A0.shape[0]
    nrhs = b.shape[1]
    return lib().orc_check_daxmb(uplo, N, nrhs, np.asfortranarray(A0), N,
                                 np.asfortranarray(b), N, np.asfortranarray(x), N)


def find_openblas():
    """Path of an optimized CBLAS/LAPACKE library, with its symbol prefix."""
    try:
        import scipy
        d = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
        libs = sorted(glob.glob(os.path.join(d, "libscipy_openblas*.so")))
        if libs:
            return libs[0], "scipy_"
    except Exception:
        pass
    return None, None


def load_openblas():
    path, prefix = find_openblas()
    if path is None:
        return None
    rc = lib().orc_blas_load(path.encode(), prefix.encode())
    return path if rc == 0 else None
