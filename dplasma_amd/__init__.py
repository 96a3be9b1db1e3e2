"""Python mirror of DPLASMA's tile Cholesky interface on one MI355X.

The product is the C-ABI pair ``lib/libdplasma.so`` (``dplasma_*`` entry
points and the PaRSEC-type shim, ``include/dplasma.h``) over
``lib/libdplasma_mi.so`` (hand-written gfx950 kernels, ``include/dplasma_mi.h``).
This module only binds them with ctypes so that tests and the benchmark can
drive the same calls a C caller (tests/testing_zpotrf.c) makes:

    ctx = Context()                       # parsec_init
    A = SymMatrix(ctx, N, NB, LOWER)      # parsec_matrix_sym_block_cyclic_init + data_allocate
    ctx.dplgsy(float(N), LOWER, A, 3872)  # dplasma_dplgsy
    info = ctx.dpotrf(LOWER, A)           # dplasma_dpotrf

There is no fallback: a missing library or GPU raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
LIB_MI = os.path.join(LIB_DIR, "libdplasma_mi.so")
LIB_DPL = os.path.join(LIB_DIR, "libdplasma.so")
INCLUDE_DIR = os.path.join(os.path.dirname(HERE), "include")

# dplasma enums (src/include/dplasma/constants.h:78-96)
NOTRANS, TRANS, CONJTRANS = 111, 112, 113
UPPER, LOWER, UPPERLOWER = 121, 122, 123
NONUNIT, UNIT = 131, 132
LEFT, RIGHT = 141, 142
# mixed-precision factor of dplasma_dposv_mixed (dplasma.h extension)
MIXED_FP32, MIXED_BF16X3 = 0, 1

MATRIX_DOUBLE = 3
MATRIX_LAPACK, MATRIX_TILE = 0, 1

_mi = None
_dpl = None


class DataCollection(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_int), ("myrank", ctypes.c_int), ("key", ctypes.c_char_p)]


class TiledMatrix(ctypes.Structure):
    _fields_ = [("super", DataCollection)] + [(n, ctypes.c_int) for n in (
        "mtype", "storage", "dtype", "mb", "nb", "bsiz", "lm", "ln", "lmt", "lnt", "i", "j",
        "m", "n", "mt", "nt", "llm", "lln", "slm", "sln", "nb_local_tiles")]


class Grid(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "rank", "rows", "cols", "krows", "kcols", "ip", "jq", "rrank", "crank")]


class BlockCyclic(ctypes.Structure):
    _fields_ = [("super", TiledMatrix), ("grid", Grid), ("mat", ctypes.c_void_p)]


class SymBlockCyclic(ctypes.Structure):
    _fields_ = [("super", TiledMatrix), ("grid", Grid), ("mat", ctypes.c_void_p),
                ("uplo", ctypes.c_int)]


class Timing(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("stage_in_ms", ctypes.c_double),
                ("stage_out_ms", ctypes.c_double)]


class CheckNorms(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("Anorm", "Rnorm", "Bnorm", "Xnorm", "Ainvnorm", "result")]


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


def libs():
    """Load (once) and return (libdplasma_mi, libdplasma)."""
    global _mi, _dpl
    if _dpl is not None:
        return _mi, _dpl
    for p in (LIB_MI, LIB_DPL):
        if not os.path.exists(p):
            raise RuntimeError(f"{p} is missing: build it with __graft_entry__.build()")
    _mi = ctypes.CDLL(LIB_MI, mode=ctypes.RTLD_GLOBAL)
    _dpl = ctypes.CDLL(LIB_DPL, mode=ctypes.RTLD_GLOBAL)
    vp, i, d, u64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_uint64
    pi, pd = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    sz = ctypes.c_size_t
    M = _mi
    _sig(M, "mi_ctx_create", i, [i, ctypes.POINTER(vp)])
    _sig(M, "mi_ctx_destroy", None, [vp])
    _sig(M, "mi_device_count", i, [pi])
    _sig(M, "mi_last_error", ctypes.c_char_p, [])
    _sig(M, "mi_synchronize", i, [vp])
    _sig(M, "mi_malloc", i, [vp, ctypes.POINTER(vp), sz])
    _sig(M, "mi_free", i, [vp, vp])
    _sig(M, "mi_host_alloc", i, [ctypes.POINTER(vp), sz])
    _sig(M, "mi_host_free", i, [vp])
    for n in ("mi_memcpy_h2d", "mi_memcpy_d2h", "mi_memcpy_d2d"):
        _sig(M, n, i, [vp, vp, vp, sz])
    _sig(M, "mi_memcpy2d_h2d", i, [vp, vp, sz, vp, sz, sz, sz])
    _sig(M, "mi_memcpy2d_d2h", i, [vp, vp, sz, vp, sz, sz, sz])
    _sig(M, "mi_memset", i, [vp, vp, i, sz])
    _sig(M, "mi_dcopy", i, [vp, sz, vp, vp])
    _sig(M, "mi_dsym_upper_to_lower", i, [vp, i, i, vp, vp])
    _sig(M, "mi_dsym_lower_to_upper", i, [vp, i, i, vp, vp])
    _sig(M, "mi_dtiles_to_colmajor", i, [vp, i, i, i, i, vp, vp, i])
    _sig(M, "mi_dcolmajor_to_tiles", i, [vp, i, i, i, i, vp, i, vp])
    _sig(M, "mi_dplgsy_tiles", i, [vp, d, i, i, vp, u64])
    _sig(M, "mi_dplrnt", i, [vp, i, i, vp, i, i, i, i, u64])
    _sig(M, "mi_dadd_diag", i, [vp, i, vp, i, d])
    _sig(M, "mi_dpotrf_tiles", i, [vp, i, i, vp, pi])
    _sig(M, "mi_dpotrf_tiles_noalloc", i, [vp, i, i, vp, pi])
    _sig(M, "mi_dpotrf_tiles_rec", i, [vp, i, i, vp, i, pi])
    _sig(M, "mi_dtd_create", i, [vp, i, i, vp, ctypes.POINTER(vp)])
    _sig(M, "mi_dtd_potrf", i, [vp, i])
    _sig(M, "mi_dtd_trsm", i, [vp, i, i])
    _sig(M, "mi_dtd_syrk", i, [vp, i, i])
    _sig(M, "mi_dtd_gemm", i, [vp, i, i, i])
    _sig(M, "mi_dtd_wait", i, [vp, pi, ctypes.POINTER(ctypes.c_long)])
    _sig(M, "mi_dtd_destroy", None, [vp])
    _sig(M, "mi_dtrsm_LL", i, [vp, i, i, d, i, i, vp, i, vp, i])
    _sig(M, "mi_dpotrs_tiles", i, [vp, i, i, vp, i, vp, i])
    _sig(M, "mi_dpotrf_lapack", i, [vp, i, i, vp, i, i, pi])
    _sig(M, "mi_dpotrs_lapack", i, [vp, i, i, i, vp, i, vp, i, i])
    _sig(M, "mi_dposv_lapack", i, [vp, i, i, i, vp, i, vp, i, i, pi])
    _sig(M, "mi_dsym_lapack_to_tiles", i, [vp, i, i, vp, i, i, vp])
    _sig(M, "mi_dsym_tiles_to_lapack", i, [vp, i, i, i, vp, vp, i])
    _sig(M, "mi_spotrf_tiles", i, [vp, i, i, vp, pi])
    _sig(M, "mi_spotrf_step_potrf", i, [vp, i, i, vp, i, pi])
    _sig(M, "mi_dsposv_tiles", i, [vp, i, i, vp, i, vp, i, vp, i, i, pi, pi])
    _sig(M, "mi_dgauss", i, [vp, ctypes.c_int64, vp, u64])
    _sig(M, "mi_dorth_haar", i, [vp, i, u64, vp])
    _sig(M, "mi_dspd_from_orth_tiles", i, [vp, i, i, vp, d, vp])
    _sig(M, "mi_dsymm_resid", i, [vp, i, i, vp, i, vp, i, vp, i, vp, i])
    _sig(M, "mi_dlag2s_tiles", i, [vp, ctypes.c_int64, vp, vp, pi])
    _sig(M, "mi_spotrf_step_trsm", i, [vp, i, i, vp, i])
    _sig(M, "mi_spotrf_step_update", i, [vp, i, i, vp, i, i, i])
    _sig(M, "mi_strsm_LL", i, [vp, i, i, ctypes.c_float, i, i, vp, i, vp, i])
    _sig(M, "mi_spotrs_tiles", i, [vp, i, i, vp, i, vp, i])
    _sig(M, "mi_dpotrf_adjoint_tiles", i, [vp, i, i, vp, vp])
    _sig(M, "mi_dpotri_tiles", i, [vp, i, i, vp, pi])
    _sig(M, "mi_dcheck_poinv", i, [vp, i, i, vp, vp, pd, pd, pd])
    _sig(M, "mi_batch_sizes", i, [u64, i, i, i, pi])
    _sig(M, "mi_batch_create", i, [vp, i, pi, ctypes.POINTER(vp)])
    _sig(M, "mi_dbatch_spd", i, [vp, i, pi, u64, vp])
    _sig(M, "mi_batch_destroy", i, [vp])
    _sig(M, "mi_batch_extent", i, [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)])
    _sig(M, "mi_batch_set_shape", i, [vp, i, i])
    _sig(M, "mi_batch_get_shape", i, [vp, pi, pi])
    _sig(M, "mi_dbatch_autotune", i, [vp, vp, vp, vp, vp, pi])
    _sig(M, "mi_sbatch_autotune", i, [vp, vp, vp, vp, vp, pi])
    _sig(M, "mi_dposv_batched", i, [vp, vp, vp, i])
    _sig(M, "mi_sposv_batched", i, [vp, vp, vp, i])
    _sig(M, "mi_batch_info", i, [vp, pi])
    _sig(M, "mi_dposv_batched_loop", i, [vp, vp, vp])
    _sig(M, "mi_sposv_batched_loop", i, [vp, vp, vp])
    _sig(M, "mi_batch_last_ms", i, [vp, pd])
    _sig(M, "mi_spotrf_adjoint_tiles", i, [vp, i, i, vp, vp])
    _sig(M, "mi_dpotrf_step_potrf", i, [vp, i, i, vp, i, pi])
    _sig(M, "mi_dpotrf_step_trsm", i, [vp, i, i, vp, i])
    _sig(M, "mi_dpotrf_step_update", i, [vp, i, i, vp, i, i, i])
    _sig(M, "mi_set_update_engine", i, [i])
    _sig(M, "mi_set_panel_engine", i, [i])
    _sig(M, "mi_set_spotrf_update", i, [i])
    _sig(M, "mi_set_bf16_panel_split", i, [i])
    _sig(M, "mi_set_symm_kernel", i, [i])
    _sig(M, "mi_set_solve_engine", i, [i])
    _sig(M, "mi_set_solve_wait_limit", i, [d])
    _sig(M, "mi_dcheck_potrf", i, [vp, i, i, vp, vp, pd, pd])
    _sig(M, "mi_dcheck_axmb", i, [vp, i, i, vp, i, vp, i, vp, i, pd, pd, pd, pd])
    _sig(M, "mi_prof_enable", i, [vp, i])
    _sig(M, "mi_prof_reset", i, [vp])
    _sig(M, "mi_prof_get", i, [vp, ctypes.c_char_p, pi, pd, pd, pd])
    _sig(M, "mi_prof_count", i, [vp])
    _sig(M, "mi_prof_busy_ms", i, [vp, ctypes.c_char_p, pd])
    _sig(M, "mi_prof_name", ctypes.c_char_p, [vp, i])
    _sig(M, "mi_last_span_ms", i, [vp, pd])

    D = _dpl
    _sig(D, "parsec_init", vp, [i, vp, vp])
    _sig(D, "parsec_fini", i, [ctypes.POINTER(vp)])
    _sig(D, "parsec_context_add_taskpool", i, [vp, vp])
    _sig(D, "parsec_context_start", i, [vp])
    _sig(D, "parsec_context_wait", i, [vp])
    _sig(D, "parsec_devices_release_memory", None, [])
    _sig(D, "parsec_matrix_block_cyclic_init", i, [vp] + [i] * 16)
    _sig(D, "parsec_matrix_sym_block_cyclic_init", i, [vp] + [i] * 13)
    _sig(D, "parsec_tiled_matrix_destroy", None, [vp])
    _sig(D, "parsec_data_allocate", vp, [sz])
    _sig(D, "parsec_data_free", None, [vp])
    _sig(D, "parsec_shim_tile_addr", vp, [vp, i, i])
    _sig(D, "parsec_shim_sizeof", sz, [i])
    _sig(D, "dplasma_dpotrf", i, [vp, i, vp])
    _sig(D, "dplasma_dpotrf_rec", i, [vp, i, vp, i])
    _sig(D, "dplasma_dpotrf_New", vp, [i, vp, pi])
    _sig(D, "dplasma_dpotrf_Destruct", None, [vp])
    _sig(D, "dplasma_dpotrf_setrecursive", None, [vp, i])
    _sig(D, "dplasma_dpotrs", i, [vp, i, vp, vp])
    _sig(D, "dplasma_dposv", i, [vp, i, vp, vp])
    _sig(D, "dplasma_dsposv", i, [vp, i, vp, vp, vp, pi])
    _sig(D, "dplasma_dposv_mixed", i, [vp, i, vp, vp, vp, i, pi])
    _sig(D, "dplasma_dpotrf_adjoint", i, [vp, i, vp, vp])
    _sig(D, "dplasma_dpotri", i, [vp, i, vp])
    _sig(D, "dplasma_dpoinv", i, [vp, i, vp])
    _sig(D, "check_dpoinv", i, [vp, i, i, vp, vp])
    _sig(D, "dplasma_dtrsm", i, [vp, i, i, i, i, d, vp, vp])
    _sig(D, "dplasma_dtrsm_New", vp, [i, i, i, i, d, vp, vp])
    _sig(D, "dplasma_dtrsm_Destruct", None, [vp])
    _sig(D, "dplasma_dplgsy", i, [vp, d, i, vp, ctypes.c_ulonglong])
    _sig(D, "dplasma_dplghe", i, [vp, d, i, vp, ctypes.c_ulonglong])
    _sig(D, "dplasma_dplrnt", i, [vp, i, vp, ctypes.c_ulonglong])
    _sig(D, "dplasma_dlacpy", i, [vp, i, vp, vp])
    _sig(D, "check_dpotrf", i, [vp, i, i, vp, vp])
    _sig(D, "check_daxmb", i, [vp, i, i, vp, vp, vp])
    _sig(D, "dplasma_advise_data_on_device", i, [vp, i, vp, vp, vp])
    _sig(D, "dplasma_data_sync_host", i, [vp, vp])
    _sig(D, "dplasma_data_release", i, [vp, vp])
    _sig(D, "dplasma_mi_last_timing", i, [vp, ctypes.POINTER(Timing)])
    _sig(D, "dplasma_mi_last_check", i, [vp, ctypes.POINTER(CheckNorms)])
    _sig(D, "dplasma_mi_context", vp, [vp])
    _sig(D, "parsec_dtd_taskpool_new", vp, [])
    _sig(D, "parsec_dtd_data_collection_init", i, [vp])
    _sig(D, "parsec_dtd_data_collection_fini", None, [vp])
    _sig(D, "parsec_shim_data_key", ctypes.c_uint64, [vp, i, i])
    _sig(D, "parsec_dtd_tile_of", vp, [vp, ctypes.c_uint64])
    _sig(D, "parsec_dtd_create_dpotrf_task_class", vp, [vp, i, i])
    _sig(D, "parsec_dtd_create_dtrsm_task_class", vp, [vp, i, i])
    _sig(D, "parsec_dtd_create_dsyrk_task_class", vp, [vp, i, i])
    _sig(D, "parsec_dtd_create_dgemm_task_class", vp, [vp, i, i])
    _sig(D, "parsec_dtd_task_class_release", None, [vp, vp])
    _sig(D, "parsec_dtd_data_flush_all", None, [vp, vp])
    _sig(D, "parsec_taskpool_wait", i, [vp])
    _sig(D, "parsec_taskpool_free", None, [vp])
    _sig(D, "parsec_shim_dtd_ntasks", ctypes.c_long, [vp])
    # variadic: (tp, tc, priority, device_type, (flag, pointer)..., 0)
    D.parsec_dtd_insert_task_with_task_class.restype = i
    return _mi, _dpl


def device_count():
    mi, _ = libs()
    n = ctypes.c_int(0)
    mi.mi_device_count(ctypes.byref(n))
    return n.value


def last_error():
    mi, _ = libs()
    return mi.mi_last_error().decode(errors="replace")


class Matrix:
    """A shim descriptor plus its host storage (parsec_data_allocate)."""

    def __init__(self, ctx, desc, nelem):
        self.ctx = ctx
        self.desc = desc
        _, D = libs()
        self.nelem = nelem
        ptr = D.parsec_data_allocate(nelem * 8)
        if not ptr:
            raise MemoryError("parsec_data_allocate failed")
        desc.mat = ptr
        self.array = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_double)),
                                           shape=(nelem,))
        self.array[:] = 0.0      # deterministic padding for comparisons

    @property
    def ptr(self):
        return ctypes.addressof(self.desc)

    @property
    def t(self):
        return self.desc.super

    def free(self):
        _, D = libs()
        if self.desc.mat:
            D.parsec_data_free(self.desc.mat)
            self.desc.mat = None
        D.parsec_tiled_matrix_destroy(self.ptr)
        self.array = None

    def __del__(self):
        try:
            if self.array is not None:
                self.free()
        except Exception:
            pass


class SymMatrix(Matrix):
    """parsec_matrix_sym_block_cyclic_t, tile storage of the uplo tiles."""

    def __init__(self, ctx, N, NB, uplo=LOWER):
        _, D = libs()
        desc = SymBlockCyclic()
        D.parsec_matrix_sym_block_cyclic_init(ctypes.byref(desc), MATRIX_DOUBLE, 0, NB, NB, N, N,
                                              0, 0, N, N, 1, 1, uplo)
        super().__init__(ctx, desc, desc.super.nb_local_tiles * desc.super.bsiz)
        self.N, self.NB, self.uplo = N, NB, uplo


class GeMatrix(Matrix):
    """parsec_matrix_block_cyclic_t, tile (MB x NB, ld MB) or LAPACK storage."""

    def __init__(self, ctx, M, N, MB, NB, storage=MATRIX_TILE):
        _, D = libs()
        desc = BlockCyclic()
        D.parsec_matrix_block_cyclic_init(ctypes.byref(desc), MATRIX_DOUBLE, storage, 0, MB, NB, M, N,
                                          0, 0, M, N, 1, 1, 1, 1, 0, 0)
        nelem = desc.super.nb_local_tiles * desc.super.bsiz if storage == MATRIX_TILE else M * N
        super().__init__(ctx, desc, max(nelem, 1))
        self.M, self.N, self.MB, self.NB, self.storage = M, N, MB, NB, storage

    def to_numpy(self):
        """Column-major M x N copy of the host storage."""
        M, N, MB, NB = self.M, self.N, self.MB, self.NB
        if self.storage == MATRIX_LAPACK:
            return self.array[:M * N].reshape((N, M)).T.copy()
        MT = (M + MB - 1) // MB
        out = np.zeros((M, N))
        for n in range(0, (N + NB - 1) // NB):
            for m in range(MT):
                tm = min(MB, M - m * MB)
                tn = min(NB, N - n * NB)
                off = (n * MT + m) * MB * NB
                t = self.array[off:off + MB * NB].reshape((NB, MB)).T
                out[m * MB:m * MB + tm, n * NB:n * NB + tn] = t[:tm, :tn]
        return out

    def from_numpy(self, A):
        M, N, MB, NB = self.M, self.N, self.MB, self.NB
        if self.storage == MATRIX_LAPACK:
            self.array[:M * N] = np.asarray(A, dtype=np.float64).T.reshape(-1)
            return
        MT = (M + MB - 1) // MB
        for n in range(0, (N + NB - 1) // NB):
            for m in range(MT):
                tm = min(MB, M - m * MB)
                tn = min(NB, N - n * NB)
                off = (n * MT + m) * MB * NB
                t = self.array[off:off + MB * NB].reshape((NB, MB))
                t.T[:tm, :tn] = A[m * MB:m * MB + tm, n * NB:n * NB + tn]


class Context:
    """parsec_context_t on one GPU."""

    def __init__(self, device=0):
        _, D = libs()
        os.environ.setdefault("PARSEC_MCA_device_hip_device", str(device))
        self.ptr = D.parsec_init(1, None, None)
        if not self.ptr:
            raise RuntimeError("parsec_init failed: no usable GPU (" + last_error() + ")")
        self.D = D

    def close(self):
        if self.ptr:
            p = ctypes.c_void_p(self.ptr)
            self.D.parsec_fini(ctypes.byref(p))
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- entry points (same argument order as the C API)
    def dpotrf(self, uplo, A):
        return self.D.dplasma_dpotrf(self.ptr, uplo, A.ptr)

    def dpotrf_rec(self, uplo, A, hmb):
        return self.D.dplasma_dpotrf_rec(self.ptr, uplo, A.ptr, hmb)

    def dpotrs(self, uplo, A, B):
        return self.D.dplasma_dpotrs(self.ptr, uplo, A.ptr, B.ptr)

    def dposv(self, uplo, A, B):
        return self.D.dplasma_dposv(self.ptr, uplo, A.ptr, B.ptr)

    def dposv_mixed(self, uplo, A, B, X, prec=MIXED_FP32):
        """Mixed-precision solve (dsposv semantics); returns (info, iter)."""
        it = ctypes.c_int(0)
        info = self.D.dplasma_dposv_mixed(self.ptr, uplo, A.ptr, B.ptr, X.ptr, prec, ctypes.byref(it))
        return info, it.value

    def dpotri(self, uplo, A):
        return self.D.dplasma_dpotri(self.ptr, uplo, A.ptr)

    def dpoinv(self, uplo, A):
        return self.D.dplasma_dpoinv(self.ptr, uplo, A.ptr)

    def check_dpoinv(self, loud, uplo, A, Ainv):
        return self.D.check_dpoinv(self.ptr, loud, uplo, A.ptr, Ainv.ptr)

    def dpotrf_adjoint(self, uplo, L, Lbar):
        """Lbar (df/dL) -> symmetric df/dA in place; returns 0 or an error code."""
        return self.D.dplasma_dpotrf_adjoint(self.ptr, uplo, L.ptr, Lbar.ptr)

    def dtrsm(self, side, uplo, trans, diag, alpha, A, B):
        return self.D.dplasma_dtrsm(self.ptr, side, uplo, trans, diag, float(alpha), A.ptr, B.ptr)

    def dplgsy(self, bump, uplo, A, seed):
        return self.D.dplasma_dplgsy(self.ptr, float(bump), uplo, A.ptr, seed)

    def dplrnt(self, diagdom, A, seed):
        return self.D.dplasma_dplrnt(self.ptr, diagdom, A.ptr, seed)

    def dlacpy(self, uplo, A, B):
        return self.D.dplasma_dlacpy(self.ptr, uplo, A.ptr, B.ptr)

    def check_dpotrf(self, loud, uplo, A, A0):
        return self.D.check_dpotrf(self.ptr, loud, uplo, A.ptr, A0.ptr)

    def check_daxmb(self, loud, uplo, A, b, x):
        return self.D.check_daxmb(self.ptr, loud, uplo, A.ptr, b.ptr, x.ptr)

    def advise(self, uplo, A):
        return self.D.dplasma_advise_data_on_device(self.ptr, uplo, A.ptr, None, None)

    def sync_host(self, A):
        return self.D.dplasma_data_sync_host(self.ptr, A.ptr)

    def release(self, A):
        return self.D.dplasma_data_release(self.ptr, A.ptr)

    def release_all(self):
        self.D.parsec_devices_release_memory()

    def last_timing(self):
        t = Timing()
        self.D.dplasma_mi_last_timing(self.ptr, ctypes.byref(t))
        return {"total_ms": t.total_ms, "stage_in_ms": t.stage_in_ms, "stage_out_ms": t.stage_out_ms}

    def last_check(self):
        """Norms and ratio behind the last check_* call on this context."""
        c = CheckNorms()
        self.D.dplasma_mi_last_check(self.ptr, ctypes.byref(c))
        return {n: getattr(c, n) for n, _ in CheckNorms._fields_}

    # -- kernel profiler of the underlying kernel context
    def prof_enable(self, on=True):
        mi, _ = libs()
        mi.mi_prof_enable(self.D.dplasma_mi_context(self.ptr), int(on))

    def prof_reset(self):
        mi, _ = libs()
        mi.mi_prof_reset(self.D.dplasma_mi_context(self.ptr))

    def prof_busy_ms(self, kernel=None):
        """Union of the launch intervals of `kernel` (all kernels if None), ms."""
        mi, _ = libs()
        ms = ctypes.c_double()
        mi.mi_prof_busy_ms(self.D.dplasma_mi_context(self.ptr), kernel.encode() if kernel else None,
                           ctypes.byref(ms))
        return ms.value

    def prof(self):
        """{kernel: (launches, total_ms, algorithmic flops, algorithmic bytes)}"""
        mi, _ = libs()
        h = self.D.dplasma_mi_context(self.ptr)
        out = {}
        for k in range(mi.mi_prof_count(h)):
            name = mi.mi_prof_name(h, k)
            n, ms, fl, by = ctypes.c_int(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
            mi.mi_prof_get(h, name, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(by))
            out[name.decode()] = (n.value, ms.value, fl.value, by.value)
        return out


class Device:
    """Direct use of the kernel C ABI (dplasma_mi.h) on device buffers."""

    def __init__(self, device=0):
        self.M, _ = libs()
        h = ctypes.c_void_p()
        rc = self.M.mi_ctx_create(device, ctypes.byref(h))
        if rc:
            raise RuntimeError("mi_ctx_create failed: " + last_error())
        self.h = h
        self._bufs = []

    def check(self, rc, what):
        if rc:
            raise RuntimeError(f"{what} failed ({rc}): {last_error()}")

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        self.check(self.M.mi_malloc(self.h, ctypes.byref(p), nbytes), "mi_malloc")
        return p.value

    def free(self, p):
        self.check(self.M.mi_free(self.h, p), "mi_free")

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        self.check(self.M.mi_memcpy_h2d(self.h, p, arr.ctypes.data, arr.nbytes), "h2d")
        return p

    def download(self, p, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        self.check(self.M.mi_memcpy_d2h(self.h, out.ctypes.data, p, out.nbytes), "d2h")
        return out

    def close(self):
        if self.h:
            self.M.mi_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
