"""The C-ABI libraries load without a GPU and export every function the
public headers declare; the ctypes mirrors of the descriptor structs match
the C layouts; host-only shim logic (descriptor geometry, tile addressing)
behaves as DPLASMA's callers expect."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set()
    for m in re.finditer(r"\b((?:mi|dplasma|parsec|check)_\w+)\s*\(", txt):
        names.add(m.group(1))
    return names


@pytest.mark.parametrize("header,lib", [("dplasma_mi.h", 0), ("dplasma.h", 1), ("parsec_shim.h", 1)])
def test_every_declared_symbol_is_exported(dp, header, lib):
    libs = dp.libs()
    names = declared(header)
    assert len(names) > 5
    missing = [n for n in sorted(names) if not hasattr(libs[lib], n)]
    assert not missing, missing


def test_struct_layouts_match(dp):
    _, D = dp.libs()
    assert D.parsec_shim_sizeof(0) == ctypes.sizeof(dp.TiledMatrix)
    assert D.parsec_shim_sizeof(1) == ctypes.sizeof(dp.BlockCyclic)
    assert D.parsec_shim_sizeof(2) == ctypes.sizeof(dp.SymBlockCyclic)


def test_sym_descriptor_geometry_and_tile_addresses(dp):
    _, D = dp.libs()
    desc = dp.SymBlockCyclic()
    N, NB = 378, 93
    D.parsec_matrix_sym_block_cyclic_init(ctypes.byref(desc), dp.MATRIX_DOUBLE, 0, NB, NB, N, N,
                                          0, 0, N, N, 1, 1, dp.LOWER)
    t = desc.super
    assert (t.mt, t.nt, t.lmt, t.bsiz) == (5, 5, 5, NB * NB)
    assert t.nb_local_tiles == 15 and t.super.nodes == 1
    desc.mat = 4096     # fake base address: only the arithmetic is checked
    addr = lambda m, n: D.parsec_shim_tile_addr(ctypes.byref(desc), m, n)
    # lower tiles column by column: (m,n) at n*MT - n(n-1)/2 + (m-n)
    for (m, n, idx) in [(0, 0, 0), (4, 0, 4), (1, 1, 5), (4, 1, 8), (2, 2, 9), (4, 4, 14)]:
        assert addr(m, n) == 4096 + idx * NB * NB * 8
    assert addr(0, 1) is None
    desc.uplo = dp.UPPER
    for (m, n, idx) in [(0, 0, 0), (0, 1, 1), (1, 1, 2), (0, 4, 10), (4, 4, 14)]:
        assert addr(m, n) == 4096 + idx * NB * NB * 8


def test_general_descriptor_geometry(dp):
    _, D = dp.libs()
    desc = dp.BlockCyclic()
    D.parsec_matrix_block_cyclic_init(ctypes.byref(desc), dp.MATRIX_DOUBLE, dp.MATRIX_TILE, 0, 64, 64,
                                      200, 3, 0, 0, 200, 3, 1, 1, 1, 1, 0, 0)
    t = desc.super
    assert (t.mt, t.nt, t.nb_local_tiles) == (4, 1, 4)
    desc.mat = 8
    assert D.parsec_shim_tile_addr(ctypes.byref(desc), 3, 0) == 8 + 3 * 64 * 64 * 8
    desc2 = dp.BlockCyclic()
    D.parsec_matrix_block_cyclic_init(ctypes.byref(desc2), dp.MATRIX_DOUBLE, dp.MATRIX_LAPACK, 0, 64, 64,
                                      200, 3, 0, 0, 200, 3, 1, 1, 1, 1, 0, 0)
    desc2.mat = 8
    # LAPACK storage: column-major with lda = llm
    assert D.parsec_shim_tile_addr(ctypes.byref(desc2), 2, 0) == 8 + 128 * 8


def test_no_gpu_is_reported_not_hidden(dp):
    if dp.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError):
        dp.Context(0)


def test_lapack_entries_reject_bad_arguments_before_touching_the_device(dp):
    """The LAPACK-layout entries validate uplo, N, lda, NB (and NRHS, ldb)
    first, as LAPACK's xerbla checks do, with no context or device needed."""
    M, _ = dp.libs()
    info = ctypes.c_int(5)
    assert M.mi_dpotrf_lapack(None, 0, 10, None, 10, 32, ctypes.byref(info)) == -2
    assert info.value == 0
    assert M.mi_dpotrf_lapack(None, 122, -1, None, 10, 32, ctypes.byref(info)) == -3
    assert M.mi_dpotrf_lapack(None, 121, 10, None, 9, 32, ctypes.byref(info)) == -5
    assert M.mi_dpotrf_lapack(None, 122, 10, None, 10, 0, ctypes.byref(info)) == -6
    assert M.mi_dpotrs_lapack(None, 122, 10, -1, None, 10, None, 10, 32) == -4
    assert M.mi_dpotrs_lapack(None, 122, 10, 1, None, 10, None, 9, 32) == -8
    assert M.mi_dposv_lapack(None, 123, 10, 1, None, 10, None, 10, 32, ctypes.byref(info)) == -2
    assert M.mi_dposv_lapack(None, 122, 10, 1, None, 10, None, 3, 32, ctypes.byref(info)) == -8
    assert M.mi_dsym_lapack_to_tiles(None, 122, 10, None, 4, 32, None) == -5
    assert "lda = 4" in dp.last_error()
    assert M.mi_dsym_tiles_to_lapack(None, 121, 10, 32, None, None, 4) == -5
    # N = 0 is a no-op that succeeds without a device
    assert M.mi_dpotrf_lapack(None, 122, 0, None, 1, 32, ctypes.byref(info)) == 0 and info.value == 0
    assert M.mi_dpotrs_lapack(None, 121, 0, 3, None, 1, None, 1, 32) == 0
