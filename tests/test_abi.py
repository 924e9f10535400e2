"""The C-ABI library and the host module load, and every symbol that
include/vbcsr_hip.h declares is exported (CPU only: no compute calls)."""
import ctypes
import os
import re

import numpy as np

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "vbcsr_hip.h")).read()
    return sorted(set(re.findall(r"VBH_API\s+[\w\s\*]+?\b(vbh_\w+)\s*\(", text)))


def test_header_declares_the_path():
    names = declared_symbols()
    for must in ("vbh_spmv", "vbh_spmm", "vbh_spmm_adjoint", "vbh_spmm_fused", "vbh_spgemm", "vbh_kpm_moments",
                 "vbh_kpm_moments_batched", "vbh_cheb_apply", "vbh_matrix_filter", "vbh_matrix_axpby",
                 "vbh_vec_bdot", "vbh_vec_set_random"):
        assert must in names, must


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "vbcsr_amd", "libvbcsr_hip.so"))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    lib.vbh_version.restype = ctypes.c_int
    assert lib.vbh_version() >= 100


def test_host_module_surface_matches_reference_names():
    import vbcsr_amd
    from vbcsr_amd import _core

    for cls in ("BlockSpMat_Double", "BlockSpMat_Complex", "DistVector_Double", "DistVector_Complex",
                "DistMultiVector_Double", "DistMultiVector_Complex", "DistGraph", "AssemblyMode"):
        assert hasattr(_core, cls), cls
    # method names of vbcsr/pybind_vbcsr.cpp:223-404 on the matrix class
    for meth in ("add_block", "assemble", "mult", "mult_dense", "mult_adjoint", "mult_dense_adjoint", "scale",
                 "shift", "add_diagonal", "axpy", "axpby", "copy_from", "fill", "duplicate", "spmm", "spmm_self",
                 "add", "transpose", "get_block", "get_values", "filter_blocks", "to_dense", "from_dense",
                 "conjugate", "row_ptr", "col_ind", "matrix_kind", "local_nnz", "local_block_nnz", "shape"):
        assert hasattr(_core.BlockSpMat_Double, meth), meth
    for meth in ("sync_ghosts", "reduce_ghosts", "set_constant", "scale", "axpy", "axpby", "pointwise_mult",
                 "pointwise_mult_vec", "bdot", "duplicate", "copy_from", "set_random", "local_rows",
                 "ghost_rows", "num_vectors"):
        assert hasattr(_core.DistMultiVector_Double, meth), meth
    for meth in ("dot", "local_size", "ghost_size", "full_size"):
        assert hasattr(_core.DistVector_Double, meth), meth
    # the Python API of vbcsr/matrix.py
    for meth in ("create_serial", "create_distributed", "create_distributed_flat", "create_random",
                 "create_vector", "create_multivector", "mult", "mult_adjoint", "spmm", "spmm_self",
                 "filter_blocks", "from_scipy", "to_scipy", "get_values", "duplicate", "axpby", "shift", "scale",
                 "__matmul__", "_matvec", "_matmat", "_rmatvec"):
        assert hasattr(vbcsr_amd.VBCSR, meth), meth


def test_graph_construction_host_only():
    from vbcsr_amd import _core

    g = _core.DistGraph(None)
    g.construct_serial(3, [2, 3, 1], [[2, 0], [1], [0, 2, 2]])
    assert list(g.adj_ptr) == [0, 2, 3, 5]
    assert list(np.asarray(g.adj_ind)) == [0, 2, 1, 0, 2]  # sorted, de-duplicated
    assert list(g.block_offsets) == [0, 2, 5, 6]
    assert g.owned_scalar_rows == 6 and g.rank == 0 and g.size == 1
    gl = _core.DistGraph(None)
    gl.construct_local([0, 1], [2, 2], [[0, 7], [1, 5]], [7, 5], [3, 1])
    assert list(gl.ghost_global_indices) == [5, 7]
    assert list(gl.block_sizes) == [2, 2, 1, 3]
    assert gl.local_scalar_cols == 8 and gl.owned_scalar_rows == 4
