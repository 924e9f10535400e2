"""vbh_kpm_moments_noalloc: the KPM entry point that allocates no device
memory (its work buffers live in a static device array) against the
allocating vbh_kpm_moments on the same inputs, through the C ABI.

Both run the same kernels in the same order, so the moments must be equal
bit for bit; they are also checked against the oracle.  Two calls in a row,
the second on a smaller input, show that the array's leftover contents from
the first call do not leak into the second.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from conftest import relerr

pytestmark = pytest.mark.gpu

VBH_ERR_ARG, VBH_ERR_STATE = 1, 4


def lib():
    import vbcsr_amd

    L = ctypes.CDLL(vbcsr_amd.__file__.replace("__init__.py", "libvbcsr_hip.so"))
    L.vbh_last_error.restype = ctypes.c_char_p
    sig = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
           ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    L.vbh_kpm_moments.argtypes = sig
    L.vbh_kpm_moments_noalloc.argtypes = sig
    return L


def mk(A):
    from vbcsr_amd import VBCSR

    return VBCSR.from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)


def run(L, fn, H, V, K, M, mode, half, prec):
    mu = np.full(M, np.nan)
    rc = fn(H._core.handle, K, V.data_ptr(), K, M, mode, 0.0, half, prec, mu.ctypes.data, None)
    assert rc == 0, L.vbh_last_error().decode()
    return mu


def cases():
    from vbcsr_amd import workloads as W

    rng = np.random.default_rng(5)
    big = W.graphene(30, 30, 4, seed=2)  # N = 7200, uniform 4x4 blocks
    bs = rng.choice(np.asarray([1, 4, 9, 13], np.int32), size=200).astype(np.int32)
    iu, ju, _ = W.random_geometric_pairs(200, 10.0, rng)
    small = W.symmetric_from_upper(bs, iu, ju, lambda g, n, r, c, sel: g.standard_normal((n, r, c)), rng)
    w = np.linalg.eigvalsh(small.to_dense())
    small = small.scaled(0.99 / max(abs(w[0]), abs(w[-1])))  # matrix-core rows (9, 13) and lane rows
    # (matrix, K, M, mode, precision, tolerance against the oracle)
    return [(big, 8, 64, 2, 0, 1e-11), (small, 5, 33, 1, 2, 1e-6)], [(big, 6, 40, 0, 1, 1e-5),
                                                                     (small, 3, 17, 2, 0, 1e-11)]


@pytest.mark.parametrize("order", [0, 1])
def test_noalloc_matches_allocating_entry_point(order):
    import torch

    L = lib()
    runs = []
    for A, K, M, mode, prec, tol in cases()[order]:
        H = mk(A)
        V0 = np.where(np.random.default_rng(K).random((A.n, K)) < 0.5, -1.0, 1.0)
        V = torch.as_tensor(V0, device="cuda")
        torch.cuda.synchronize()
        ref = run(L, L.vbh_kpm_moments, H, V, K, M, mode, 1.0, prec)  # also builds the lazy structures
        runs.append((A, H, V0, V, K, M, mode, prec, tol, ref))
    # two no-allocation calls in a row, the second on the smaller input
    got = [run(L, L.vbh_kpm_moments_noalloc, H, V, K, M, mode, 1.0, prec)
           for A, H, V0, V, K, M, mode, prec, tol, ref in runs]
    for g, (A, H, V0, V, K, M, mode, prec, tol, ref) in zip(got, runs):
        assert np.array_equal(g, ref), (A.n, K, M, mode, prec, np.abs(g - ref).max())
        Am = A if prec != 2 else A.__class__(A.bs, A.row_ptr, A.col_ind, A.vals.astype(np.float32).astype(np.float64))
        assert relerr(g, O.kpm_moments(O.Mat.of(Am), V0, M, mode=max(mode, 1))) < tol


def test_noalloc_errors():
    import torch

    from vbcsr_amd import workloads as W

    L = lib()
    A = W.graphene(12, 12, 4, seed=3)
    H = mk(A)
    K = 4
    V = torch.ones((A.n, K), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    mu = np.zeros(8)
    # a fresh matrix has no row-class lists yet: building them would allocate
    rc = L.vbh_kpm_moments_noalloc(H._core.handle, K, V.data_ptr(), K, 8, 2, 0.0, 1.0, 0, mu.ctypes.data, None)
    assert rc == VBH_ERR_STATE and b"not prepared" in L.vbh_last_error()
    run(L, L.vbh_kpm_moments, H, V, K, 8, 2, 1.0, 0)
    # FP32 copy not built yet
    rc = L.vbh_kpm_moments_noalloc(H._core.handle, K, V.data_ptr(), K, 8, 2, 0.0, 1.0, 1, mu.ctypes.data, None)
    assert rc == VBH_ERR_STATE and b"FP32" in L.vbh_last_error()
    # beyond the static array: rejected before any launch
    B = W.graphene(110, 110, 4, seed=4)  # N = 96,800: three FP64 64-column vectors need 149 MB
    HB = mk(B)
    VB = torch.ones((B.n, 64), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    run(L, L.vbh_kpm_moments, HB, VB, 64, 4, 2, 1.0, 0)
    mb = np.zeros(4)
    rc = L.vbh_kpm_moments_noalloc(HB._core.handle, 64, VB.data_ptr(), 64, 4, 2, 0.0, 1.0, 0, mb.ctypes.data, None)
    assert rc == VBH_ERR_ARG and b"static scratch" in L.vbh_last_error()
