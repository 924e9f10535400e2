"""GPU parity of the SpMM kernel families across K, leading dimensions,
alignment and storage precision, called through the C ABI (vbh_spmm,
vbh_spmm_fused) on torch-allocated device buffers, against the oracle.

K from 2 to 64 runs the lane-per-row family (every instantiated shape:
real KT 4/8/10/12/16 with 1, 2 or 4 lanes per row, complex KT 8 with up to 8
lanes); K = 1 and K > 64 run the row-loop families, except for precision 2
(FP32 values, FP64 vectors), which is lanes only.  Leading dimensions K,
K+1 (odd: scalar loads) and K rounded up to 16 bytes (vector loads) pick the
vector width.  FP64 tolerance 1e-12 relative, FP32 storage 2e-6."""
import ctypes

import numpy as np
import pytest

import oracle as O
from conftest import relerr

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-6


def lib():
    import vbcsr_amd

    path = vbcsr_amd.__file__.replace("__init__.py", "libvbcsr_hip.so")
    L = ctypes.CDLL(path)
    L.vbh_spmm.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                           ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.vbh_spmm_fused.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                 ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                 ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                 ctypes.c_void_p]
    L.vbh_last_error.restype = ctypes.c_char_p
    return L


def check(L, rc):
    assert rc == 0, L.vbh_last_error().decode()


def matrices():
    from vbcsr_amd import workloads as W

    rng = np.random.default_rng(17)
    g = W.graphene(12, 12, 4, seed=3)  # uniform 4x4 blocks, 10 per row
    bs = rng.choice(np.asarray([1, 4, 9, 13], np.int32), size=70).astype(np.int32)
    iu, ju, _ = W.random_geometric_pairs(70, 8.0, rng)
    v = W.symmetric_from_upper(bs, iu, ju, lambda gg, n, r, c, sel: gg.standard_normal((n, r, c)), rng)
    return {"bsr4": g, "vbcsr": v}


MATS = None


def get(name, dtype=np.float64):
    global MATS
    if MATS is None:
        MATS = matrices()
    A = MATS[name]
    if dtype == np.complex128:
        from vbcsr_amd.workloads import BlockCSR

        rng = np.random.default_rng(5)
        vals = A.vals + 1j * rng.standard_normal(A.vals.shape)
        A = BlockCSR(A.bs, A.row_ptr, A.col_ind, vals)
    from vbcsr_amd import VBCSR

    return A, VBCSR.from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)


def padded(ld_kind, K, elem):
    if ld_kind == "K":
        return K
    if ld_kind == "odd":
        return K + 1 if (K + 1) % 2 == 1 else K + 2
    per = max(1, 16 // elem)
    return (K + per - 1) // per * per


def dev(X, ld, torch_dtype):
    import torch

    n, K = X.shape
    buf = torch.zeros((max(n, 1), ld), dtype=torch_dtype, device="cuda")
    buf[:n, :K] = torch.from_numpy(np.ascontiguousarray(X)).to(torch_dtype)
    return buf


@pytest.mark.parametrize("name", ["bsr4", "vbcsr"])
@pytest.mark.parametrize("ld_kind", ["K", "odd", "vec"])
@pytest.mark.parametrize("K", [1, 2, 3, 5, 8, 10, 12, 13, 16, 17, 32, 33, 64, 65])
def test_spmm_real_fp64_fp32(name, ld_kind, K):
    import torch

    L = lib()
    A, H = get(name)
    m = O.Mat.of(A)
    rng = np.random.default_rng(K * 31 + len(ld_kind))
    X = rng.uniform(-1, 1, (A.n, K))
    ref = O.spmm(m, X)
    for prec, tdt, elem, tol in ((0, torch.float64, 8, FP64_TOL), (1, torch.float32, 4, FP32_TOL)):
        ld = padded(ld_kind, K, elem)
        Xd = dev(X, ld, tdt)
        Yd = torch.full((A.n, ld), 7.0, dtype=tdt, device="cuda")
        check(L, L.vbh_spmm(H._core.handle, K, Xd.data_ptr(), ld, Yd.data_ptr(), ld, prec, None))
        torch.cuda.synchronize()
        Y = Yd.cpu().numpy()
        assert relerr(Y[:, :K], ref) < tol, (name, ld_kind, K, prec)
        # columns past K belong to the caller and stay untouched
        assert np.all(Y[:, K:] == 7.0)


@pytest.mark.parametrize("name", ["bsr4", "vbcsr"])
@pytest.mark.parametrize("K", [2, 5, 8, 9, 16, 17, 40, 64])
def test_spmm_complex(name, K):
    import torch

    L = lib()
    A, H = get(name, np.complex128)
    m = O.Mat.of(A)
    rng = np.random.default_rng(K)
    X = rng.uniform(-1, 1, (A.n, K)) + 1j * rng.uniform(-1, 1, (A.n, K))
    ref = O.spmm(m, X)
    for prec, tdt, elem, tol in ((0, torch.complex128, 16, FP64_TOL), (1, torch.complex64, 8, FP32_TOL)):
        for ld in (K, padded("vec", K, elem)):
            Xd = dev(X, ld, tdt)
            Yd = torch.zeros((A.n, ld), dtype=tdt, device="cuda")
            check(L, L.vbh_spmm(H._core.handle, K, Xd.data_ptr(), ld, Yd.data_ptr(), ld, prec, None))
            torch.cuda.synchronize()
            assert relerr(Yd.cpu().numpy()[:, :K], ref) < tol, (name, K, prec, ld)


@pytest.mark.parametrize("name", ["bsr4", "vbcsr"])
@pytest.mark.parametrize("K", [1, 4, 10, 32])
@pytest.mark.parametrize("with_dots", [False, True])
def test_spmm_fused(name, K, with_dots):
    """Y = alpha AX + beta X + gamma Z with per-column dots <W,Y>, <X,Y>."""
    import torch

    L = lib()
    A, H = get(name)
    m = O.Mat.of(A)
    rng = np.random.default_rng(K + 100)
    X, Z, Wm = (rng.uniform(-1, 1, (A.n, K)) for _ in range(3))
    alpha, beta, gamma = 1.7, -0.3, -1.0
    Yr = alpha * O.spmm(m, X) + beta * X + gamma * Z
    for prec, tdt, tol in ((0, torch.float64, FP64_TOL), (1, torch.float32, FP32_TOL)):
        Xd, Zd, Wd = (dev(M_, K, tdt) for M_ in (X, Z, Wm))
        Yd = torch.zeros((A.n, K), dtype=tdt, device="cuda")
        dots = torch.zeros(2 * K, dtype=torch.float64, device="cuda")
        check(L, L.vbh_spmm_fused(H._core.handle, K, Xd.data_ptr(), K, Yd.data_ptr(), K, alpha, beta, gamma,
                                  Zd.data_ptr(), K, Wd.data_ptr(), K, dots.data_ptr() if with_dots else None, prec,
                                  None))
        torch.cuda.synchronize()
        Y = Yd.cpu().numpy().astype(np.float64)
        assert relerr(Y, Yr) < tol * 10, (name, K, prec)
        if with_dots:
            d = dots.cpu().numpy()
            # dots are of the stored values Y; X, W as stored
            Xs = Xd.cpu().numpy().astype(np.float64)
            Ws = Wd.cpu().numpy().astype(np.float64)
            assert np.allclose(d[:K], (Ws * Y).sum(0), rtol=1e-12, atol=1e-9)
            assert np.allclose(d[K:], (Xs * Y).sum(0), rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("name", ["bsr4", "vbcsr"])
@pytest.mark.parametrize("K", [1, 2, 10, 32, 64])
@pytest.mark.parametrize("cplx", [False, True])
def test_spmm_mixed_precision(name, K, cplx):
    """precision 2: FP32 matrix values against FP64 vectors equals the FP64
    product with the FP32-rounded matrix."""
    import torch

    L = lib()
    A, H = get(name, np.complex128 if cplx else np.float64)
    v = A.vals
    v32 = (v.real.astype(np.float32).astype(np.float64) + 1j * v.imag.astype(np.float32).astype(np.float64)
           if cplx else v.astype(np.float32).astype(np.float64))
    m = O.Mat(A.bs, A.row_ptr, A.col_ind, v32)
    rng = np.random.default_rng(K + 7)
    X = rng.uniform(-1, 1, (A.n, K))
    if cplx:
        X = X + 1j * rng.uniform(-1, 1, (A.n, K))
    ref = O.spmm(m, X)
    tdt = torch.complex128 if cplx else torch.float64
    for ld in (K, K + 3):
        Xd = dev(X, ld, tdt)
        Yd = torch.zeros((A.n, ld), dtype=tdt, device="cuda")
        check(L, L.vbh_spmm(H._core.handle, K, Xd.data_ptr(), ld, Yd.data_ptr(), ld, 2, None))
        torch.cuda.synchronize()
        assert relerr(Yd.cpu().numpy()[:, :K], ref) < FP64_TOL, (name, K, ld)
