"""GPU tests of the boundary's state handling:

* block norms through the C ABI against the reference's golden norms
  (BlockSpMat::get_block_norms, block_csr.hpp:2486-2502);
* the FP32 value copy follows every device-side write of the values
  (fill, copy_from, writes through values_ptr), so FP32 and mixed-precision
  applies never read stale values;
* calls on different streams from different threads (per-stream scratch and
  graph state, per-matrix lazily built structures) give the oracle's results,
  through the C ABI with explicit streams and through the Python API with the
  GIL released.
"""
import ctypes
import threading

import numpy as np
import pytest

import oracle as O
from conftest import relerr

pytestmark = pytest.mark.gpu


def lib():
    import vbcsr_amd

    L = ctypes.CDLL(vbcsr_amd.__file__.replace("__init__.py", "libvbcsr_hip.so"))
    L.vbh_last_error.restype = ctypes.c_char_p
    L.vbh_matrix_block_norms.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.vbh_spmm.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                           ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.vbh_kpm_moments.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                  ctypes.c_void_p, ctypes.c_void_p]
    L.vbh_stream_release.argtypes = [ctypes.c_void_p]
    L.vbh_stream_sync.argtypes = [ctypes.c_void_p]
    return L


def check(L, rc):
    assert rc == 0, L.vbh_last_error().decode()


def mk(A):
    from vbcsr_amd import VBCSR

    return VBCSR.from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)


def test_block_norms_golden(golden_real):
    import torch

    from vbcsr_amd import VBCSR

    d = golden_real
    H = VBCSR.from_block_csr(d["A_bs"], d["A_row_ptr"], d["A_col_ind"], d["A_vals"])
    L = lib()
    out = torch.full((d["A_col_ind"].shape[0],), -1.0, dtype=torch.float64, device="cuda")
    check(L, L.vbh_matrix_block_norms(H._core.handle, out.data_ptr(), None))
    torch.cuda.synchronize()
    assert relerr(out.cpu().numpy(), d["norms"]) < 1e-14


def _kpm32(H, V0, M=24):
    from vbcsr_amd import kpm as K

    return K.kpm_moments(H, V0, M, mode="one", precision="fp32"), K.kpm_moments(H, V0, M, mode="one",
                                                                                 precision="mixed")


def test_fp32_copy_follows_device_writes():
    import torch

    from vbcsr_amd import workloads as W
    from vbcsr_amd.workloads import BlockCSR

    A = W.graphene(6, 6, 4, seed=1)
    B = W.graphene(6, 6, 4, seed=2)
    H, H2 = mk(A), mk(B)
    V0 = H.create_multivector(3)
    V0.set_random(9, normalize=False)
    v0 = np.asarray(V0.to_numpy())

    def ref(vals):
        return O.kpm_moments(O.Mat.of(BlockCSR(A.bs, A.row_ptr, A.col_ind, vals)), v0, 24, mode=1)

    def check_both(vals):
        # fp32: FP32 values and vectors; mixed: FP32-rounded values, FP64 vectors,
        # so the oracle on the rounded values pins it to FP64 rounding
        f32, mixed = _kpm32(H, V0)
        assert relerr(f32, ref(vals)) < 1e-5
        assert relerr(mixed, ref(vals.astype(np.float32).astype(np.float64))) < 1e-11

    _kpm32(H, V0)  # builds the FP32 copy
    # fill
    H.fill(0.01)
    check_both(np.full_like(A.vals, 0.01))
    # copy_from
    H.copy_from(H2)
    check_both(B.vals)
    # a write through the device pointer, then mark_values_written
    core = H._core
    src = torch.as_tensor(A.vals, device="cuda")
    torch.cuda.synchronize()
    L = lib()
    L.vbh_memcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    check(L, L.vbh_memcpy(core.values_ptr, src.data_ptr(), 8 * A.vals.shape[0], 2, None))
    check(L, L.vbh_stream_sync(None))
    core.mark_values_written()
    check_both(A.vals)


def _fresh_pair():
    from vbcsr_amd import workloads as W

    rng = np.random.default_rng(41)
    bs = rng.choice(np.asarray([1, 4, 9, 13], np.int32), size=300).astype(np.int32)
    iu, ju, _ = W.random_geometric_pairs(300, 12.0, rng)
    A = W.symmetric_from_upper(bs, iu, ju, lambda g, n, r, c, sel: 0.05 * g.standard_normal((n, r, c)), rng)
    B = W.graphene(20, 20, 4, seed=6)
    return A, B


def test_threads_on_two_streams_c_abi():
    """Two host threads, two HIP streams, fresh matrices: SpMM (precision 0
    and 2) on one stream and KPM (FP32, precision 1) on the other, a few
    rounds each; then both threads on the same fresh matrix, so its FP32 copy
    and row-class lists are built while the other stream uses it."""
    import torch

    A, B = _fresh_pair()
    L = lib()
    rng = np.random.default_rng(3)
    XA = rng.uniform(-1, 1, (A.n, 32))
    ref_spmm = O.spmm(O.Mat.of(A), XA)
    V0B = np.where(rng.random((B.n, 8)) < 0.5, -1.0, 1.0)
    ref_kpm_b = O.kpm_moments(O.Mat.of(B), V0B, 40, mode=1)
    V0A = np.where(rng.random((A.n, 8)) < 0.5, -1.0, 1.0)
    # KPM needs the spectrum inside [-1, 1]: scale A by its dense bounds
    w = np.linalg.eigvalsh(A.to_dense())
    half = 1.01 * max(abs(w[0]), abs(w[-1]))
    ref_kpm_a = O.kpm_moments(O.Mat.of(A.scaled(1.0 / half)), V0A, 40, mode=1)

    for same in (False, True):
        HA, HB = mk(A), mk(B)
        ha, hb = HA._core.handle, HB._core.handle  # values on the device before the threads start
        sA, sB = torch.cuda.Stream(), torch.cuda.Stream()
        torch.cuda.synchronize()
        errors = []

        def spmm_worker():
            try:
                with torch.cuda.stream(sA):
                    X = torch.as_tensor(XA, device="cuda")
                    Y = torch.empty_like(X)
                    torch.cuda.current_stream().synchronize()
                    for prec in (0, 2, 0, 2):
                        check(L, L.vbh_spmm(ha, 32, X.data_ptr(), 32, Y.data_ptr(), 32, prec,
                                            ctypes.c_void_p(sA.cuda_stream)))
                        check(L, L.vbh_stream_sync(ctypes.c_void_p(sA.cuda_stream)))
                        tol = 1e-12 if prec == 0 else 1e-6
                        e = relerr(Y.cpu().numpy(), ref_spmm)
                        if e > tol:
                            errors.append(("spmm", prec, e))
            except Exception as ex:  # reported below
                errors.append(("spmm", repr(ex)))

        def kpm_worker():
            try:
                h, V0, ref, hw = (ha, V0A, ref_kpm_a, half) if same else (hb, V0B, ref_kpm_b, 1.0)
                with torch.cuda.stream(sB):
                    V = torch.as_tensor(V0, device="cuda")
                    torch.cuda.current_stream().synchronize()
                    mu = np.zeros(40)
                    for _ in range(3):
                        check(L, L.vbh_kpm_moments(h, V0.shape[1], V.data_ptr(), V0.shape[1], 40, 1,
                                                   0.0, hw, 1, mu.ctypes.data, ctypes.c_void_p(sB.cuda_stream)))
                        e = relerr(mu, ref)
                        if e > 1e-5:
                            errors.append(("kpm", e))
            except Exception as ex:
                errors.append(("kpm", repr(ex)))

        ts = [threading.Thread(target=spmm_worker), threading.Thread(target=kpm_worker)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        torch.cuda.synchronize()
        for s in (sA, sB):
            check(L, L.vbh_stream_release(ctypes.c_void_p(s.cuda_stream)))
        assert not errors, (same, errors)


def test_threads_python_api():
    """Two Python threads through the drop-in API (GIL released inside the
    calls): SpMM on one matrix, KPM on another, each against the oracle."""
    from vbcsr_amd import kpm as K

    A, B = _fresh_pair()
    HA, HB = mk(A), mk(B)
    HA._core.handle, HB._core.handle  # noqa: B018 (values on the device before the threads start)
    rng = np.random.default_rng(8)
    XA = rng.uniform(-1, 1, (A.n, 16))
    ref_a = O.spmm(O.Mat.of(A), XA)
    V0 = HB.create_multivector(6)
    V0.set_random(4, normalize=False)
    ref_b = O.kpm_moments(O.Mat.of(B), np.asarray(V0.to_numpy()), 30, mode=1)
    errors = []

    def w1():
        for _ in range(5):
            Y = HA.mult(XA).to_numpy()
            e = relerr(Y, ref_a)
            if e > 1e-12:
                errors.append(("spmm", e))

    def w2():
        for _ in range(5):
            e = relerr(K.kpm_moments(HB, V0, 30, mode="two"), ref_b)
            if e > 1e-11:
                errors.append(("kpm", e))

    ts = [threading.Thread(target=w1), threading.Thread(target=w2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
