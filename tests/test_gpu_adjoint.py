"""GPU checks of the Chebyshev-propagation adjoint (vbcsr_amd.autograd):
the SDDMM, basis and Clenshaw kernels against dense numpy, and the gradient
with respect to the block values against a dense FP64 torch autograd gradient
(tests/adjoint_ref.py, itself checked against finite differences on the CPU)
and against central finite differences of the GPU forward.  Tolerances: 1e-10
relative for FP64, 2e-4 for complex64 against FP64."""
import ctypes

import numpy as np
import pytest

from adjoint_ref import dense, overlap_loss, propagate_dense, value_coordinates

pytestmark = pytest.mark.gpu


def lib():
    from vbcsr_amd import autograd as AG

    return AG._lib()


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def system(cells=3, seed=0):
    from vbcsr_amd import VBCSR
    from vbcsr_amd import workloads as W

    A = W.graphene(cells, cells, 4, seed=seed, complex_blocks=True)
    H = VBCSR.from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)
    return A, H


@pytest.mark.parametrize("cplx", [False, True])
def test_sddmm_vs_dense(cplx):
    import torch

    from vbcsr_amd import VBCSR
    from vbcsr_amd import workloads as W

    rng = np.random.default_rng(3)
    bs = rng.choice(np.asarray([1, 4, 9, 13], np.int32), size=30).astype(np.int32)
    iu, ju, _ = W.random_geometric_pairs(30, 6.0, rng)
    A = W.symmetric_from_upper(bs, iu, ju, lambda g, n, r, c, sel: g.standard_normal((n, r, c)), rng,
                               hermitian_complex=cplx)
    H = VBCSR.from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)
    Lq = 37
    U = rng.standard_normal((A.n, Lq))
    V = rng.standard_normal((A.n, Lq))
    if cplx:
        U = U + 1j * rng.standard_normal((A.n, Lq))
        V = V + 1j * rng.standard_normal((A.n, Lq))
    tdt = torch.complex128 if cplx else torch.float64
    Ud, Vd = (torch.as_tensor(M_, dtype=tdt, device="cuda") for M_ in (U, V))
    G0 = rng.standard_normal(A.nnz) + (1j * rng.standard_normal(A.nnz) if cplx else 0)
    Gd = torch.as_tensor(G0, dtype=tdt, device="cuda")
    alpha = np.array([0.7, -0.2 if cplx else 0.0])
    rc = lib().vbh_sddmm(ctypes.c_void_p(H._core.handle), Lq, ctypes.c_void_p(Ud.data_ptr()), Lq,
                         ctypes.c_void_p(Vd.data_ptr()), Lq, alpha.ctypes.data_as(ctypes.c_void_p),
                         ctypes.c_void_p(Gd.data_ptr()), 0, None)
    assert rc == 0
    torch.cuda.synchronize()
    rows, cols = value_coordinates(A)
    a = alpha[0] + 1j * alpha[1] if cplx else alpha[0]
    ref = G0 + a * (U[rows] * V[cols].conj()).sum(1)
    assert relerr(Gd.cpu().numpy(), ref) < 1e-12


def test_basis_and_clenshaw_vs_dense():
    import torch

    A, H = system(3)
    D = dense(A)
    half = float(np.abs(np.linalg.eigvalsh(D)).max()) * 1.05
    center, nt, K = 0.2, 12, 2
    Ht = (D - center * np.eye(A.n)) / half
    rng = np.random.default_rng(0)
    X = rng.standard_normal((A.n, K)) + 1j * rng.standard_normal((A.n, K))
    lam = rng.standard_normal((A.n, K)) + 1j * rng.standard_normal((A.n, K))
    coef = rng.standard_normal(nt) + 1j * rng.standard_normal(nt)
    L = lib()
    h = ctypes.c_void_p(H._core.handle)
    Xd = torch.as_tensor(X, device="cuda")
    Td = torch.zeros((A.n, nt * K), dtype=torch.complex128, device="cuda")
    assert L.vbh_cheb_basis(h, K, ctypes.c_void_p(Xd.data_ptr()), K, nt, ctypes.c_void_p(Td.data_ptr()), nt * K,
                            center, half, 0, None) == 0
    ld = torch.as_tensor(lam, device="cuda")
    Bd = torch.zeros((A.n, nt * K), dtype=torch.complex128, device="cuda")
    cf = np.ascontiguousarray(coef.astype(np.complex128).view(np.float64))
    assert L.vbh_cheb_adjoint(h, K, ctypes.c_void_p(ld.data_ptr()), K, nt, cf.ctypes.data_as(ctypes.c_void_p),
                              ctypes.c_void_p(Bd.data_ptr()), nt * K, center, half, 0, None) == 0
    torch.cuda.synchronize()
    T = Td.cpu().numpy()
    t = [X, Ht @ X]
    for n in range(2, nt):
        t.append(2 * Ht @ t[-1] - t[-2])
    for n in range(nt):
        assert relerr(T[:, n * K:(n + 1) * K], t[n]) < 1e-12, n
    # b_0 = sum_n conj(c_n) T_n(H~)^H lam
    u = [lam, Ht.conj().T @ lam]
    for n in range(2, nt):
        u.append(2 * Ht.conj().T @ u[-1] - u[-2])
    b0 = sum(np.conj(coef[n]) * u[n] for n in range(nt))
    assert relerr(Bd.cpu().numpy()[:, :K], b0) < 1e-12


def _setup(cells, seed, K=1):
    import torch

    A, H = system(cells, seed)
    D = dense(A)
    half = float(np.abs(np.linalg.eigvalsh(D)).max()) * 1.01
    rng = np.random.default_rng(seed + 10)
    psi0 = rng.standard_normal((A.n, K)) + 1j * rng.standard_normal((A.n, K))
    psi0 /= np.linalg.norm(psi0, axis=0)
    tgt = rng.standard_normal((A.n, K)) + 1j * rng.standard_normal((A.n, K))
    tgt /= np.linalg.norm(tgt, axis=0)
    vals = torch.as_tensor(A.vals, device="cuda").clone()
    return A, H, half, psi0, tgt, vals


@pytest.mark.parametrize("K", [1, 3])
def test_gradient_vs_dense_torch(K):
    import torch

    from vbcsr_amd import autograd as AG

    A, H, half, psi0, tgt, vals = _setup(3, 1, K)
    center, dt, nsteps, nt = 0.05, 0.5, 7, 20
    v = vals.clone().requires_grad_(True)
    p0 = torch.as_tensor(psi0, device="cuda").clone().requires_grad_(True)
    out = AG.propagate(H, v, p0, dt, nsteps, nt, center=center, half_width=half, checkpoint=3)
    assert abs(torch.linalg.norm(out[:, 0]).item() - 1.0) < 1e-12  # unitary
    loss = overlap_loss(out, torch.as_tensor(tgt, device="cuda"))
    loss.backward()
    rows, cols = (torch.as_tensor(x) for x in value_coordinates(A))
    coef = AG.propagator_coefficients(dt, center, half, nt)
    vr = torch.as_tensor(A.vals).clone().requires_grad_(True)
    pr = torch.as_tensor(psi0).clone().requires_grad_(True)
    ref_out = propagate_dense(vr, rows, cols, A.n, pr, coef, center, half, nsteps)
    assert relerr(out.detach().cpu().numpy(), ref_out.detach().numpy()) < 1e-11
    overlap_loss(ref_out, torch.as_tensor(tgt)).backward()
    assert relerr(v.grad.cpu().numpy(), vr.grad.numpy()) < 1e-10
    assert relerr(p0.grad.cpu().numpy(), pr.grad.numpy()) < 1e-10


def test_gradient_vs_finite_differences():
    import torch

    from vbcsr_amd import autograd as AG

    A, H, half, psi0, tgt, vals = _setup(3, 2)
    center, dt, nsteps, nt = 0.0, 0.5, 10, 24
    p0 = torch.as_tensor(psi0, device="cuda")
    tg = torch.as_tensor(tgt, device="cuda")

    def loss(v):
        return overlap_loss(AG.propagate(H, v, p0, dt, nsteps, nt, center=center, half_width=half, checkpoint=4),
                            tg)

    v = vals.clone().requires_grad_(True)
    loss(v).backward()
    g = v.grad.cpu().numpy()
    rng = np.random.default_rng(4)
    eps = 1e-6
    with torch.no_grad():
        for idx in rng.choice(A.nnz, 20, replace=False):
            for unit in (1.0, 1j):
                vp = vals.clone()
                vm = vals.clone()
                vp[idx] += eps * unit
                vm[idx] -= eps * unit
                fd = (loss(vp) - loss(vm)).item() / (2 * eps)
                an = g[idx].real if unit == 1.0 else g[idx].imag
                assert abs(fd - an) < 1e-7 + 1e-5 * abs(an), (idx, unit, fd, an)


def test_gradient_complex64_vs_fp64():
    import torch

    from vbcsr_amd import autograd as AG

    A, H, half, psi0, tgt, vals = _setup(4, 3)
    center, dt, nsteps, nt = 0.0, 0.5, 10, 24
    tg = torch.as_tensor(tgt, device="cuda")
    grads = {}
    for prec, cdt in (("fp64", torch.complex128), ("fp32", torch.complex64)):
        v = vals.clone().requires_grad_(True)
        p0 = torch.as_tensor(psi0, device="cuda").to(cdt)
        out = AG.propagate(H, v, p0, dt, nsteps, nt, center=center, half_width=half, checkpoint=5,
                           precision=prec)
        overlap_loss(out.to(torch.complex128), tg).backward()
        grads[prec] = v.grad.cpu().numpy()
    assert relerr(grads["fp32"], grads["fp64"]) < 2e-4
