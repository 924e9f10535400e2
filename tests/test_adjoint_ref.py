"""CPU checks of the dense torch reference used by the GPU adjoint tests: its
autograd gradient against central finite differences, the propagator against
scipy's expm, and unitarity."""
import numpy as np
import scipy.linalg
import torch

from adjoint_ref import dense, overlap_loss, propagate_dense, value_coordinates


def small_system(seed=0):
    from vbcsr_amd import workloads as W

    A = W.graphene(2, 2, 2, seed=seed, complex_blocks=True)
    rng = np.random.default_rng(seed)
    psi0 = rng.standard_normal((A.n, 1)) + 1j * rng.standard_normal((A.n, 1))
    psi0 /= np.linalg.norm(psi0)
    tgt = rng.standard_normal((A.n, 1)) + 1j * rng.standard_normal((A.n, 1))
    tgt /= np.linalg.norm(tgt)
    return A, psi0, tgt


def coefs(dt, center, half, nterms):
    from scipy.special import jv

    n = np.arange(nterms)
    return np.exp(-1j * center * dt) * (2.0 - (n == 0)) * (-1j) ** n * jv(n, half * dt)


def test_dense_propagator_is_expm_and_unitary():
    A, psi0, _ = small_system()
    D = dense(A)
    assert np.allclose(D, D.conj().T)
    half = float(np.abs(np.linalg.eigvalsh(D)).max()) * 1.01
    rows, cols = value_coordinates(A)
    c = coefs(0.5, 0.0, half, 40)
    out = propagate_dense(torch.as_tensor(A.vals), torch.as_tensor(rows), torch.as_tensor(cols), A.n,
                          torch.as_tensor(psi0), c, 0.0, half, 3).numpy()
    ref = np.linalg.matrix_power(scipy.linalg.expm(-1j * 0.5 * D), 3) @ psi0
    assert np.abs(out - ref).max() < 1e-12
    assert abs(np.linalg.norm(out) - 1.0) < 1e-12


def test_dense_reference_gradient_matches_finite_differences():
    A, psi0, tgt = small_system(1)
    half = float(np.abs(np.linalg.eigvalsh(dense(A))).max()) * 1.01
    rows, cols = (torch.as_tensor(x) for x in value_coordinates(A))
    c = coefs(0.5, 0.1, half, 24)
    psi0_t, tgt_t = torch.as_tensor(psi0), torch.as_tensor(tgt)

    def loss(v):
        return overlap_loss(propagate_dense(v, rows, cols, A.n, psi0_t, c, 0.1, half, 3), tgt_t)

    v = torch.as_tensor(A.vals).clone().requires_grad_(True)
    loss(v).backward()
    g = v.grad.numpy()
    rng = np.random.default_rng(2)
    eps = 1e-6
    for idx in rng.choice(A.nnz, 12, replace=False):
        for unit in (1.0, 1j):
            vp = torch.as_tensor(A.vals).clone()
            vm = vp.clone()
            vp[idx] += eps * unit
            vm[idx] -= eps * unit
            fd = (loss(vp) - loss(vm)).item() / (2 * eps)
            an = g[idx].real if unit == 1.0 else g[idx].imag
            assert abs(fd - an) < 1e-7 + 1e-6 * abs(an), (idx, unit, fd, an)
