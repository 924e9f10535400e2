"""GPU KPM recursion and Chebyshev series against the oracle, the reference's
golden moments (ex08 composition) and closed forms.  FP64 tolerance 1e-11
relative on the moments; the FP32-mixed recursion is bounded by its DOS error
(<= 1e-5 against FP64, the precision bar of the FP32 modes)."""
import numpy as np
import pytest

import oracle as O
from conftest import relerr

pytestmark = pytest.mark.gpu


def mk(A):
    from vbcsr_amd import VBCSR

    return VBCSR.from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)


@pytest.mark.parametrize("mode", ["plain", "one", "two"])
def test_golden_moments(golden_kpm, mode):
    from vbcsr_amd import VBCSR
    from vbcsr_amd import kpm as K

    d = golden_kpm
    H = VBCSR.from_block_csr(d["H_bs"], d["H_row_ptr"], d["H_col_ind"], d["H_vals"])
    V0 = H.create_multivector(d["V0"].shape[1])
    V0.from_numpy(d["V0"])
    mu = K.kpm_moments(H, V0, 24, mode=mode)
    assert relerr(mu, d["mu"]) < 1e-11


@pytest.mark.parametrize("M", [1, 2, 3, 7, 50])
def test_moment_counts_and_modes_agree(M):
    from vbcsr_amd import kpm as K
    from vbcsr_amd import workloads as W

    A = W.graphene(6, 5, 4, seed=2)
    H = mk(A)
    V0 = H.create_multivector(6)
    V0.set_random(11, normalize=False)
    ref = O.kpm_moments(O.Mat.of(A), np.asarray(V0.to_numpy()), M, mode=1)
    for mode in ("plain", "one", "two"):
        assert relerr(K.kpm_moments(H, V0, M, mode=mode), ref) < 1e-11, (mode, M)


def test_diagonal_closed_form():
    from vbcsr_amd import VBCSR
    from vbcsr_amd import kpm as K

    e = np.linspace(-0.97, 0.93, 41)
    H = VBCSR.create_serial(41, [1] * 41, [[i] for i in range(41)])
    H._set_values(e)
    V0 = H.create_multivector(3)
    V0.set_random(5, normalize=False)
    expect = np.array([3 * np.sum(np.cos(n * np.arccos(e))) for n in range(20)])
    for mode in ("one", "two"):
        assert np.abs(K.kpm_moments(H, V0, 20, mode=mode) - expect).max() < 2e-12
    # affine map inside the recursion
    H2 = VBCSR.create_serial(41, [1] * 41, [[i] for i in range(41)])
    H2._set_values(3.0 * e - 1.0)
    assert np.abs(K.kpm_moments(H2, V0, 20, mode="two", center=-1.0, half_width=3.0) - expect).max() < 1e-11


def test_fp32_mixed_dos_error():
    from vbcsr_amd import kpm as K
    from vbcsr_amd import workloads as W

    # the FP32 rounding errors of the moments average over N*K entries, so the
    # DOS error falls with the system size; 100x100 cells (N = 80,000) is a
    # quarter of the benchmark lattice
    A = W.graphene(100, 100, 4, seed=4)
    H = mk(A)
    V0 = H.create_multivector(10)
    V0.set_random(42, normalize=False)
    mu64 = K.kpm_moments(H, V0, 200, mode="two")
    mu32 = K.kpm_moments(H, V0, 200, mode="two", precision="fp32")
    _, r64 = K.dos_from_moments(mu64)
    _, r32 = K.dos_from_moments(mu32)
    assert K.dos_error(r32, r64) < 1e-5
    mu1 = K.kpm_moments(H, V0, 200, mode="one", precision="fp32")
    assert K.dos_error(K.dos_from_moments(mu1)[1], r64) < 1e-5


@pytest.mark.parametrize("graph", [False, True])
def test_batched_matches_single(graph):
    from vbcsr_amd import kpm as K
    from vbcsr_amd import workloads as W

    Hs = W.batched_hamiltonians(count=5, nb_lo=40, nb_hi=200, neighbours=8.0, seed=3)
    mats = [mk(A) for A in Hs]
    V0s = []
    for h, H in enumerate(mats):
        v = H.create_multivector(7)
        v.set_random(100 + h, normalize=False)
        V0s.append(v)
    for mode in ("one", "two"):
        mub = K.kpm_moments_batched(mats, V0s, 33, mode=mode, graph=graph)
        for h in range(len(mats)):
            ref = O.kpm_moments(O.Mat.of(Hs[h]), np.asarray(V0s[h].to_numpy()), 33, mode=1)
            assert relerr(mub[h], ref) < 1e-11, (mode, h)
    mu32 = K.kpm_moments_batched(mats, V0s, 33, mode="two", precision="fp32", graph=graph)
    for h in range(len(mats)):
        ref = O.kpm_moments(O.Mat.of(Hs[h]), np.asarray(V0s[h].to_numpy()), 33, mode=1)
        assert relerr(mu32[h], ref) < 1e-5


def rounded32(A):
    """The same matrix with its values rounded to FP32 (the value copy that
    precision "mixed" reads), kept in FP64 for the oracle."""
    from vbcsr_amd.workloads import BlockCSR

    v = A.vals
    if np.iscomplexobj(v):
        v32 = v.real.astype(np.float32).astype(np.float64) + 1j * v.imag.astype(np.float32).astype(np.float64)
    else:
        v32 = v.astype(np.float32).astype(np.float64)
    return BlockCSR(A.bs, A.row_ptr, A.col_ind, v32)


@pytest.mark.parametrize("K", [1, 7, 32])
def test_mixed_precision_single_and_batched(K):
    """precision "mixed": FP32 values against FP64 vectors -- equal, to FP64
    rounding, to the FP64 recursion on the FP32-rounded matrix."""
    from vbcsr_amd import kpm as KP
    from vbcsr_amd import workloads as W

    Hs = W.batched_hamiltonians(count=4, nb_lo=40, nb_hi=200, neighbours=8.0, seed=5)
    mats = [mk(A) for A in Hs]
    V0s = []
    for h, H in enumerate(mats):
        v = H.create_multivector(K)
        v.set_random(200 + h, normalize=False)
        V0s.append(v)
    refs = [O.kpm_moments(O.Mat.of(rounded32(A)), np.asarray(v.to_numpy()).reshape(A.n, K), 40, mode=1)
            for A, v in zip(Hs, V0s)]
    for mode in ("plain", "one", "two"):
        for h in range(len(mats)):
            mu = KP.kpm_moments(mats[h], V0s[h], 40, mode=mode, precision="mixed")
            assert relerr(mu, refs[h]) < 1e-11, (mode, h)
    for mode in ("one", "two"):
        mub = KP.kpm_moments_batched(mats, V0s, 40, mode=mode, precision="mixed")
        for h in range(len(mats)):
            assert relerr(mub[h], refs[h]) < 1e-11, (mode, h)


def test_kpm_apply_exp_closed_form():
    """tests/test_kpm.py:7-58 through the drop-in kpm_apply."""
    from vbcsr_amd import VBCSR
    from vbcsr_amd.utils import compute_kpm_coeffs, kpm_apply

    n = 50
    H = VBCSR.create_serial(n, [1] * n, [[i] for i in range(n)])
    for i in range(n):
        H.add_block(i, i, np.array([[(i + 1) * 0.1]]))
    H.assemble()
    c = compute_kpm_coeffs(200, lambda x: np.exp(x * 2.5 + 2.55))
    v = H.create_vector()
    v.set_constant(1.0)
    y = kpm_apply(H, v, c, 2.5, 2.55).to_numpy()
    assert np.max(np.abs(y - np.exp(np.arange(1, n + 1) * 0.1))) < 0.05


def test_cheb_apply_vs_oracle_complex():
    from scipy.special import jv

    from vbcsr_amd import workloads as W
    from vbcsr_amd.utils import kpm_apply

    A = W.graphene(5, 6, 4, seed=8, complex_blocks=True)
    H = mk(A)
    coef = np.array([(2 if n else 1) * (-1j) ** n * jv(n, 2.0) for n in range(30)])
    psi = W.wavepacket(5, 6, 4, 2.0, 1)
    v = H.create_vector()
    v.from_numpy(psi)
    y = kpm_apply(H, v, coef, 1.0, 0.0).to_numpy()
    ref = O.cheb_apply(O.Mat.of(A), psi, coef)
    assert relerr(y, ref) < 1e-12
    # e^{-iHt} is unitary: the norm is conserved to the truncation error
    assert abs(np.linalg.norm(y) - 1.0) < 1e-10
