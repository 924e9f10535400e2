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


def test_spectral_bounds_device_lanczos():
    """Device Lanczos bounds against dense eigenvalues (small H, enough steps
    to converge) and scipy's eigsh (larger H)."""
    import scipy.sparse.linalg as sla

    from vbcsr_amd import kpm as K
    from vbcsr_amd import workloads as W

    Hs = W.batched_hamiltonians(count=2, nb_lo=30, nb_hi=40, neighbours=8.0, seed=21, scale=False)
    for A in Hs:
        w = np.linalg.eigvalsh(A.to_dense())
        lo, hi = K.spectral_bounds(mk(A), num_iter=A.n)
        assert abs(lo - w[0]) < 1e-8 * np.abs(w).max() and abs(hi - w[-1]) < 1e-8 * np.abs(w).max()
    A = W.batched_hamiltonians(count=1, nb_lo=600, nb_hi=601, seed=22, scale=False)[0]
    S = A.to_scipy()
    e_hi = sla.eigsh(S, k=1, which="LA", tol=1e-10, return_eigenvectors=False)[0]
    e_lo = sla.eigsh(S, k=1, which="SA", tol=1e-10, return_eigenvectors=False)[0]
    lo, hi = K.spectral_bounds(mk(A), num_iter=160)
    width = e_hi - e_lo
    # Ritz values converge from inside the spectrum
    assert e_lo - 1e-9 * width <= lo < e_lo + 1e-3 * width
    assert e_hi - 1e-3 * width < hi <= e_hi + 1e-9 * width


def test_batched_c4_dos_bar_spectrum_scaled():
    """Batched uneven KPM as configured: random-geometric Hamiltonians with
    blocks {1,4,9,13}, ~20 neighbours, each spectrum scaled to [-0.99, 0.99]
    by its own Lanczos bounds, 512 moments, 32 Rademacher vectors each.  The
    FP32-values / FP64-vectors recursion ("mixed") meets the DOS-error bar
    1e-5 against the plain FP64 path on every Hamiltonian; the all-FP32
    recursion is recorded against a looser bound (its error is set by the
    FP32 rounding of the vectors, largest on the smallest Hamiltonians)."""
    from vbcsr_amd import kpm as K
    from vbcsr_amd import workloads as W

    Hs = W.batched_hamiltonians(count=8, nb_lo=200, nb_hi=1000, seed=31, scale=False)
    mats = [mk(A) for A in Hs]
    for H in mats:
        K.scale_to_unit(H, 0.99)
    # the scaled spectra sit at the configured edges
    for H in mats[:3]:
        lo, hi = K.spectral_bounds(H)
        assert -0.991 < lo < -0.985 and 0.985 < hi < 0.991, (lo, hi)
    V0s = []
    for h, H in enumerate(mats):
        v = H.create_multivector(32)
        v.set_random(1000 + h, normalize=False)
        V0s.append(v)
    M = 512
    mu64 = [K.kpm_moments(H, v, M, mode="plain") for H, v in zip(mats, V0s)]
    mub = K.kpm_moments_batched(mats, V0s, M, mode="two", precision="mixed")
    mu32 = K.kpm_moments_batched(mats, V0s, M, mode="two", precision="fp32")
    worst_mixed = worst_fp32 = 0.0
    for h in range(len(mats)):
        _, r64 = K.dos_from_moments(mu64[h])
        worst_mixed = max(worst_mixed, K.dos_error(K.dos_from_moments(mub[h])[1], r64))
        worst_fp32 = max(worst_fp32, K.dos_error(K.dos_from_moments(mu32[h])[1], r64))
    print(f"C4 DOS error vs FP64: mixed {worst_mixed:.2e}, fp32 {worst_fp32:.2e}")
    assert worst_mixed <= 1e-5
    assert worst_fp32 <= 2e-4


def test_c2_fp32_dos_error_full_lattice():
    """Graphene 200 x 200 cells, 4 orbitals (N = 320,000), spectrum scaled to
    [-0.99, 0.99], 200 moments, 10 Rademacher vectors: the FP32 recursion's
    DOS error against FP64 stays within 1e-5 at the configured size
    (examples/ex08_graphene_dos.py:145-155 reconstruction, interior nodes)."""
    from vbcsr_amd import kpm as K
    from vbcsr_amd import workloads as W

    A = W.graphene(200, 200, 4, seed=0, scale=False)
    H = mk(A)
    K.scale_to_unit(H, 0.99)
    V0 = H.create_multivector(10)
    V0.set_random(42, normalize=False)
    mu64 = K.kpm_moments(H, V0, 200, mode="plain")
    mu32 = K.kpm_moments(H, V0, 200, mode="two", precision="fp32")
    _, r64 = K.dos_from_moments(mu64)
    err = K.dos_error(K.dos_from_moments(mu32)[1], r64)
    print(f"C2 full lattice FP32 DOS error {err:.2e}")
    assert err <= 1e-5


def test_large_fp64_quads_four_slot_batches():
    """FP64 4x4-block launches with one partition per row (at least 2^19
    scalar rows) run the quad slot loop with 4-slot batches: graphene 300 x
    300 cells (N = 720,000), one SpMM with K = 10 against the oracle's SpMM,
    and 8 fused moments (two per matvec) against the oracle's recursion."""
    from vbcsr_amd import kpm as K
    from vbcsr_amd import workloads as W

    A = W.graphene(300, 300, 4, seed=4)
    assert A.n >= 1 << 19
    H = mk(A)
    m = O.Mat.of(A)
    X = np.random.default_rng(5).standard_normal((A.n, 10))
    Xv = H.create_multivector(10)
    Xv.from_numpy(X)
    Y = H.mult(Xv)
    assert relerr(np.asarray(Y.to_numpy()), O.spmm(m, X)) < 1e-12
    V0 = H.create_multivector(10)
    V0.set_random(3, normalize=False)
    ref = O.kpm_moments(m, np.asarray(V0.to_numpy()), 8, mode=1)
    assert relerr(K.kpm_moments(H, V0, 8, mode="two"), ref) < 1e-11
