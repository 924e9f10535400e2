"""CPU pins of the KPM density-of-states reconstruction and the Jackson kernel.

These define the DOS-error bar of the FP32 recursions, so each is checked
against a fact the mathematics fixes rather than against a retyped formula:

* Jackson g_0 = 1 and g_n -> 0 at the end of the expansion
  (vbcsr/utils.py:5-17, examples/ex08_graphene_dos.py:145-146);
* the Jackson-damped expansion of a delta function is non-negative and has
  unit weight (the defining property of the Jackson kernel); the undamped
  (Dirichlet) one is not non-negative -- a dropped or mis-indexed damping
  factor fails the first check;
* the Gauss-Chebyshev integral of the reconstructed DOS equals mu_0, i.e. the
  orbital count for unnormalised Rademacher vectors
  (tests/test_inp_nanocrystal_example.py:172-188,
  examples/inp_nanocrystal_dos/kpm.py:135-151);
* for a diagonal H the moments are mu_n = sum_i cos(n arccos eps_i)
  (tests/test_inp_nanocrystal_example.py:238-263), and the DOS built from
  them by the T_n recursion equals the cosine-form sum
  sum_i [g_0 + 2 sum_n g_n cos(n t_i) cos(n t)] / (pi sin t), x = cos t.
"""
import numpy as np

from vbcsr_amd.kpm import chebyshev_nodes, dos_error, dos_from_moments
from vbcsr_amd.utils import jackson_kernel


def test_jackson_endpoints():
    for M in (2, 16, 200, 512):
        g = jackson_kernel(M)
        assert g.shape == (M,)
        assert abs(g[0] - 1.0) < 1e-14
        assert np.all(np.diff(g) < 0), "Jackson damping must decrease with n"
        # g_M (one past the last kept moment) vanishes: the series ends smoothly
        n, q = M, np.pi / (M + 1)
        g_end = ((M - n + 1) * np.cos(q * n) + np.sin(q * n) / np.tan(q)) / (M + 1)
        assert abs(g_end) < 1e-12


def _delta_moments(x0, M):
    return np.cos(np.arange(M) * np.arccos(x0))


def test_jackson_delta_is_positive_with_unit_weight():
    K = 4000
    x = chebyshev_nodes(K)
    w = np.pi * np.sqrt(1.0 - x * x) / K  # Gauss-Chebyshev weights for int f(x) dx
    for M in (64, 200, 512):
        for x0 in (-0.97, -0.3, 0.0, 0.55, 0.9):
            _, rho = dos_from_moments(_delta_moments(x0, M), x=x)
            assert rho.min() > -1e-10 * rho.max(), (M, x0, rho.min())
            assert abs(np.sum(rho * w) - 1.0) < 1e-12
            # the peak sits at x0 (resolution ~ pi/M)
            assert abs(x[np.argmax(rho)] - x0) < 2 * np.pi / M
    # without damping the truncated series rings below zero
    _, rho_d = dos_from_moments(_delta_moments(0.3, 200), x=x, kernel="none")
    assert rho_d.min() < -1e-3 * rho_d.max()


def test_gauss_chebyshev_integral_is_mu0():
    rng = np.random.default_rng(5)
    for K in (100, 2000):
        x = chebyshev_nodes(K)
        w = np.pi * np.sqrt(1.0 - x * x) / K
        for M in (24, 200):
            mu = rng.standard_normal(M) * 0.1
            mu[0] = 1234.0  # orbital count for unnormalised +/-1 vectors
            _, rho = dos_from_moments(mu, x=x)
            # exact while M < 2K: sum_k T_n(x_k) = 0 for 0 < n < 2K
            assert abs(np.sum(rho * w) - 1234.0) < 1e-9


def test_diagonal_hamiltonian_dos_closed_form():
    rng = np.random.default_rng(11)
    eps = rng.uniform(-0.95, 0.95, 37)
    M = 128
    mu = np.array([np.sum(np.cos(n * np.arccos(eps))) for n in range(M)])
    x = chebyshev_nodes(500)
    _, rho = dos_from_moments(mu, x=x)
    g = jackson_kernel(M)
    t, ti = np.arccos(x), np.arccos(eps)
    n = np.arange(1, M)
    series = g[0] + 2.0 * np.einsum("n,in,kn->ik", g[1:], np.cos(np.outer(ti, n)), np.cos(np.outer(t, n)))
    ref = series.sum(0) / (np.pi * np.sin(t))
    assert np.max(np.abs(rho - ref)) < 1e-10 * np.max(np.abs(ref))


def test_dos_error_metric():
    a = np.array([1.0, 2.0, 4.0])
    assert dos_error(a, a) == 0.0
    assert abs(dos_error(a + np.array([0.0, 0.04, 0.0]), a) - 0.01) < 1e-15
