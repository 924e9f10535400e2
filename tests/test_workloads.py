"""Synthetic workload generators: shapes, symmetry and neighbour counts (CPU)."""
import numpy as np

from vbcsr_amd import workloads as W


def test_config1_shape_and_symmetry():
    A, x = W.config1_spmv(nb=60, extra=6, seed=1)
    D = A.to_dense()
    assert np.abs(D - D.T).max() == 0.0
    assert A.bs.min() >= 4 and A.bs.max() <= 8
    assert x.shape == (A.n,) and np.abs(x).max() <= 1.0
    rows = np.diff(A.row_ptr)
    assert rows.min() >= 7  # diagonal + 6 distinct extra, before symmetrisation


def test_graphene_blocks_and_scaling():
    A = W.graphene(8, 6, 4, seed=2)
    assert A.nb == 2 * 8 * 6
    assert np.all(np.diff(A.row_ptr) == 10)  # on-site + 3 NN + 6 NNN
    w = np.linalg.eigvalsh(A.to_dense())
    assert w.max() <= 0.99 + 1e-9 and w.min() >= -0.99 - 1e-9
    assert max(abs(w.max()), abs(w.min())) > 0.97
    C = W.graphene(3, 4, 2, seed=3, complex_blocks=True, scale=False)
    Dc = C.to_dense()
    assert np.abs(Dc - Dc.conj().T).max() < 1e-15


def test_random_geometric_neighbours():
    rng = np.random.default_rng(0)
    iu, ju, r = W.random_geometric_pairs(4000, 30.0, rng)
    deg = 2 * np.sum(iu != ju) / 4000
    assert 27 < deg < 33
    P = W.mcweeny_start(nb=500, neighbours=20.0, seed=1)
    D = P.to_dense()
    assert np.abs(D - D.T).max() < 1e-15


def test_batched_sizes():
    Hs = W.batched_hamiltonians(count=4, nb_lo=50, nb_hi=120, neighbours=6.0, seed=5)
    for A in Hs:
        assert 50 <= A.nb <= 120
        assert set(np.unique(A.bs)).issubset({1, 4, 9, 13})
