"""GPU Hermitian SpGEMM (spmm_hermitian = upper block triangle + Hermitian
completion) on the reference's own cases (tests/test_spmm_hermitian.py:23-68)
and against the oracle: FP64 values 1e-12 relative, patterns identical, and
the result exactly Hermitian (lower blocks are exact conjugate transposes and
the diagonal average is symmetric in its operands)."""
import numpy as np
import pytest

import oracle as O
from conftest import relerr
from test_oracle_hermitian import build

pytestmark = pytest.mark.gpu


def dev(m: O.Mat):
    from vbcsr_amd import VBCSR

    return VBCSR.from_block_csr(m.rbs, m.row_ptr, m.col_ind, m.vals)


def dense(C):
    return np.asarray(C.to_dense())


def test_congruence_matches_full_product():
    rng = np.random.default_rng(7)
    S = dev(build(6, 3, rng, True))
    Z = dev(build(6, 3, rng, False))
    SZ = S.spmm(Z, 0.0)
    full = dense(Z.spmm(SZ, 0.0, True))      # Z^H (S Z)
    half = dense(Z.spmm_hermitian(SZ, 0.0, True))
    assert np.abs(full - half).max() < 1e-12 * np.abs(full).max()
    assert np.abs(half - half.conj().T).max() == 0.0


def test_square_of_hermitian_matches():
    rng = np.random.default_rng(11)
    X = dev(build(5, 4, rng, True))
    full = dense(X.spmm(X, 0.0))
    half = dense(X.spmm_hermitian(X, 0.0))
    assert np.abs(full - half).max() < 1e-12 * np.abs(full).max()
    assert np.abs(half - half.conj().T).max() == 0.0


@pytest.mark.parametrize("thr", [0.0, 0.05, 0.5])
def test_sparse_upper_and_completion_vs_oracle(thr):
    from vbcsr_amd import workloads as W

    A = W.batched_hamiltonians(count=1, nb_lo=120, nb_hi=121, neighbours=8.0, seed=21)[0]
    m = O.Mat.of(A)
    H = dev(m)
    U = H.spmm_hermitian_upper(H, thr)
    Ur = O.spgemm(m, m, thr, upper=True)
    assert np.array_equal(U.row_ptr, Ur.row_ptr) and np.array_equal(U.col_ind, Ur.col_ind)
    assert relerr(U.get_values(), Ur.vals) < 1e-12
    C = H.spmm_hermitian(H, thr)
    Cr = O.complete_hermitian(Ur)
    assert np.array_equal(C.row_ptr, Cr.row_ptr) and np.array_equal(C.col_ind, Cr.col_ind)
    assert relerr(C.get_values(), Cr.vals) < 1e-12
    D = dense(C)
    assert np.abs(D - D.T).max() == 0.0
    # completion of the device upper triangle on its own
    C2 = U.complete_hermitian()
    assert np.array_equal(C2.get_values(), C.get_values())


def test_complex_sparse_vs_oracle():
    rng = np.random.default_rng(5)
    X = build(7, 3, rng, True)
    # drop some blocks symmetrically so the pattern is sparse
    keep = np.ones(X.nbr * X.nbr, bool)
    for i in range(X.nbr):
        for j in range(X.nbr):
            if (i + 2 * j) % 3 == 1 and i != j:
                keep[i * X.nbr + j] = keep[j * X.nbr + i] = False
    vo = X.voff
    rp = np.zeros(X.nbr + 1, np.int32)
    ci, vals = [], []
    for i in range(X.nbr):
        for j in range(X.nbr):
            k = i * X.nbr + j
            if keep[k]:
                ci.append(j)
                vals.append(X.vals[vo[k]:vo[k + 1]])
        rp[i + 1] = len(ci)
    Xs = O.Mat(X.rbs, rp, np.asarray(ci, np.int32), np.concatenate(vals))
    H = dev(Xs)
    for thr in (0.0, 2.0):
        C = H.spmm_hermitian(H, thr)
        Cr = O.complete_hermitian(O.spgemm(Xs, Xs, thr, upper=True))
        assert np.array_equal(C.row_ptr, Cr.row_ptr) and np.array_equal(C.col_ind, Cr.col_ind)
        assert relerr(C.get_values(), Cr.vals) < 1e-12
        D = dense(C)
        assert np.abs(D - D.conj().T).max() == 0.0


@pytest.mark.parametrize("numeric,tol", [("fp32", 1e-5), ("bf16", 2e-2)])
def test_reduced_precision_numeric(numeric, tol):
    from vbcsr_amd import workloads as W

    A = W.batched_hamiltonians(count=1, nb_lo=80, nb_hi=81, neighbours=8.0, seed=22, sizes=(4, 6, 8))[0]
    m = O.Mat.of(A)
    H = dev(m)
    C = H.spmm_hermitian(H, 0.0, numeric=numeric)
    Cr = O.complete_hermitian(O.spgemm(m, m, 0.0, upper=True))
    assert np.array_equal(C.row_ptr, Cr.row_ptr) and np.array_equal(C.col_ind, Cr.col_ind)
    assert relerr(C.get_values(), Cr.vals) < tol
    D = dense(C)
    assert np.abs(D - D.T).max() == 0.0


def test_hermitian_edge_cases():
    """A threshold above every estimate gives an empty result; an empty matrix
    gives an empty result."""
    from vbcsr_amd import VBCSR
    from vbcsr_amd import workloads as W

    A = W.batched_hamiltonians(count=1, nb_lo=20, nb_hi=21, neighbours=4.0, seed=23)[0]
    H = VBCSR.from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)
    C = H.spmm_hermitian(H, 1e30)
    assert len(C.get_values()) == 0 and np.all(np.asarray(C.row_ptr) == 0)
    E = VBCSR.from_block_csr(A.bs, np.zeros(A.nb + 1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    assert len(E.spmm_hermitian(E, 0.0).get_values()) == 0
