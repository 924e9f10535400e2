"""GPU thresholded SpGEMM, filter, union axpby and McWeeny purification
against the reference's golden outputs and the oracle.  FP64 values 1e-12
relative; patterns identical (the estimates differ from the reference's only
in summation order, so a block may flip only when its estimate lies within
rounding of the threshold -- none does in these cases).  Reduced-precision
numeric phases (FP32, BF16 on the matrix cores) keep the FP64 symbolic
pattern; their values are checked to 1e-5 (FP32) and 2e-2 (BF16) relative."""
import numpy as np
import pytest

import oracle as O
from conftest import relerr

pytestmark = pytest.mark.gpu


def from_golden(d, p):
    from vbcsr_amd import VBCSR

    return VBCSR.from_block_csr(d[f"{p}_bs"], d[f"{p}_row_ptr"], d[f"{p}_col_ind"], d[f"{p}_vals"])


@pytest.mark.parametrize("tag", ["t0", "t1", "t2"])
def test_golden_spgemm(golden_real, tag):
    d = golden_real
    A = from_golden(d, "A")
    C = A.spmm(A, float(d[f"C{tag}_thr"]))
    assert np.array_equal(C.row_ptr, d[f"C{tag}_row_ptr"])
    assert np.array_equal(C.col_ind, d[f"C{tag}_col_ind"])
    assert relerr(C.get_values(), d[f"C{tag}_vals"]) < 1e-12


def test_golden_spgemm_complex(golden_complex):
    d = golden_complex
    G = from_golden(d, "G")
    C = G @ G
    assert np.array_equal(C.row_ptr, d["C_row_ptr"]) and np.array_equal(C.col_ind, d["C_col_ind"])
    assert relerr(C.get_values(), d["C_vals"]) < 1e-12


def test_golden_filter_axpby_shift(golden_real):
    d = golden_real
    A = from_golden(d, "A")
    F = A.duplicate()
    F.filter_blocks(float(d["F_thr"]))
    assert np.array_equal(F.row_ptr, d["F_row_ptr"]) and np.array_equal(F.col_ind, d["F_col_ind"])
    assert np.array_equal(F.get_values(), d["F_vals"])
    C1 = A.spmm(A, float(d["Ct1_thr"]))
    C1.axpby(-2.0, A, 3.0)
    assert np.array_equal(C1.row_ptr, d["U_row_ptr"]) and np.array_equal(C1.col_ind, d["U_col_ind"])
    assert relerr(C1.get_values(), d["U_vals"]) < 1e-12
    S = A.duplicate()
    S.shift(0.75)
    assert np.array_equal(S.get_values(), d["S_vals"])


def test_config1_spgemm_vs_oracle():
    from vbcsr_amd import VBCSR
    from vbcsr_amd import workloads as W

    A, _ = W.config1_spmv()
    H = VBCSR.from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)
    m = O.Mat.of(A)
    norms = O.block_norms(m)
    for thr in (0.0, float(np.median(norms)) * 4, float(norms.max()) * 3):
        C = H.spmm(H, thr)
        Cr = O.spgemm(m, m, thr)
        assert np.array_equal(C.row_ptr, Cr.row_ptr) and np.array_equal(C.col_ind, Cr.col_ind), thr
        assert relerr(C.get_values(), Cr.vals) < 1e-12


def test_asymmetric_row_eps():
    from vbcsr_amd import VBCSR

    adj = [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3, 4], [3, 4]]
    val = {(0, 0): 1, (0, 1): 0.8, (0, 2): 0.2, (1, 0): 0.8, (1, 1): 1, (1, 3): 1, (2, 0): 0.2, (2, 2): 1,
           (2, 3): 1, (3, 1): 1, (3, 2): 1, (3, 3): 1, (3, 4): 1, (4, 3): 1, (4, 4): 1}
    H = VBCSR.create_serial(5, [1] * 5, adj)
    for (i, j), v in val.items():
        H.add_block(i, j, np.array([[float(v)]]))
    H.assemble()
    D = H.spmm(H, 0.9).to_dense()
    assert abs(D[0, 3] - D[3, 0]) < 1e-12


@pytest.mark.parametrize("numeric,tol", [("fp32", 1e-5), ("bf16", 2e-2)])
def test_reduced_precision_numeric(numeric, tol):
    from vbcsr_amd import VBCSR
    from vbcsr_amd import workloads as W

    P = W.mcweeny_start(nb=300, neighbours=12.0, seed=2)
    H = VBCSR.from_block_csr(P.bs, P.row_ptr, P.col_ind, P.vals)
    m = O.Mat.of(P)
    thr = 1e-5
    C = H.spmm(H, 0.0, numeric=numeric)
    Cr = O.spgemm(m, m, 0.0)
    assert np.array_equal(C.row_ptr, Cr.row_ptr) and np.array_equal(C.col_ind, Cr.col_ind)
    assert relerr(C.get_values(), Cr.vals) < tol
    C2 = H.spmm(H, thr, numeric=numeric)
    assert C2.local_block_nnz <= Cr.col_ind.shape[0]


def test_mcweeny_iterations_vs_oracle():
    """P <- 3 P^2 - 2 P^3 with thresholded products (README.md:21-34 form)."""
    from vbcsr_amd import VBCSR
    from vbcsr_amd import workloads as W

    P0 = W.mcweeny_start(nb=400, neighbours=14.0, seed=1)
    thr = 1e-5
    P = VBCSR.from_block_csr(P0.bs, P0.row_ptr, P0.col_ind, P0.vals)
    pm = O.Mat.of(P0)
    for it in range(3):
        P2 = P.spmm(P, thr)
        P = 3.0 * P2 - 2.0 * P2.spmm(P, thr)
        p2 = O.spgemm(pm, pm, thr)
        p3 = O.spgemm(p2, pm, thr)
        pm = O.axpby(p2, -2.0, p3, 3.0)
        assert np.array_equal(P.row_ptr, pm.row_ptr) and np.array_equal(P.col_ind, pm.col_ind), it
        assert relerr(P.get_values(), pm.vals) < 1e-11, it
    assert abs(P.trace() - np.trace(W.BlockCSR(P0.bs, pm.row_ptr, pm.col_ind, pm.vals).to_dense())) < 1e-8


@pytest.mark.parametrize("thr", [0.0, 0.05, -1.0])
def test_wide_rows_symbolic_paths(thr):
    """Rows whose hash table would exceed the LDS classes (min(paths, nbc) >
    2047): with thr >= 0 they take the dense per-column estimate kernel
    (columns emitted in order), with thr < 0 the global hash tables (every
    reached column kept, est > thr).  Both against the oracle: 2200 block
    rows of sizes 1-2, 47 blocks per row, so about 2200 rows have 2209 paths."""
    from vbcsr_amd import VBCSR
    from vbcsr_amd import workloads as W

    A = W.random_vbcsr(2200, bs_lo=1, bs_hi=2, per_row=46, seed=3)
    rows = np.diff(A.row_ptr)
    paths = np.array([rows[A.col_ind[A.row_ptr[i]:A.row_ptr[i + 1]]].sum() for i in range(A.nb)])
    assert (np.minimum(paths, A.nb) > 2047).sum() > 2000
    H = VBCSR.from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)
    m = O.Mat.of(A)
    C = H.spmm(H, thr)
    Cr = O.spgemm(m, m, thr)
    assert np.array_equal(C.row_ptr, Cr.row_ptr) and np.array_equal(C.col_ind, Cr.col_ind)
    assert relerr(C.get_values(), Cr.vals) < 1e-12


@pytest.mark.parametrize("numeric,tol", [("fp64", 1e-12), ("fp32", 1e-5), ("bf16", 2e-2)])
def test_many_block_columns_numeric(numeric, tol):
    """More than 8192 block columns: no LDS column map, so the FP64/FP32
    numeric phase locates C blocks by a per-lane binary search and BF16 uses
    the binary-search strip kernel.  9000 block rows of sizes 1-3, 9 blocks
    per row, against the oracle."""
    from vbcsr_amd import VBCSR
    from vbcsr_amd import workloads as W

    A = W.random_vbcsr(9000, bs_lo=1, bs_hi=3, per_row=8, seed=11)
    assert A.nb > 8192
    H = VBCSR.from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)
    m = O.Mat.of(A)
    C = H.spmm(H, 0.0, numeric=numeric)
    Cr = O.spgemm(m, m, 0.0)
    assert np.array_equal(C.row_ptr, Cr.row_ptr) and np.array_equal(C.col_ind, Cr.col_ind)
    assert relerr(C.get_values(), Cr.vals) < tol


@pytest.mark.parametrize("numeric,tol", [("fp64", 1e-12), ("fp32", 1e-5), ("bf16", 2e-2)])
def test_dense_panel_products(numeric, tol):
    """Products whose C fills more than a fifth of its block pattern take the
    dense-panel numeric kernel (B staged per inner block for a tile of rows,
    skip rule per pair, block norms for the filter from the same registers).
    Against the oracle: the pattern (after the threshold filter) and the
    values, for P^2 and P^2 P of a purification start, at threshold 1e-5 and 0,
    and the numeric pair count of the reference's rule.  BF16 products take
    one dense matrix-core GEMM instead (every stored pair multiplied; the
    pattern is still the symbolic one and the filter the same)."""
    from vbcsr_amd import VBCSR
    from vbcsr_amd import workloads as W

    P0 = W.mcweeny_start(nb=150, neighbours=30.0, seed=4, normalize="spectral")
    P = VBCSR.from_block_csr(P0.bs, P0.row_ptr, P0.col_ind, P0.vals)
    pm = O.Mat.of(P0)
    for thr in (1e-5, 0.0):
        P2 = P.spmm(P, thr, numeric=numeric)
        st = P.spmm_stats
        assert st[2] >= 0.8 * P0.nb * P0.nb, st  # C filled enough for the panel kernel
        p2 = O.spgemm(pm, pm, thr)
        if numeric == "bf16":  # blocks near the filter threshold may flip under BF16 rounding
            assert relerr(P2.to_dense(), W.BlockCSR(P0.bs, p2.row_ptr, p2.col_ind, p2.vals).to_dense()) < tol
            continue
        assert np.array_equal(P2.row_ptr, p2.row_ptr) and np.array_equal(P2.col_ind, p2.col_ind), thr
        assert relerr(P2.get_values(), p2.vals) < tol, thr
        P3 = P2.spmm(P, thr, numeric=numeric)
        p3 = O.spgemm(p2, pm, thr)
        assert np.array_equal(P3.row_ptr, p3.row_ptr) and np.array_equal(P3.col_ind, p3.col_ind), thr
        assert relerr(P3.get_values(), p3.vals) < tol, thr
    thr = 1e-5
    want = _numeric_pairs(P0, thr)
    P.spmm(P, thr, numeric=numeric)
    got = P.spmm_stats[1]
    if numeric == "bf16":  # the dense BF16 products multiply every candidate pair
        assert got == P.spmm_stats[0]
        return
    assert abs(got - want) <= 1e-3 * want, (got, want)


def _numeric_pairs(P0, thr):
    """Numeric pairs of P0 P0: (i, k, j) with A_ik and B_kj stored, C_ij in
    the symbolic pattern (est_ij > thr) and ||A_ik|| ||B_kj|| >= thr /
    nnz_row(A_i) (ops/spmm/vbcsr.hpp:371-387), counted on the CPU from the
    block norms."""
    nb, rp = P0.nb, P0.row_ptr
    nrm = np.zeros((nb, nb))
    pres = np.zeros((nb, nb), bool)
    bn = O.block_norms(O.Mat.of(P0))
    for i in range(nb):
        for k in range(rp[i], rp[i + 1]):
            nrm[i, P0.col_ind[k]] = bn[k]
            pres[i, P0.col_ind[k]] = True
    sym = (nrm @ nrm) > thr
    eps = thr / np.maximum(np.diff(rp), 1)
    want = 0
    for i in range(nb):
        prod = nrm[i][:, None] * nrm  # [k, j]
        ok = pres[i][:, None] & pres & sym[i][None, :] & ~(prod < eps[i])
        want += int(ok.sum())
    return want


@pytest.mark.parametrize("numeric,tol", [("fp64", 1e-12), ("fp32", 1e-5), ("bf16", 2e-2)])
def test_sparse_products_fused_filter_and_pairs(numeric, tol):
    """Sparse products (C far below the dense paths' fill): the numeric
    kernels write C's block norms for the threshold filter themselves (and
    the FP32 kernel widens its FP32 accumulators into C in the same pass).
    The filtered pattern and values against the oracle, and the numeric pair
    count of the reference's rule, for all three numeric types."""
    from vbcsr_amd import VBCSR
    from vbcsr_amd import workloads as W

    P0 = W.mcweeny_start(nb=600, neighbours=8.0, seed=9, normalize="spectral")  # C fill ~0.05
    P = VBCSR.from_block_csr(P0.bs, P0.row_ptr, P0.col_ind, P0.vals)
    pm = O.Mat.of(P0)
    for thr in (1e-3, 1e-5):
        C = P.spmm(P, thr, numeric=numeric)
        st = P.spmm_stats
        assert st[2] < 0.1 * P0.nb * P0.nb, st  # the sparse kernels, not the dense paths
        assert st[3] <= st[2], st
        ref = O.spgemm(pm, pm, thr)
        if numeric == "bf16":  # blocks near the threshold may flip under BF16 rounding
            assert relerr(C.to_dense(), W.BlockCSR(P0.bs, ref.row_ptr, ref.col_ind, ref.vals).to_dense()) < tol
        else:
            assert np.array_equal(C.row_ptr, ref.row_ptr) and np.array_equal(C.col_ind, ref.col_ind), thr
            assert relerr(C.get_values(), ref.vals) < tol, thr
        want = _numeric_pairs(P0, thr)
        assert abs(int(st[1]) - want) <= 1e-3 * want, (numeric, thr, int(st[1]), want)


def test_device_transpose_and_trans_products(golden_complex):
    """Conjugate transpose on the device (vbh_matrix_transpose) against the
    dense conjugate transpose, for a non-symmetric variable-block matrix and
    the reference's complex BSR golden; spmm with transA / transB against
    the oracle on the explicit transposes (block_csr.hpp:1530-1543)."""
    from vbcsr_amd import VBCSR
    from vbcsr_amd import workloads as W

    A = W.random_vbcsr(60, per_row=7, seed=3)  # not symmetric
    H = VBCSR.from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)
    Ht = H.transpose()
    assert np.array_equal(Ht.to_dense(), A.to_dense().T)
    for i in range(A.nb):  # columns ascending per row
        cols = Ht.col_ind[Ht.row_ptr[i]:Ht.row_ptr[i + 1]]
        assert np.all(np.diff(cols) > 0)
    d = golden_complex
    G = from_golden(d, "G")
    Gd = G.to_dense()
    assert np.array_equal(G.transpose().to_dense(), Gd.conj().T)
    # A^T A (transA) and A A^T (transB) against the oracle on the explicit
    # transpose
    At = O.Mat.of(W.BlockCSR(A.bs, Ht.row_ptr, Ht.col_ind, Ht.get_values()))
    Am = O.Mat.of(A)
    for thr in (0.0, 0.5):
        C = H.spmm(H, thr, transA=True)
        Cr = O.spgemm(At, Am, thr)
        assert np.array_equal(C.row_ptr, Cr.row_ptr) and np.array_equal(C.col_ind, Cr.col_ind), thr
        assert relerr(C.get_values(), Cr.vals) < 1e-12
        C2 = H.spmm(H, thr, transB=True)
        Cr2 = O.spgemm(Am, At, thr)
        assert np.array_equal(C2.row_ptr, Cr2.row_ptr) and np.array_equal(C2.col_ind, Cr2.col_ind), thr
        assert relerr(C2.get_values(), Cr2.vals) < 1e-12
