"""GPU parity of block SpMV / SpMM / adjoint against the oracle and the
reference's golden outputs.  FP64 tolerance 1e-12 relative (max-abs error over
max-abs reference); FP32-mixed tolerance 2e-6 relative."""
import numpy as np
import pytest

import oracle as O
from conftest import relerr

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-6


def vb():
    from vbcsr_amd import VBCSR

    return VBCSR


def from_golden(d, p):
    return vb().from_block_csr(d[f"{p}_bs"], d[f"{p}_row_ptr"], d[f"{p}_col_ind"], d[f"{p}_vals"])


def test_golden_real(golden_real):
    d = golden_real
    H = from_golden(d, "A")
    assert H.matrix_kind == str(d["kind"])
    assert relerr(H.mult(d["x"]).to_numpy(), d["y"]) < FP64_TOL
    assert relerr(H.mult(d["X"]).to_numpy(), d["Y"]) < FP64_TOL
    assert relerr(H.mult_adjoint(d["X"]).to_numpy(), d["Yadj"]) < FP64_TOL
    assert relerr(H.mult_adjoint(d["x"]).to_numpy(), d["yadj"]) < FP64_TOL
    assert relerr(H @ d["x"], d["y"]) < FP64_TOL if isinstance(H @ d["x"], np.ndarray) else True


def test_golden_complex(golden_complex):
    d = golden_complex
    H = from_golden(d, "G")
    assert H.matrix_kind == "bsr"
    assert relerr(H.mult(d["z"]).to_numpy(), d["y"]) < FP64_TOL
    assert relerr(H.mult(d["Z"]).to_numpy(), d["Y"]) < FP64_TOL
    assert relerr(H.mult_adjoint(d["Z"]).to_numpy(), d["Yadj"]) < FP64_TOL


@pytest.mark.parametrize("K", [1, 3, 10, 32, 70])
def test_config1_vs_oracle(K):
    from vbcsr_amd import workloads as W

    A, x = W.config1_spmv()
    H = vb().from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)
    m = O.Mat.of(A)
    X = np.random.default_rng(K).uniform(-1, 1, (A.n, K))
    Y = H.mult(X if K > 1 else X[:, 0]).to_numpy()
    Yr = O.spmm(m, X)
    assert relerr(Y.reshape(Yr.shape), Yr) < FP64_TOL
    # the adjoint of a symmetric matrix equals the forward product
    Ya = H.mult_adjoint(X if K > 1 else X[:, 0]).to_numpy()
    assert relerr(Ya.reshape(Yr.shape), Yr) < FP64_TOL


def test_config1_fp32_mixed():
    from vbcsr_amd import workloads as W

    A, x = W.config1_spmv()
    H = vb().from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)
    X = H.create_multivector(8)
    X.from_numpy(np.random.default_rng(5).uniform(-1, 1, (A.n, 8)))
    Y = H.create_multivector(8)
    # FP32 through the raw entry point: float buffers made by torch
    import torch

    Xf = torch.tensor(X.to_numpy(), dtype=torch.float32, device="cuda").contiguous()
    Yf = torch.zeros_like(Xf)
    H._core.spmm_raw(8, Xf.data_ptr(), 8, Yf.data_ptr(), 8, 1, False)
    torch.cuda.synchronize()
    Yr = O.spmm(O.Mat.of(A), np.asarray(X.to_numpy()))
    assert relerr(Yf.cpu().numpy(), Yr) < FP32_TOL
    H.mult(X, Y)
    assert relerr(Y.to_numpy(), Yr) < FP64_TOL


@pytest.mark.parametrize("sizes", [(1, 13, 20), (1, 4, 9, 13), (2,), (8,), (4,)])
def test_block_size_families(sizes):
    """Row-pass tiling (r > 64/Kp), compile-time BSR paths and CSR rows."""
    from vbcsr_amd import workloads as W

    rng = np.random.default_rng(len(sizes) * 7 + sizes[0])
    nb = 60
    bs = rng.choice(np.asarray(sizes), size=nb).astype(np.int32)
    iu, ju, _ = W.random_geometric_pairs(nb, 8.0, rng)
    A = W.symmetric_from_upper(bs, iu, ju, lambda g, n, r, c, sel: g.standard_normal((n, r, c)), rng)
    H = vb().from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)
    m = O.Mat.of(A)
    for K in (1, 5, 40):
        X = rng.standard_normal((A.n, K))
        Y = H.mult(X if K > 1 else X[:, 0]).to_numpy().reshape(A.n, K)
        assert relerr(Y, O.spmm(m, X)) < FP64_TOL, (sizes, K)


def test_local_share_with_ghost_columns():
    """Owned block rows plus explicitly sized ghost columns (one rank's share)."""
    rng = np.random.default_rng(3)
    owned = [0, 1, 2, 3]
    osz = [3, 2, 4, 3]
    ghosts = [10, 7, 12]
    gsz = [2, 5, 1]
    adj = [[0, 1, 10], [1, 7], [2, 3, 12, 7], [0, 3, 10, 12]]
    H = vb().create_local(owned, osz, adj, ghosts, gsz)
    # local column order: owned 0..3, then ghosts sorted by global id: 7, 10, 12
    col_sizes = osz + [5, 2, 1]
    lmap = {0: 0, 1: 1, 2: 2, 3: 3, 7: 4, 10: 5, 12: 6}
    rp = [0]
    ci = []
    vals = []
    for i, row in enumerate(adj):
        cols = sorted(lmap[g] for g in row)
        ci += cols
        rp.append(len(ci))
        for lc in cols:
            g = [k for k, v in lmap.items() if v == lc][0]
            blk = rng.standard_normal((osz[i], col_sizes[lc]))
            H.add_block(i, g, blk)
            vals.append(blk.ravel())
    H.assemble()
    m = O.Mat(osz, rp, ci, np.concatenate(vals), cbs=col_sizes)
    xfull = rng.standard_normal(sum(col_sizes))
    v = H.create_vector()
    assert v.local_size == sum(osz) and v.ghost_size == 8
    buf = np.asarray(memoryview(v._core))
    buf[:] = xfull  # owned entries and the already-synchronised ghost entries
    y = H.mult(v)
    assert relerr(y.to_numpy(), O.spmv(m, xfull)) < FP64_TOL
    assert np.all(np.asarray(memoryview(y._core))[sum(osz):] == 0.0)  # ghost tail zeroed
    ya = H.mult_adjoint(y)
    assert relerr(np.asarray(memoryview(ya._core)), O.spmm_adjoint(m, y.to_numpy()[:, None])[:, 0]) < FP64_TOL


def test_empty_rows_and_empty_matrix():
    V = vb()
    H = V.create_serial(4, [2, 3, 1, 2], [[0, 1], [], [2], []])
    H.add_block(0, 0, np.eye(2))
    H.add_block(0, 1, np.ones((2, 3)))
    H.add_block(2, 2, np.array([[5.0]]))
    H.assemble()
    x = np.arange(8, dtype=float) + 1
    y = H.mult(x).to_numpy()
    assert np.allclose(y, [1 + 3 + 4 + 5, 2 + 3 + 4 + 5, 0, 0, 0, 30, 0, 0])
    E = V.create_serial(3, [2, 2, 2], [[], [], []])
    assert np.all(E.mult(np.ones(6)).to_numpy() == 0.0)
    Z = V.create_serial(0, [], [])
    assert Z.mult(np.zeros(0)).to_numpy().shape == (0,)


def test_graphene_bsr_fp64_fp32():
    from vbcsr_amd import workloads as W

    A = W.graphene(20, 20, 4, seed=1)
    H = vb().from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)
    assert H.matrix_kind == "bsr"
    m = O.Mat.of(A)
    X = np.random.default_rng(2).choice([-1.0, 1.0], size=(A.n, 10))
    assert relerr(H.mult(X).to_numpy(), O.spmm(m, X)) < FP64_TOL
    assert relerr(H.mult(X[:, 0]).to_numpy(), O.spmv(m, X[:, 0])) < FP64_TOL


@pytest.mark.parametrize("bs_hi,cplx", [(8, False), (4, False), (4, True)])
def test_large_spmv_vs_oracle(bs_hi, cplx):
    """SpMV on matrices whose x exceeds one XCD's L2 (over 2 MB in FP64),
    4-8 and 2-4 row blocks, real and complex, against the oracle."""
    from vbcsr_amd import workloads as W

    A = W.random_vbcsr(60000 if bs_hi == 8 else 90000, bs_lo=bs_hi // 2, bs_hi=bs_hi, per_row=6, seed=5)
    assert A.n * 8 > (2 << 20)
    rng = np.random.default_rng(9)
    if cplx:
        A.vals = A.vals + 1j * rng.uniform(-1, 1, A.vals.shape)
    H = vb().from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)
    x = rng.uniform(-1, 1, A.n) + (1j * rng.uniform(-1, 1, A.n) if cplx else 0)
    y = H.mult(x).to_numpy()
    yr = O.spmm(O.Mat.of(A), x[:, None])[:, 0]
    assert relerr(y, yr) < FP64_TOL
