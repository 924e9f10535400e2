"""CPU pins of the oracle's Hermitian SpGEMM (upper block triangle + Hermitian
completion) on the reference's own cases (tests/test_spmm_hermitian.py:23-68):
the product of a Hermitian congruence Z^H (S Z) and the square of a Hermitian
X, against the dense product; the result is exactly Hermitian.  Plus the
upper-triangle pattern against the dense product's upper blocks, and the
threshold rule on a sparse case against the full oracle product."""
import numpy as np

import oracle as O


def build(n_blocks, bs, rng, hermitian):
    """Full block pattern, Hermitian or not (after the reference's _build)."""
    blocks = {}
    for i in range(n_blocks):
        for j in range(i, n_blocks):
            b = rng.standard_normal((bs, bs)) + 1j * rng.standard_normal((bs, bs))
            if hermitian and i == j:
                b = 0.5 * (b + b.conj().T)
            blocks[(i, j)] = b
            if i != j:
                blocks[(j, i)] = b.conj().T if hermitian else (rng.standard_normal((bs, bs)) +
                                                                1j * rng.standard_normal((bs, bs)))
    rp = np.arange(n_blocks + 1, dtype=np.int32) * n_blocks
    ci = np.tile(np.arange(n_blocks, dtype=np.int32), n_blocks)
    vals = np.concatenate([blocks[(i, j)].ravel() for i in range(n_blocks) for j in range(n_blocks)])
    return O.Mat(np.full(n_blocks, bs, np.int32), rp, ci, vals)


def conj_transpose(m: O.Mat) -> O.Mat:
    D = m.dense().conj().T
    n = m.nbr
    bs = int(m.rbs[0])
    vals = np.concatenate([D[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs].ravel() for i in range(n) for j in range(n)])
    return O.Mat(m.rbs, m.row_ptr, m.col_ind, vals)


def hermitian_product(A, B, thr=0.0):
    return O.complete_hermitian(O.spgemm(A, B, thr, upper=True))


def test_congruence_matches_full_product():
    rng = np.random.default_rng(7)
    S = build(6, 3, rng, True)
    Z = build(6, 3, rng, False)
    SZ = O.spgemm(S, Z, 0.0)
    Zh = conj_transpose(Z)
    full = Zh.dense() @ SZ.dense()
    half = hermitian_product(Zh, SZ).dense()
    assert np.abs(full - half).max() < 1e-12 * np.abs(full).max()
    assert np.abs(half - half.conj().T).max() == 0.0


def test_square_of_hermitian_matches():
    rng = np.random.default_rng(11)
    X = build(5, 4, rng, True)
    full = X.dense() @ X.dense()
    half = hermitian_product(X, X).dense()
    assert np.abs(full - half).max() < 1e-12 * np.abs(full).max()
    assert np.abs(half - half.conj().T).max() == 0.0


def test_upper_pattern_and_threshold_on_sparse_product():
    """Upper blocks of the upper-only product equal the full product's upper
    blocks (same symbolic estimate, same threshold rules)."""
    from vbcsr_amd import workloads as W

    A = W.batched_hamiltonians(count=1, nb_lo=40, nb_hi=41, neighbours=6.0, seed=4)[0]
    m = O.Mat.of(A)
    for thr in (0.0, 1e-2, 0.5):
        full = O.spgemm(m, m, thr)
        up = O.spgemm(m, m, thr, upper=True)
        fv = full.voff
        uv = up.voff
        for i in range(m.nbr):
            fcols = [int(c) for c in full.col_ind[full.row_ptr[i]:full.row_ptr[i + 1]] if c >= i]
            ucols = list(up.col_ind[up.row_ptr[i]:up.row_ptr[i + 1]])
            assert fcols == ucols, (thr, i)
            for j in ucols:
                kf = full.row_ptr[i] + list(full.col_ind[full.row_ptr[i]:full.row_ptr[i + 1]]).index(j)
                ku = up.row_ptr[i] + ucols.index(j)
                assert np.array_equal(full.vals[fv[kf]:fv[kf + 1]], up.vals[uv[ku]:uv[ku + 1]])
