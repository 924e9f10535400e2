"""GPU checks of the batched-KPM tile plans: every plan the autotuner can
pick (class threshold, the class rows on the matrix cores or -- negative
thresholds -- on the block-row wave kernel, x lane-family slot partitions or
row-wave rows per work-group, the latter given as negative parts) gives the
oracle's moments
(FP64: 1e-11 relative; precision "mixed": 1e-11 against the oracle on the
FP32-rounded matrix), the autotuner reports a timed plan from its candidate
set, and a second call on the same batch reuses that plan bit for bit."""
import numpy as np
import pytest

import oracle as O
from conftest import relerr
from test_gpu_kpm import mk, rounded32

pytestmark = pytest.mark.gpu


def batch(seed, count=4, K=8):
    from vbcsr_amd import workloads as W

    Hs = W.batched_hamiltonians(count=count, nb_lo=40, nb_hi=160, neighbours=8.0, seed=seed)
    mats = [mk(A) for A in Hs]
    V0s = []
    for h, H in enumerate(mats):
        v = H.create_multivector(K)
        v.set_random(300 + h, normalize=False)
        V0s.append(v)
    return Hs, mats, V0s


@pytest.fixture
def kpm_api():
    from vbcsr_amd import kpm as K

    yield K
    K.set_batch_plan(0)
    K.set_autotune(True)


@pytest.mark.parametrize("threshold", [4, 8, 17])
@pytest.mark.parametrize("parts", [1, 2, 8, -8, -32])
def test_every_plan_matches_oracle(kpm_api, threshold, parts):
    K = kpm_api
    Hs, mats, V0s = batch(11)
    K.set_batch_plan(threshold, parts)
    M = 30
    for precision, graph in (("fp64", True), ("fp64", False), ("mixed", True)):
        mu = K.kpm_moments_batched(mats, V0s, M, mode="two", precision=precision, graph=graph)
        assert K.last_plan()[0] == threshold and not K.last_plan()[2]
        for h in range(len(mats)):
            A = Hs[h] if precision == "fp64" else rounded32(Hs[h])
            ref = O.kpm_moments(O.Mat.of(A), np.asarray(V0s[h].to_numpy()), M, mode=1)
            assert relerr(mu[h], ref) < 1e-11, (precision, graph, h)


def test_autotuner_picks_a_candidate_and_keeps_it(kpm_api):
    K = kpm_api
    K.set_batch_plan(0)
    K.set_autotune(True)
    Hs, mats, V0s = batch(12)
    M = 24
    mu1 = K.kpm_moments_batched(mats, V0s, M, mode="two", precision="mixed")
    t, p, tuned = K.last_plan()
    assert tuned and t in (4, 8, 17, -1, -4, -8) and (p >= 1 or p in (-8, -16, -32))
    mu2 = K.kpm_moments_batched(mats, V0s, M, mode="two", precision="mixed")
    assert K.last_plan() == (t, p, True)
    assert np.array_equal(mu1, mu2)
    for h in range(len(mats)):
        ref = O.kpm_moments(O.Mat.of(rounded32(Hs[h])), np.asarray(V0s[h].to_numpy()), M, mode=1)
        assert relerr(mu1[h], ref) < 1e-11


def test_autotune_off_uses_the_default_rule(kpm_api):
    K = kpm_api
    K.set_batch_plan(0)
    K.set_autotune(False)
    Hs, mats, V0s = batch(13)
    mu = K.kpm_moments_batched(mats, V0s, 20, mode="two")
    t, p, tuned = K.last_plan()
    assert not tuned and t == 4
    for h in range(len(mats)):
        ref = O.kpm_moments(O.Mat.of(Hs[h]), np.asarray(V0s[h].to_numpy()), 20, mode=1)
        assert relerr(mu[h], ref) < 1e-11


def test_single_matrix_calls_keep_their_row_classes(kpm_api):
    """A batched call under another plan leaves each matrix's own lists active."""
    K = kpm_api
    Hs, mats, V0s = batch(14, count=2)
    K.set_batch_plan(17, 2)
    K.kpm_moments_batched(mats, V0s, 10, mode="two")
    K.set_batch_plan(0)
    X = np.random.default_rng(0).standard_normal((Hs[0].n, 8))
    Xv = mats[0].create_multivector(8)
    Xv.from_numpy(X)
    Y = np.asarray(mats[0].mult(Xv).to_numpy())
    ref = O.spmm(O.Mat.of(Hs[0]), X)
    assert relerr(Y, ref) < 1e-12


@pytest.mark.parametrize("threshold", [4, 17])
def test_batch_with_only_class_rows_or_only_lane_rows(kpm_api, threshold):
    """Edge layouts: one matrix whose rows are all matrix-core rows (no lane
    rows: an empty lane launch for it) next to one with lane rows only."""
    from vbcsr_amd import workloads as W

    K = kpm_api
    Hs = [W.batched_hamiltonians(count=1, nb_lo=30, nb_hi=31, neighbours=6.0, seed=31, sizes=(9, 13))[0],
          W.batched_hamiltonians(count=1, nb_lo=30, nb_hi=31, neighbours=6.0, seed=32, sizes=(1, 2))[0]]
    mats = [mk(A) for A in Hs]
    V0s = []
    for h, H in enumerate(mats):
        v = H.create_multivector(6)
        v.set_random(500 + h, normalize=False)
        V0s.append(v)
    K.set_batch_plan(threshold, 2)
    for graph in (False, True):
        mu = K.kpm_moments_batched(mats, V0s, 16, mode="two", graph=graph)
        for h in range(2):
            ref = O.kpm_moments(O.Mat.of(Hs[h]), np.asarray(V0s[h].to_numpy()), 16, mode=1)
            assert relerr(mu[h], ref) < 1e-11, (graph, h)
    # only matrix-core rows in the whole batch: no lane launch at all
    K.set_batch_plan(4, 1)
    mu = K.kpm_moments_batched(mats[:1], V0s[:1], 16, mode="two")
    ref = O.kpm_moments(O.Mat.of(Hs[0]), np.asarray(V0s[0].to_numpy()), 16, mode=1)
    assert relerr(mu[0], ref) < 1e-11


@pytest.mark.parametrize("K", [4, 16, 32, 64])
@pytest.mark.parametrize("threshold", [4, 17])
def test_row_wave_plans_wide(kpm_api, K, threshold):
    """The row-wave kernel at the column counts it serves (FP64: K = 2 * 2^k,
    FP32 recursion: K = 4 * 2^k), with rows of more than 64 blocks (LDS entry
    lists refilled within a row) against the oracle."""
    from vbcsr_amd import workloads as W

    Kapi = kpm_api
    Hs = W.batched_hamiltonians(count=3, nb_lo=60, nb_hi=200, neighbours=70.0, seed=41)
    mats = [mk(A) for A in Hs]
    V0s = []
    for h, H in enumerate(mats):
        v = H.create_multivector(K)
        v.set_random(900 + h, normalize=False)
        V0s.append(v)
    Kapi.set_batch_plan(threshold, -16)
    M = 12
    for precision, tol in (("fp64", 1e-11), ("mixed", 1e-11), ("fp32", 2e-5)):
        mu = Kapi.kpm_moments_batched(mats, V0s, M, mode="two", precision=precision)
        assert Kapi.last_plan()[1] == -16, precision
        for h in range(len(mats)):
            A = Hs[h] if precision == "fp64" else rounded32(Hs[h])
            ref = O.kpm_moments(O.Mat.of(A), np.asarray(V0s[h].to_numpy()), M, mode=1)
            assert relerr(mu[h], ref) < tol, (precision, h)


@pytest.mark.parametrize("K", [16, 32, 64])
@pytest.mark.parametrize("threshold", [-4, -8, -1])
def test_block_row_wave_plans(kpm_api, K, threshold):
    """The class rows on the block-row wave kernel (vector ALUs, DPP row
    broadcasts of the column-major FP32 values) at the column counts it
    serves, with the other rows on the lane family or the row-wave kernel;
    threshold -1 puts every block row in the class (no lane launch).  Long
    rows (about 70 neighbours) and rows of every size 1..13.  Precision
    "mixed" against the oracle on the FP32-rounded matrix (1e-11), FP32
    vectors at 2e-5; graph and plain launches."""
    from vbcsr_amd import workloads as W

    Kapi = kpm_api
    Hs = W.batched_hamiltonians(count=3, nb_lo=60, nb_hi=200, neighbours=70.0, seed=43,
                                sizes=(1, 2, 3, 4, 5, 7, 9, 13))
    mats = [mk(A) for A in Hs]
    V0s = []
    for h, H in enumerate(mats):
        v = H.create_multivector(K)
        v.set_random(700 + h, normalize=False)
        V0s.append(v)
    M = 12
    for parts in (1, -16):
        Kapi.set_batch_plan(threshold, parts)
        for precision, tol, graph in (("mixed", 1e-11, True), ("mixed", 1e-11, False), ("fp32", 2e-5, True)):
            mu = Kapi.kpm_moments_batched(mats, V0s, M, mode="two", precision=precision, graph=graph)
            assert Kapi.last_plan()[:2] == (threshold, parts), (precision, Kapi.last_plan())
            for h in range(len(mats)):
                ref = O.kpm_moments(O.Mat.of(rounded32(Hs[h])), np.asarray(V0s[h].to_numpy()), M, mode=1)
                assert relerr(mu[h], ref) < tol, (parts, precision, graph, h)
    # values changed after the column-major copy was built: it follows them
    Kapi.set_batch_plan(threshold, 1)
    for H in mats:
        H.scale(0.5)
    mu = Kapi.kpm_moments_batched(mats, V0s, M, mode="two", precision="mixed")
    for h in range(len(mats)):
        A = W.BlockCSR(Hs[h].bs, Hs[h].row_ptr, Hs[h].col_ind, 0.5 * Hs[h].vals)
        ref = O.kpm_moments(O.Mat.of(rounded32(A)), np.asarray(V0s[h].to_numpy()), M, mode=1)
        assert relerr(mu[h], ref) < 1e-11, h


def _drop_block_row(A, i):
    """A with every block of block row i and block column i removed (a
    symmetric matrix with an empty block row)."""
    from vbcsr_amd.workloads import BlockCSR

    rp, ci, bs = A.row_ptr, A.col_ind, A.bs
    voff = np.zeros(len(ci) + 1, np.int64)
    for r in range(A.nb):
        for k in range(rp[r], rp[r + 1]):
            voff[k + 1] = voff[k] + bs[r] * bs[ci[k]]
    new_rp, new_ci, vals = [0], [], []
    for r in range(A.nb):
        for k in range(rp[r], rp[r + 1]):
            if r != i and ci[k] != i:
                new_ci.append(ci[k])
                vals.append(A.vals[voff[k]:voff[k + 1]])
        new_rp.append(len(new_ci))
    return BlockCSR(bs, np.asarray(new_rp, np.int32), np.asarray(new_ci, np.int32), np.concatenate(vals))


@pytest.mark.parametrize("threshold", [-1, -4])
def test_block_row_wave_empty_block_row(kpm_api, threshold):
    """A class row without stored blocks still gets Y = beta X + gamma Z."""
    from vbcsr_amd import workloads as W

    Hs = W.batched_hamiltonians(count=2, nb_lo=40, nb_hi=80, neighbours=8.0, seed=47, sizes=(4, 9, 13))
    Hs[1] = _drop_block_row(Hs[1], 7)
    mats = [mk(A) for A in Hs]
    V0s = []
    for h, H in enumerate(mats):
        v = H.create_multivector(32)
        v.set_random(60 + h, normalize=False)
        V0s.append(v)
    Kapi = kpm_api
    Kapi.set_batch_plan(threshold, 1)
    mu = Kapi.kpm_moments_batched(mats, V0s, 16, mode="two", precision="mixed")
    assert Kapi.last_plan()[0] == threshold
    for h in range(2):
        ref = O.kpm_moments(O.Mat.of(rounded32(Hs[h])), np.asarray(V0s[h].to_numpy()), 16, mode=1)
        assert relerr(mu[h], ref) < 1e-11, h
