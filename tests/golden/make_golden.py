"""Write tests/golden/*.npz from the compiled reference core (oracle/_ref).

Run in the container that holds /root/reference after `make -C oracle`:

    python tests/golden/make_golden.py

Every fixture is data only: seeded inputs and the reference's outputs for
them, stored with numpy (no pickles).  The inputs come from the seeded
generators in vbcsr_amd/workloads.py.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from vbcsr_amd import workloads as W  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def pack(prefix, m):
    return {f"{prefix}_bs": m.rbs, f"{prefix}_row_ptr": m.row_ptr, f"{prefix}_col_ind": m.col_ind,
            f"{prefix}_vals": m.vals}


def main():
    # 1. variable-block real matrix: SpMV / SpMM / adjoint / norms
    A, x = W.config1_spmv(nb=40, extra=3, seed=11)
    m = oracle.Mat.of(A)
    R = oracle.RefMat(m)
    X = np.random.default_rng(12).uniform(-1, 1, (A.n, 3))
    d = pack("A", m)
    d.update(x=x, y=R.mult(x), X=X, Y=R.mult_dense(X), Yadj=R.mult_dense(X, adjoint=True),
             yadj=R.mult_adjoint(x), norms=R.block_norms(), kind=np.array(R.kind))
    # 2. SpGEMM A*A at three thresholds (patterns and values of the reference)
    norms = R.block_norms()
    for tag, thr in (("t0", 0.0), ("t1", float(np.median(norms))), ("t2", float(np.percentile(norms, 90) * 3))):
        C = R.spmm(R, thr).export()
        d.update({f"C{tag}_thr": np.array(thr), f"C{tag}_row_ptr": C.row_ptr, f"C{tag}_col_ind": C.col_ind,
                  f"C{tag}_vals": C.vals})
    # 3. union axpby: 3*C_t1 - 2*A (different patterns)
    C1 = R.spmm(R, d["Ct1_thr"].item())
    C1.axpby(-2.0, R, 3.0)
    U = C1.export()
    d.update(U_row_ptr=U.row_ptr, U_col_ind=U.col_ind, U_vals=U.vals)
    # 4. shift, filter
    S = R.duplicate()
    S.shift(0.75)
    d.update(S_vals=S.export().vals)
    F = R.duplicate()
    F.filter(float(np.median(norms)))
    Fe = F.export()
    d.update(F_thr=np.array(float(np.median(norms))), F_row_ptr=Fe.row_ptr, F_col_ind=Fe.col_ind, F_vals=Fe.vals)
    # 5. Rademacher multivector (seed 2026, 4 columns, normalised)
    d.update(rademacher=R.random_mv(4, 2026, True))
    np.savez_compressed(os.path.join(OUT, "vbcsr_real.npz"), **d)

    # 6. complex Hermitian BSR (graphene 3x4 cells, 4 orbitals)
    G = W.graphene(3, 4, 4, seed=5, complex_blocks=True, scale=False)
    gm = oracle.Mat.of(G)
    GR = oracle.RefMat(gm)
    z = np.random.default_rng(6).standard_normal(G.n) + 1j * np.random.default_rng(7).standard_normal(G.n)
    Z = np.random.default_rng(8).standard_normal((G.n, 2)) + 1j * np.random.default_rng(9).standard_normal((G.n, 2))
    e = pack("G", gm)
    e.update(z=z, y=GR.mult(z), Z=Z, Y=GR.mult_dense(Z), Yadj=GR.mult_dense(Z, adjoint=True),
             kind=np.array(GR.kind))
    C = GR.spmm(GR, 0.0).export()
    e.update(C_row_ptr=C.row_ptr, C_col_ind=C.col_ind, C_vals=C.vals)
    np.savez_compressed(os.path.join(OUT, "bsr_complex.npz"), **e)

    # 7. KPM moments of a scaled real graphene sample (ex08 composition)
    H = W.graphene(4, 4, 4, seed=9)
    hm = oracle.Mat.of(H)
    HR = oracle.RefMat(hm)
    V0 = HR.random_mv(5, 42, False)
    mu, _ = HR.kpm_moments(V0, 24)
    f = pack("H", hm)
    f.update(V0=V0, mu=mu)
    np.savez_compressed(os.path.join(OUT, "kpm_graphene.npz"), **f)
    print("wrote fixtures to", OUT)


if __name__ == "__main__":
    main()
