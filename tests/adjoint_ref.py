"""Dense FP64 torch reference of the Chebyshev propagation and its loss, used
by the adjoint tests as the independent gradient (torch autograd through dense
complex128 products, on the CPU)."""
import numpy as np
import torch


def value_coordinates(A):
    """(row, col) scalar coordinates of every stored value, slot order."""
    off = A.offsets()
    voff = A.value_offsets()
    rows, cols = [], []
    for i in range(A.nb):
        for k in range(A.row_ptr[i], A.row_ptr[i + 1]):
            j = A.col_ind[k]
            r, c = int(A.bs[i]), int(A.bs[j])
            a, b = np.meshgrid(np.arange(r), np.arange(c), indexing="ij")
            rows.append(off[i] + a.ravel())
            cols.append(off[j] + b.ravel())
            assert voff[k + 1] - voff[k] == r * c
    return np.concatenate(rows), np.concatenate(cols)


def dense(A):
    rows, cols = value_coordinates(A)
    D = np.zeros((A.n, A.n), A.vals.dtype)
    D[rows, cols] = A.vals
    return D


def propagate_dense(values, rows, cols, n, psi0, coef, center, half, nsteps):
    """psi(T) with H assembled densely from ``values`` (differentiable)."""
    H = torch.zeros((n, n), dtype=torch.complex128).index_put((rows, cols), values)
    Ht = (H - center * torch.eye(n, dtype=torch.complex128)) / half
    c = torch.as_tensor(coef, dtype=torch.complex128)
    psi = psi0
    for _ in range(nsteps):
        t0 = psi
        t1 = Ht @ t0
        y = c[0] * t0 + c[1] * t1
        for k in range(2, len(coef)):
            t2 = 2.0 * (Ht @ t1) - t0
            y = y + c[k] * t2
            t0, t1 = t1, t2
        psi = y
    return psi


def overlap_loss(psi, target):
    """|<target|psi>|^2, summed over columns."""
    ov = (target.conj() * psi).sum(0)
    return (ov.real ** 2 + ov.imag ** 2).sum()
