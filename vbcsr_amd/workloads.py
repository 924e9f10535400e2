"""Seeded synthetic workloads for the variable-block sparse hot path.

Every generator returns a :class:`BlockCSR`: block-row CSR with per-block
row-major values packed in slot order -- the layout ``VBCSR.get_values()``,
``row_ptr`` and ``col_ind`` export (vbcsr/matrix.py:1008-1017,
vbcsr/core/block_csr.hpp:1202-1248).  Hamiltonian builders play the role of
the reference's example lattice builders (examples/ex08_graphene_dos.py:30-63,
vbcsr/utils.py:159-203) at the sizes of the benchmark configurations.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np


@dataclass
class BlockCSR:
    """Square block matrix in block-row CSR form (one rank's owned rows)."""

    bs: np.ndarray  # int32 [nb] block sizes (rows == cols)
    row_ptr: np.ndarray  # int32 [nb+1]
    col_ind: np.ndarray  # int32 [nnzb], sorted per row
    vals: np.ndarray  # float64/complex128 [nnz], row-major blocks in slot order
    meta: dict = field(default_factory=dict)

    @property
    def nb(self) -> int:
        return int(self.bs.shape[0])

    @property
    def nnzb(self) -> int:
        return int(self.col_ind.shape[0])

    @property
    def nnz(self) -> int:
        return int(self.vals.shape[0])

    @property
    def n(self) -> int:
        return int(self.bs.sum())

    def offsets(self) -> np.ndarray:
        off = np.zeros(self.nb + 1, np.int64)
        np.cumsum(self.bs, out=off[1:])
        return off

    def value_offsets(self) -> np.ndarray:
        rows = np.repeat(np.arange(self.nb), np.diff(self.row_ptr))
        sz = self.bs[rows].astype(np.int64) * self.bs[self.col_ind].astype(np.int64)
        off = np.zeros(self.nnzb + 1, np.int64)
        np.cumsum(sz, out=off[1:])
        return off

    def adjacency(self):
        return [self.col_ind[self.row_ptr[i]:self.row_ptr[i + 1]].tolist() for i in range(self.nb)]

    def to_dense(self) -> np.ndarray:
        off = self.offsets()
        voff = self.value_offsets()
        A = np.zeros((self.n, self.n), self.vals.dtype)
        for i in range(self.nb):
            r = self.bs[i]
            for k in range(self.row_ptr[i], self.row_ptr[i + 1]):
                j = self.col_ind[k]
                c = self.bs[j]
                A[off[i]:off[i] + r, off[j]:off[j] + c] = self.vals[voff[k]:voff[k + 1]].reshape(r, c)
        return A

    def to_scipy(self):
        import scipy.sparse as sp

        off = self.offsets()
        voff = self.value_offsets()
        rows = np.repeat(np.arange(self.nb), np.diff(self.row_ptr))
        r = self.bs[rows].astype(np.int64)
        c = self.bs[self.col_ind].astype(np.int64)
        sz = r * c
        blk = np.repeat(np.arange(self.nnzb), sz)
        e = np.arange(self.nnz) - voff[blk]
        ri = off[rows][blk] + e // c[blk]
        ci = off[self.col_ind][blk] + e % c[blk]
        return sp.csr_matrix((self.vals, (ri, ci)), shape=(self.n, self.n))

    def scaled(self, alpha: float, shift: float = 0.0) -> "BlockCSR":
        """alpha * A + shift * I (shift applied to stored diagonal blocks)."""
        vals = self.vals * alpha
        if shift != 0.0:
            voff = self.value_offsets()
            rows = np.repeat(np.arange(self.nb), np.diff(self.row_ptr))
            for k in np.nonzero(rows == self.col_ind)[0]:
                b = self.bs[rows[k]]
                vals[voff[k] + np.arange(b) * (b + 1)] += shift
        return BlockCSR(self.bs, self.row_ptr, self.col_ind, vals, dict(self.meta))


def _csr_from_pairs(nb: int, rows: np.ndarray, cols: np.ndarray):
    order = np.lexsort((cols, rows))
    rows, cols = rows[order], cols[order]
    keep = np.ones(rows.shape[0], bool)
    keep[1:] = (rows[1:] != rows[:-1]) | (cols[1:] != cols[:-1])
    rows, cols = rows[keep], cols[keep]
    row_ptr = np.zeros(nb + 1, np.int32)
    np.add.at(row_ptr, rows + 1, 1)
    np.cumsum(row_ptr, out=row_ptr)
    return row_ptr, cols.astype(np.int32), rows


def symmetric_from_upper(
    bs: np.ndarray,
    iu: np.ndarray,
    ju: np.ndarray,
    draw: Callable[[np.random.Generator, int, int, int, np.ndarray], np.ndarray],
    rng: np.random.Generator,
    hermitian_complex: bool = False,
) -> BlockCSR:
    """Symmetric (Hermitian) block matrix from upper pairs i <= j.

    ``draw(rng, n, r, c, pair_index)`` returns an (n, r, c) array of blocks for
    the n upper pairs of one (r, c) shape class.  The lower block (j, i) is the
    (conjugate) transpose; a diagonal block is symmetrised as (B + B^H) / 2.
    """
    nb = bs.shape[0]
    iu = np.asarray(iu, np.int64)
    ju = np.asarray(ju, np.int64)
    rows = np.concatenate([iu, ju[iu != ju]])
    cols = np.concatenate([ju, iu[iu != ju]])
    row_ptr, col_ind, srow = _csr_from_pairs(nb, rows, cols)
    dtype = np.complex128 if hermitian_complex else np.float64
    mat = BlockCSR(bs.astype(np.int32), row_ptr, col_ind, np.zeros(0, dtype))
    voff = mat.value_offsets()
    vals = np.zeros(voff[-1], dtype)
    # slot lookup for (i, j)
    key = srow.astype(np.int64) * nb + col_ind
    upper_slot = np.searchsorted(key, iu * nb + ju)
    lower_slot = np.searchsorted(key, ju * nb + iu)
    r_all = bs[iu]
    c_all = bs[ju]
    shape_key = r_all.astype(np.int64) * 1024 + c_all
    for sk in np.unique(shape_key):
        sel = np.nonzero(shape_key == sk)[0]
        r, c = int(sk // 1024), int(sk % 1024)
        B = np.asarray(draw(rng, sel.shape[0], r, c, sel), dtype)
        diag = iu[sel] == ju[sel]
        if diag.any():
            Bd = B[diag]
            B[diag] = 0.5 * (Bd + np.conj(np.transpose(Bd, (0, 2, 1))))
        idx = voff[upper_slot[sel]][:, None] + np.arange(r * c)[None, :]
        vals[idx] = B.reshape(sel.shape[0], r * c)
        off = ~diag
        if off.any():
            BT = np.conj(np.transpose(B[off], (0, 2, 1)))
            idx = voff[lower_slot[sel[off]]][:, None] + np.arange(r * c)[None, :]
            vals[idx] = BT.reshape(-1, r * c)
    mat.vals = vals
    return mat


def config1_spmv(nb: int = 500, bs_lo: int = 4, bs_hi: int = 8, extra: int = 20, seed: int = 0):
    """Variable-block FP64 SpMV case: block sizes U{bs_lo..bs_hi}; each block row
    holds its diagonal block plus ``extra`` distinct random block columns;
    entries U[-1,1]; then A <- (A + A^T)/2 (pattern P u P^T); x ~ U[-1,1].
    Returns (matrix, x)."""
    rng = np.random.default_rng(seed)
    bs = rng.integers(bs_lo, bs_hi + 1, size=nb).astype(np.int32)
    extra = min(extra, nb - 1)
    rows, cols = [], []
    for i in range(nb):
        choice = rng.choice(nb - 1, size=extra, replace=False)
        choice = choice + (choice >= i)
        rows.append(np.full(extra + 1, i))
        cols.append(np.concatenate([[i], choice]))
    rows = np.concatenate(rows).astype(np.int64)
    cols = np.concatenate(cols).astype(np.int64)
    # directed blocks A_ij ~ U[-1,1]; symmetric S_ij = (A_ij + A_ji^T) / 2
    row_ptr, col_ind, srow = _csr_from_pairs(nb, rows, cols)
    directed = BlockCSR(bs, row_ptr, col_ind, np.zeros(0))
    dvoff = directed.value_offsets()
    dvals = rng.uniform(-1.0, 1.0, size=dvoff[-1])
    key = srow.astype(np.int64) * nb + col_ind
    a, b = np.minimum(srow, col_ind), np.maximum(srow, col_ind)
    up = np.unique(a.astype(np.int64) * nb + b)
    iu, ju = up // nb, up % nb

    def draw(_rng, n, r, c, sel):
        out = np.zeros((n, r, c))
        for t, p in enumerate(sel):
            i, j = iu[p], ju[p]
            s = np.searchsorted(key, i * nb + j)
            if s < key.shape[0] and key[s] == i * nb + j:
                out[t] += dvals[dvoff[s]:dvoff[s + 1]].reshape(r, c)
            if i != j:
                s2 = np.searchsorted(key, j * nb + i)
                if s2 < key.shape[0] and key[s2] == j * nb + i:
                    out[t] += dvals[dvoff[s2]:dvoff[s2 + 1]].reshape(c, r).T
            else:
                pass
        # diagonal blocks: symmetric_from_upper symmetrises (B + B^T)/2 itself
        out[iu[sel] != ju[sel]] *= 0.5
        return out

    mat = symmetric_from_upper(bs, iu, ju, draw, rng)
    x = np.random.default_rng(seed + 1).uniform(-1.0, 1.0, size=mat.n)
    mat.meta.update(name="config1_spmv", seed=seed)
    return mat, x


def honeycomb_pairs(nx: int, ny: int, nnn: bool = True):
    """Upper (i <= j) atom pairs of a periodic honeycomb lattice with 2 atoms per
    cell (atom index 2*(ix*ny+iy)+s): on-site, 3 nearest neighbours (distance
    1) and 6 next-nearest neighbours (distance sqrt(3)) per atom."""
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ix, iy = ix.ravel(), iy.ravel()
    idx = lambda x, y, s: 2 * (((x % nx) * ny) + (y % ny)) + s
    a = idx(ix, iy, 0)
    pairs = [(a, a, 0.0), (a + 1, a + 1, 0.0)]
    # nearest neighbours: A(x,y) - B(x,y), B(x-1,y), B(x,y-1)
    for dx, dy in ((0, 0), (-1, 0), (0, -1)):
        pairs.append((a, idx(ix + dx, iy + dy, 1), 1.0))
    if nnn:
        # next-nearest: same sublattice, lattice vectors a1, a2, a1-a2
        for s in (0, 1):
            for dx, dy in ((1, 0), (0, 1), (1, -1)):
                pairs.append((idx(ix, iy, s), idx(ix + dx, iy + dy, s), np.sqrt(3.0)))
    I = np.concatenate([p[0] for p in pairs]).astype(np.int64)
    J = np.concatenate([p[1] for p in pairs]).astype(np.int64)
    D = np.concatenate([np.full(p[0].shape, p[2]) for p in pairs])
    lo, hi = np.minimum(I, J), np.maximum(I, J)
    key = lo * (2 * nx * ny) + hi
    key, first = np.unique(key, return_index=True)
    return lo[first], hi[first], D[first]


def spectral_bounds(mat: BlockCSR, tol: float = 1e-4):
    """Extreme eigenvalues of a symmetric/Hermitian matrix (Lanczos, scipy)."""
    import scipy.sparse.linalg as sla

    A = mat.to_scipy()
    if A.shape[0] < 400:
        w = np.linalg.eigvalsh(A.toarray())
        return float(w[0]), float(w[-1])
    v0 = np.random.default_rng(7).standard_normal(A.shape[0])
    lmax = sla.eigsh(A, k=1, which="LA", tol=tol, v0=v0, return_eigenvectors=False)[0]
    lmin = sla.eigsh(A, k=1, which="SA", tol=tol, v0=v0, return_eigenvectors=False)[0]
    return float(np.real(lmin)), float(np.real(lmax))


def scale_to_unit(mat: BlockCSR, edge: float = 0.99, bounds=None) -> BlockCSR:
    """Affine map of the spectrum onto [-edge, edge] (1e-3 relative safety on
    the Lanczos bounds)."""
    lo, hi = bounds if bounds is not None else spectral_bounds(mat)
    center = 0.5 * (lo + hi)
    half = 0.5 * (hi - lo) * (1.0 + 1e-3)
    out = mat.scaled(edge / half, -center * edge / half)
    out.meta.update(spectrum=(lo, hi), center=center, half_width=half / edge)
    return out


def graphene(nx: int = 200, ny: int = 200, orbitals: int = 4, seed: int = 0,
             complex_blocks: bool = False, scale: bool = True) -> BlockCSR:
    """Graphene-like tight-binding Hamiltonian: periodic honeycomb lattice,
    ``orbitals`` per atom, on-site + NN + NNN blocks with N(0,1) entries scaled
    by 1/distance (on-site unscaled), Hermitian, optionally scaled to
    [-0.99, 0.99]."""
    rng = np.random.default_rng(seed)
    iu, ju, dist = honeycomb_pairs(nx, ny)
    nb = 2 * nx * ny
    bs = np.full(nb, orbitals, np.int32)
    w = np.where(dist > 0, 1.0 / np.where(dist > 0, dist, 1.0), 1.0)

    def draw(g, n, r, c, sel):
        B = g.standard_normal((n, r, c))
        if complex_blocks:
            B = (B + 1j * g.standard_normal((n, r, c))) / np.sqrt(2.0)
        return B * w[sel][:, None, None]

    mat = symmetric_from_upper(bs, iu, ju, draw, rng, hermitian_complex=complex_blocks)
    if scale:
        mat = scale_to_unit(mat)
    mat.meta.update(name="graphene", nx=nx, ny=ny, orbitals=orbitals, seed=seed)
    return mat


def random_geometric_pairs(n: int, neighbours: float, rng: np.random.Generator):
    """Upper pairs (i <= j, with i == j included) of n uniform points in a
    periodic unit cube joined within the cutoff that gives ``neighbours``
    expected neighbours; returns (iu, ju, r)."""
    from scipy.spatial import cKDTree

    pts = rng.random((n, 3))
    rc = (neighbours * 3.0 / (4.0 * np.pi * n)) ** (1.0 / 3.0)
    rc = min(rc, 0.499)
    tree = cKDTree(pts, boxsize=1.0)
    pairs = tree.query_pairs(rc, output_type="ndarray")
    d = pts[pairs[:, 0]] - pts[pairs[:, 1]]
    d -= np.round(d)
    r = np.sqrt((d * d).sum(1))
    iu = np.concatenate([np.arange(n), pairs.min(1)])
    ju = np.concatenate([np.arange(n), pairs.max(1)])
    rr = np.concatenate([np.zeros(n), r]) / rc  # distances in cutoff units
    return iu, ju, rr


def mcweeny_start(nb: int = 5000, bs_lo: int = 4, bs_hi: int = 8, neighbours: float = 30.0,
                  seed: int = 0, h_scale: float = 0.1, h_shift: float = 0.5,
                  normalize=False) -> BlockCSR:
    """McWeeny purification start P0 = h_scale * H + h_shift * I, with H a
    symmetric random-geometric Hamiltonian: blocks N(0,1) * exp(-r) (r in
    cutoff units, r = 0 on the diagonal).

    As written (normalize=False) the spectrum of P0 reaches well outside
    [-0.5, 1.5] at these sizes (about [-1.1, 2.1] at 600 block rows), where
    P <- 3P^2 - 2P^3 diverges.  normalize="spectral" first divides H by its
    largest eigenvalue magnitude (Lanczos), so P0's spectrum spans
    [0.5 - h_scale, 0.5 + h_scale] and the iteration purifies;
    normalize=True divides by the Gershgorin bound instead (cheaper, but the
    bound overestimates the spectrum several-fold, so P0 starts much closer
    to 1/2 and purifies slowly)."""
    rng = np.random.default_rng(seed)
    bs = rng.integers(bs_lo, bs_hi + 1, size=nb).astype(np.int32)
    iu, ju, rr = random_geometric_pairs(nb, neighbours, rng)

    def draw(g, n, r, c, sel):
        return g.standard_normal((n, r, c)) * np.exp(-rr[sel])[:, None, None]

    H = symmetric_from_upper(bs, iu, ju, draw, rng)
    if normalize == "spectral":
        lo, hi = spectral_bounds(H, tol=1e-8)
        H = H.scaled(1.0 / max(abs(lo), abs(hi)))
        H.meta.update(spectrum=(lo, hi))
    elif normalize:
        H = gershgorin_scale(H, edge=1.0)
    P0 = H.scaled(h_scale, h_shift)
    P0.meta.update(name="mcweeny", nb=nb, seed=seed)
    return P0


def batched_hamiltonians(count: int = 64, nb_lo: int = 200, nb_hi: int = 5000,
                         sizes: Sequence[int] = (1, 4, 9, 13), neighbours: float = 20.0,
                         seed: int = 0, scale: bool = True):
    """``count`` independent symmetric random-geometric Hamiltonians with
    log-uniform block-row counts and block sizes drawn from ``sizes``, each
    scaled to [-0.99, 0.99]."""
    rng = np.random.default_rng(seed)
    out = []
    for h in range(count):
        nb = int(np.floor(np.exp(rng.uniform(np.log(nb_lo), np.log(nb_hi)))))
        bs = rng.choice(np.asarray(sizes, np.int32), size=nb).astype(np.int32)
        iu, ju, _ = random_geometric_pairs(nb, neighbours, rng)
        mat = symmetric_from_upper(bs, iu, ju, lambda g, n, r, c, sel: g.standard_normal((n, r, c)), rng)
        if scale:
            mat = scale_to_unit(mat)
        mat.meta.update(name="batched", index=h)
        out.append(mat)
    return out


def rademacher(n: int, k: int, seed: int) -> np.ndarray:
    """Unnormalised +/-1 columns (n x k) from numpy's generator."""
    g = np.random.default_rng(seed)
    return (2.0 * g.integers(0, 2, size=(n, k)) - 1.0)


def wavepacket(nx: int, ny: int, orbitals: int, width: float, seed: int) -> np.ndarray:
    """Normalised Gaussian wave packet on the honeycomb lattice (centre of the
    sample, ``width`` in cells) with a random phase per orbital."""
    g = np.random.default_rng(seed)
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    x = (ix + 0.5 * iy).ravel()
    y = (iy * np.sqrt(3.0) / 2.0).ravel()
    cx, cy = x.mean(), y.mean()
    env = np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2.0 * width ** 2))
    env = np.repeat(env, 2)  # two atoms per cell
    env = np.repeat(env, orbitals)
    phase = np.exp(2j * np.pi * g.random(env.shape[0]))
    psi = env * phase
    return psi / np.linalg.norm(psi)


def random_vbcsr(nb: int, bs_lo: int = 4, bs_hi: int = 8, per_row: int = 40, seed: int = 0) -> BlockCSR:
    """Large variable-block matrix for bandwidth runs (fully vectorised):
    block sizes U{bs_lo..bs_hi}, the diagonal block plus ``per_row`` random
    distinct columns per row, entries U[-1, 1].  Not symmetric."""
    rng = np.random.default_rng(seed)
    bs = rng.integers(bs_lo, bs_hi + 1, size=nb).astype(np.int32)
    rows = np.repeat(np.arange(nb, dtype=np.int64), per_row + 1)
    cols = rng.integers(0, nb, size=(nb, per_row + 1)).astype(np.int64)
    cols[:, 0] = np.arange(nb)
    key = np.unique(rows * nb + cols.ravel())
    r, c = key // nb, key % nb
    row_ptr = np.zeros(nb + 1, np.int32)
    np.add.at(row_ptr, r + 1, 1)
    np.cumsum(row_ptr, out=row_ptr)
    mat = BlockCSR(bs, row_ptr, c.astype(np.int32), np.zeros(0))
    nnz = int(mat.value_offsets()[-1])
    mat.vals = rng.uniform(-1.0, 1.0, size=nnz)
    mat.meta.update(name="random_vbcsr", nb=nb, per_row=per_row, seed=seed)
    return mat


def gershgorin_scale(mat: BlockCSR, edge: float = 0.99) -> BlockCSR:
    """Scale by the Gershgorin bound max_i sum_j |A_ij| so the spectrum lies in
    [-edge, edge] (cheap, no eigen-solve)."""
    rows = np.repeat(np.arange(mat.nb), np.diff(mat.row_ptr))
    r = mat.bs[rows].astype(np.int64)
    c = mat.bs[mat.col_ind].astype(np.int64)
    voff = mat.value_offsets()
    blk = np.repeat(np.arange(mat.nnzb), r * c)
    e = np.arange(mat.nnz) - voff[blk]
    off = mat.offsets()
    srow = off[rows][blk] + e // c[blk]
    rowsum = np.zeros(mat.n)
    np.add.at(rowsum, srow, np.abs(mat.vals))
    bound = float(rowsum.max()) if mat.n else 1.0
    out = mat.scaled(edge / bound)
    out.meta.update(half_width=bound / edge, center=0.0)
    return out
