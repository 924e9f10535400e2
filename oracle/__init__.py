"""CPU checkers for the variable-block sparse hot path -- TEST INFRASTRUCTURE.

Two libraries, both loaded with ctypes:

* ``liboracle.so``  -- this repository's scalar restatement of the reference
  semantics (oracle/vbcsr_oracle.cpp), every function citing the reference
  file:line it follows;
* ``_ref/libvbcsr_ref.so`` -- the reference vbcsr C++ core itself, compiled
  from /root/reference/vbcsr/core by oracle/Makefile (present only where the
  reference tree was available at build time; it travels with the tree).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker or the CPU baseline.  The product
(vbcsr_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_ORC_PATH = os.path.join(HERE, "liboracle.so")
_REF_PATH = os.path.join(HERE, "_ref", "libvbcsr_ref.so")

_c_int_p = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and _ref when /root/reference exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "-f", "Makefile", "all"], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


_ORC = None
_REF = None


def lib() -> ctypes.CDLL:
    global _ORC
    if _ORC is None:
        if not os.path.exists(_ORC_PATH):
            build()
        _ORC = ctypes.CDLL(_ORC_PATH)
        _ORC.orc_entry_hash.restype = ctypes.c_uint64
        _ORC.orc_entry_hash.argtypes = [ctypes.c_uint64] * 4
        for s in ("d", "z"):
            getattr(_ORC, f"orc_spgemm_{s}").restype = _vp
            getattr(_ORC, f"orc_filter_{s}").restype = _vp
            getattr(_ORC, f"orc_axpby_{s}").restype = _vp
            getattr(_ORC, f"orc_result_nnzb_{s}").restype = ctypes.c_long
            getattr(_ORC, f"orc_result_nvals_{s}").restype = ctypes.c_long
            for fn in ("orc_result_nnzb", "orc_result_nvals", "orc_result_copy", "orc_result_free"):
                getattr(_ORC, f"{fn}_{s}").argtypes = [_vp] + ([_vp, _vp, _vp] if fn == "orc_result_copy" else [])
    return _ORC


def ref_available() -> bool:
    return os.path.exists(_REF_PATH)


def ref() -> Optional[ctypes.CDLL]:
    global _REF
    if _REF is None and ref_available():
        _REF = ctypes.CDLL(_REF_PATH)
        for s in ("d", "z"):
            for fn in ("create", "spmm", "duplicate"):
                getattr(_REF, f"ref_{fn}_{s}").restype = _vp
            getattr(_REF, f"ref_kind_{s}").restype = ctypes.c_char_p
            getattr(_REF, f"ref_kind_{s}").argtypes = [_vp]
            getattr(_REF, f"ref_nnzb_{s}").restype = ctypes.c_long
            getattr(_REF, f"ref_nnzb_{s}").argtypes = [_vp]
            getattr(_REF, f"ref_nnz_{s}").restype = ctypes.c_long
            getattr(_REF, f"ref_nnz_{s}").argtypes = [_vp]
            getattr(_REF, f"ref_free_{s}").argtypes = [_vp]
            getattr(_REF, f"ref_spmm_{s}").argtypes = [_vp, _vp, ctypes.c_double]
            getattr(_REF, f"ref_filter_{s}").argtypes = [_vp, ctypes.c_double]
            getattr(_REF, f"ref_time_mult_dense_{s}").restype = ctypes.c_double
            getattr(_REF, f"ref_time_mult_{s}").restype = ctypes.c_double
            getattr(_REF, f"ref_kpm_moments_{s}").restype = ctypes.c_double
        _REF.ref_entry_hash.restype = ctypes.c_uint64
        _REF.ref_entry_hash.argtypes = [ctypes.c_uint64] * 4
        _REF.ref_last_error.restype = ctypes.c_char_p
    return _REF


def _p(a: np.ndarray):
    return a.ctypes.data_as(_vp)


def _sfx(dtype) -> str:
    return "z" if np.iscomplexobj(np.zeros(0, dtype)) else "d"


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


class Mat:
    """Minimal block-CSR holder used by the checkers (square block structure
    unless ``cbs`` is given)."""

    def __init__(self, bs, row_ptr, col_ind, vals, cbs=None):
        self.rbs = _i32(bs)
        self.cbs = _i32(cbs if cbs is not None else bs)
        self.row_ptr = _i32(row_ptr)
        self.col_ind = _i32(col_ind)
        self.vals = np.ascontiguousarray(vals)
        self.s = _sfx(self.vals.dtype)

    @classmethod
    def of(cls, m) -> "Mat":
        return cls(m.bs, m.row_ptr, m.col_ind, m.vals)

    @property
    def nbr(self):
        return int(self.rbs.shape[0])

    @property
    def nbc(self):
        return int(self.cbs.shape[0])

    def pat(self):
        return (self.nbr, self.nbc, _p(self.rbs), _p(self.cbs), _p(self.row_ptr), _p(self.col_ind))

    @property
    def voff(self):
        """Value offset of every slot (slot order, row-major blocks)."""
        rows = np.repeat(np.arange(self.nbr), np.diff(self.row_ptr))
        sz = self.rbs[rows].astype(np.int64) * self.cbs[self.col_ind].astype(np.int64)
        return np.concatenate([[0], np.cumsum(sz)])

    def dense(self):
        """Dense copy (small matrices only)."""
        ro = np.concatenate([[0], np.cumsum(self.rbs)])
        co = np.concatenate([[0], np.cumsum(self.cbs)])
        D = np.zeros((ro[-1], co[-1]), self.vals.dtype)
        vo = self.voff
        for i in range(self.nbr):
            for k in range(self.row_ptr[i], self.row_ptr[i + 1]):
                j = self.col_ind[k]
                D[ro[i]:ro[i + 1], co[j]:co[j + 1]] = self.vals[vo[k]:vo[k + 1]].reshape(self.rbs[i], self.cbs[j])
        return D


# ---------------------------------------------------------------- restatement
def spmv(m: Mat, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, m.vals.dtype)
    y = np.zeros(int(m.rbs.sum()), m.vals.dtype)
    getattr(lib(), f"orc_spmv_{m.s}")(*m.pat(), _p(m.vals), _p(x), _p(y))
    return y


def spmm(m: Mat, X: np.ndarray) -> np.ndarray:
    X = np.ascontiguousarray(X, m.vals.dtype)
    K = X.shape[1]
    Y = np.zeros((int(m.rbs.sum()), K), m.vals.dtype)
    getattr(lib(), f"orc_spmm_{m.s}")(*m.pat(), _p(m.vals), K, _p(X), K, _p(Y), K)
    return Y


def spmm_adjoint(m: Mat, X: np.ndarray) -> np.ndarray:
    X = np.ascontiguousarray(X, m.vals.dtype)
    K = X.shape[1]
    Y = np.zeros((int(m.cbs.sum()), K), m.vals.dtype)
    getattr(lib(), f"orc_spmm_adjoint_{m.s}")(*m.pat(), _p(m.vals), K, _p(X), K, _p(Y), K)
    return Y


def block_norms(m: Mat) -> np.ndarray:
    out = np.zeros(m.col_ind.shape[0])
    getattr(lib(), f"orc_block_norms_{m.s}")(*m.pat(), _p(m.vals), _p(out))
    return out


def _take_result(h, s: str, nbr: int, dtype):
    L = lib()
    nnzb = getattr(L, f"orc_result_nnzb_{s}")(h)
    nv = getattr(L, f"orc_result_nvals_{s}")(h)
    rp = np.zeros(nbr + 1, np.int32)
    ci = np.zeros(nnzb, np.int32)
    v = np.zeros(nv, dtype)
    getattr(L, f"orc_result_copy_{s}")(h, _p(rp), _p(ci), _p(v))
    getattr(L, f"orc_result_free_{s}")(h)
    return rp, ci, v


def spgemm(A: Mat, B: Mat, thr: float, upper: bool = False) -> Mat:
    """Thresholded SpGEMM; upper: only block columns >= the row
    (spmm_hermitian_upper, common.hpp:502-513,625)."""
    h = getattr(lib(), f"orc_spgemm_{A.s}")(
        A.nbr, A.nbc, B.nbc, _p(A.rbs), _p(A.cbs), _p(B.cbs), _p(A.row_ptr), _p(A.col_ind),
        _p(A.vals), _p(B.row_ptr), _p(B.col_ind), _p(B.vals), ctypes.c_double(thr), ctypes.c_int(int(upper)))
    rp, ci, v = _take_result(h, A.s, A.nbr, A.vals.dtype)
    return Mat(A.rbs, rp, ci, v, cbs=B.cbs)


def complete_hermitian(U: Mat) -> Mat:
    """U + strict_lower(U^H) for an upper-triangle U (blocks below the
    diagonal ignored); each diagonal block D becomes (D + D^H) * 0.5
    (transpose.hpp:139-200 mirror; block_csr.hpp:1948-1968).  Python loops:
    small matrices only."""
    n = U.nbr
    voff = U.voff
    rows = [[] for _ in range(n)]  # (col, block values)
    for i in range(n):
        for k in range(U.row_ptr[i], U.row_ptr[i + 1]):
            j = int(U.col_ind[k])
            if j < i:
                continue
            r, c = int(U.rbs[i]), int(U.cbs[j])
            blk = U.vals[voff[k]:voff[k + 1]].reshape(r, c)
            if j == i:
                rows[i].append((j, (blk + blk.conj().T) * 0.5))
            else:
                rows[i].append((j, blk.copy()))
                rows[j].append((i, blk.conj().T.copy()))
    rp = np.zeros(n + 1, np.int32)
    ci, vals = [], []
    for i in range(n):
        for j, b in sorted(rows[i], key=lambda t: t[0]):
            ci.append(j)
            vals.append(b.ravel())
        rp[i + 1] = len(ci)
    v = np.concatenate(vals) if vals else np.zeros(0, U.vals.dtype)
    return Mat(U.rbs, rp, np.asarray(ci, np.int32), v.astype(U.vals.dtype), cbs=U.cbs)


def filter_blocks(m: Mat, thr: float) -> Mat:
    h = getattr(lib(), f"orc_filter_{m.s}")(*m.pat(), _p(m.vals), ctypes.c_double(thr))
    rp, ci, v = _take_result(h, m.s, m.nbr, m.vals.dtype)
    return Mat(m.rbs, rp, ci, v, cbs=m.cbs)


def axpby(Y: Mat, alpha, X: Mat, beta) -> Mat:
    a = np.array([alpha], Y.vals.dtype)
    b = np.array([beta], Y.vals.dtype)
    h = getattr(lib(), f"orc_axpby_{Y.s}")(*Y.pat(), _p(Y.vals), _p(X.row_ptr), _p(X.col_ind),
                                          _p(X.vals), _p(a), _p(b))
    rp, ci, v = _take_result(h, Y.s, Y.nbr, Y.vals.dtype)
    return Mat(Y.rbs, rp, ci, v, cbs=Y.cbs)


def shift(m: Mat, alpha) -> Mat:
    v = m.vals.copy()
    a = np.array([alpha], v.dtype)
    diag = np.arange(m.nbr, dtype=np.int32)
    getattr(lib(), f"orc_shift_{m.s}")(*m.pat(), _p(v), _p(a), _p(diag))
    return Mat(m.rbs, m.row_ptr, m.col_ind, v, cbs=m.cbs)


def entry_hash(seed: int, block: int, row: int, col: int) -> int:
    return int(lib().orc_entry_hash(seed, block, row, col))


def rademacher(bs, gids, K: int, seed: int, normalize: bool, n_global: int) -> np.ndarray:
    bs = _i32(bs)
    gids = _i32(gids)
    out = np.zeros((int(bs.sum()), K))
    lib().orc_random_rademacher(int(bs.shape[0]), _p(bs), _p(gids), K, ctypes.c_uint64(seed),
                                int(normalize), ctypes.c_int64(n_global), _p(out), K)
    return out


def kpm_moments(m: Mat, V0: np.ndarray, M: int, mode: int = 1, center: float = 0.0,
                half_width: float = 1.0) -> np.ndarray:
    V0 = np.ascontiguousarray(V0, m.vals.dtype)
    mu = np.zeros(M)
    getattr(lib(), f"orc_kpm_moments_{m.s}")(*m.pat(), _p(m.vals), V0.shape[1], _p(V0), M, mode,
                                             ctypes.c_double(center), ctypes.c_double(half_width), _p(mu))
    return mu


def cheb_apply(m: Mat, X: np.ndarray, coef: np.ndarray, center: float = 0.0,
               half_width: float = 1.0) -> np.ndarray:
    X = np.ascontiguousarray(X, m.vals.dtype)
    squeeze = X.ndim == 1
    if squeeze:
        X = X[:, None].copy()
    coef = np.ascontiguousarray(coef, np.complex128)
    Y = np.zeros_like(X)
    getattr(lib(), f"orc_cheb_apply_{m.s}")(*m.pat(), _p(m.vals), X.shape[1], _p(X), coef.shape[0],
                                            _p(coef), ctypes.c_double(center),
                                            ctypes.c_double(half_width), _p(Y))
    return Y[:, 0] if squeeze else Y


def jackson(M: int) -> np.ndarray:
    """Jackson damping g_n, n = 0..M-1 (examples/ex08_graphene_dos.py:145-146)."""
    n = np.arange(M)
    return ((M - n + 1) * np.cos(np.pi * n / (M + 1))
            + np.sin(np.pi * n / (M + 1)) / np.tan(np.pi / (M + 1))) / (M + 1)


# ------------------------------------------------------------------ reference
class RefMat:
    """A matrix held by the compiled reference core (serial, one rank)."""

    def __init__(self, m: Mat, _handle=None):
        self.L = ref()
        if self.L is None:
            raise RuntimeError("reference build oracle/_ref/libvbcsr_ref.so is absent")
        self.s = m.s if m is not None else None
        self.dtype = m.vals.dtype if m is not None else None
        if _handle is not None:
            self.h = _handle
            return
        self.h = getattr(self.L, f"ref_create_{m.s}")(m.nbr, _p(m.rbs), _p(m.row_ptr), _p(m.col_ind))
        if not self.h:
            raise RuntimeError(self.L.ref_last_error().decode())
        rc = getattr(self.L, f"ref_set_values_{m.s}")(_vp(self.h), _p(m.vals))
        if rc != 0:
            raise RuntimeError(self.L.ref_last_error().decode())
        self.rbs = m.rbs

    def __del__(self):
        if getattr(self, "h", None) and self.L is not None:
            getattr(self.L, f"ref_free_{self.s}")(self.h)
            self.h = None

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(self.L.ref_last_error().decode())

    @property
    def kind(self) -> str:
        return getattr(self.L, f"ref_kind_{self.s}")(self.h).decode()

    def mult(self, x):
        x = np.ascontiguousarray(x, self.dtype)
        y = np.zeros(int(self.rbs.sum()), self.dtype)
        self._chk(getattr(self.L, f"ref_mult_{self.s}")(_vp(self.h), _p(x), _p(y)))
        return y

    def mult_adjoint(self, x):
        x = np.ascontiguousarray(x, self.dtype)
        y = np.zeros(int(self.rbs.sum()), self.dtype)
        self._chk(getattr(self.L, f"ref_mult_adjoint_{self.s}")(_vp(self.h), _p(x), _p(y)))
        return y

    def mult_dense(self, X, adjoint=False):
        X = np.ascontiguousarray(X, self.dtype)
        Y = np.zeros_like(X)
        fn = "ref_mult_dense_adjoint_" if adjoint else "ref_mult_dense_"
        self._chk(getattr(self.L, fn + self.s)(_vp(self.h), X.shape[1], _p(X), _p(Y)))
        return Y

    def time_mult_dense(self, X, iters):
        X = np.ascontiguousarray(X, self.dtype)
        return getattr(self.L, f"ref_time_mult_dense_{self.s}")(_vp(self.h), X.shape[1], _p(X), iters)

    def time_mult(self, x, iters):
        x = np.ascontiguousarray(x, self.dtype)
        return getattr(self.L, f"ref_time_mult_{self.s}")(_vp(self.h), _p(x), iters)

    def kpm_moments(self, V0, M):
        V0 = np.ascontiguousarray(V0, self.dtype)
        mu = np.zeros(M)
        t = getattr(self.L, f"ref_kpm_moments_{self.s}")(_vp(self.h), V0.shape[1], _p(V0), M, _p(mu))
        return mu, t

    def _wrap(self, h):
        if not h:
            raise RuntimeError(self.L.ref_last_error().decode())
        out = RefMat(None, _handle=h)
        out.s, out.dtype, out.rbs = self.s, self.dtype, self.rbs
        return out

    def spmm(self, B: "RefMat", thr: float) -> "RefMat":
        return self._wrap(getattr(self.L, f"ref_spmm_{self.s}")(_vp(self.h), _vp(B.h), thr))

    def duplicate(self) -> "RefMat":
        return self._wrap(getattr(self.L, f"ref_duplicate_{self.s}")(_vp(self.h)))

    def filter(self, thr: float):
        self._chk(getattr(self.L, f"ref_filter_{self.s}")(_vp(self.h), thr))

    def scale(self, a):
        a = np.array([a], self.dtype)
        self._chk(getattr(self.L, f"ref_scale_{self.s}")(_vp(self.h), _p(a)))

    def shift(self, a):
        a = np.array([a], self.dtype)
        self._chk(getattr(self.L, f"ref_shift_{self.s}")(_vp(self.h), _p(a)))

    def axpby(self, alpha, X: "RefMat", beta):
        a = np.array([alpha], self.dtype)
        b = np.array([beta], self.dtype)
        self._chk(getattr(self.L, f"ref_axpby_{self.s}")(_vp(self.h), _p(a), _vp(X.h), _p(b)))

    def block_norms(self):
        out = np.zeros(self.nnzb)
        self._chk(getattr(self.L, f"ref_block_norms_{self.s}")(_vp(self.h), _p(out)))
        return out

    @property
    def nnzb(self) -> int:
        return int(getattr(self.L, f"ref_nnzb_{self.s}")(self.h))

    def export(self) -> Mat:
        nbr = int(self.rbs.shape[0])
        rp = np.zeros(nbr + 1, np.int32)
        ci = np.zeros(self.nnzb, np.int32)
        getattr(self.L, f"ref_get_pattern_{self.s}")(_vp(self.h), _p(rp), _p(ci))
        nnz = int(getattr(self.L, f"ref_nnz_{self.s}")(self.h))
        v = np.zeros(nnz, self.dtype)
        getattr(self.L, f"ref_get_values_{self.s}")(_vp(self.h), _p(v))
        return Mat(self.rbs, rp, ci, v)

    def random_mv(self, K, seed, normalize=True):
        out = np.zeros((int(self.rbs.sum()), K), self.dtype)
        self._chk(getattr(self.L, f"ref_random_mv_{self.s}")(_vp(self.h), K, ctypes.c_uint64(seed),
                                                              int(normalize), _p(out)))
        return out
