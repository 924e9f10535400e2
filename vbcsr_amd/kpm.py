"""Native KPM: moments, batched moments and the DOS reconstruction.

The reference computes KPM moments in user-level Python loops over
``H @ V``, ``bdot``, ``axpby`` and ``scale``
(examples/ex08_graphene_dos.py:66-156, examples/inp_nanocrystal_dos/kpm.py:
46-164).  Here the whole recursion runs on the device: one fused kernel per
step (SpMM + three-term update + FP64 dot partials), optionally two moments
per matvec, optionally in FP32 with FP64 reductions, and -- for many
Hamiltonians -- one launch per step over the whole batch, captured in one HIP
graph.
"""
from __future__ import annotations

from typing import Sequence, Tuple

import numpy as np

from . import _core
from .matrix import VBCSR
from .multivector import DistMultiVector
from .utils import jackson_kernel

_MODES = {"plain": 0, "one": 1, "two": 2}
# fp64: FP64 throughout; fp32: FP32 vectors and values; mixed: FP32 values, FP64 vectors
_PREC = {"fp64": 0, "fp32": 1, "mixed": 2}


def kpm_moments(H: VBCSR, V0: DistMultiVector, num_moments: int, mode: str = "two",
                precision: str = "fp64", center: float = 0.0, half_width: float = 1.0) -> np.ndarray:
    """mu_n = sum_v Re <v0_v, T_n(H~) v0_v>, n < num_moments, H~ = (H - center)/half_width.

    mode "plain": separate SpMM / axpby / dot kernels, one moment per matvec
    (the unfused reference composition, kept as a baseline); "one": fused,
    one moment per matvec; "two": fused, two moments per matvec (needs a
    Hermitian H).  precision "fp32" runs the recursion in FP32 (vectors and
    matrix values), "mixed" with FP32 matrix values and FP64 vectors; the dot
    products are FP64 in every precision.  Columns beyond 64 are processed in
    groups of 64."""
    return np.asarray(H._core.kpm_moments(V0._core, int(num_moments), _MODES[mode], float(center),
                                          float(half_width), _PREC[precision]))


def kpm_moments_batched(Hs: Sequence[VBCSR], V0s: Sequence[DistMultiVector], num_moments: int,
                        mode: str = "two", precision: str = "fp64", center: float = 0.0,
                        half_width: float = 1.0, graph: bool = True) -> np.ndarray:
    """Moments of many independent Hamiltonians (rows: matrices), one kernel
    launch per recursion step over the whole batch, steps in one HIP graph."""
    mats = [h._core for h in Hs]
    vecs = [v._core for v in V0s]
    return np.asarray(_core.kpm_moments_batched(mats, vecs, int(num_moments), _MODES[mode], float(center),
                                                float(half_width), _PREC[precision], bool(graph)))


def set_autotune(on: bool) -> None:
    """Tile autotuning of kpm_moments_batched (on by default): the first call
    on a batch times one recursion step per candidate plan (matrix-core
    threshold x slot partitions) and keeps the fastest for that batch."""
    _core.set_autotune(bool(on))


def set_batch_plan(threshold: int = 0, parts: int = 1) -> None:
    """Pin the tile plan of later batched calls (class threshold 1..17, the
    class rows on the matrix cores, 17 = no class rows; -16..-1: class rows
    of -threshold..16 rows on the block-row wave kernel, FP32 values and K in
    16, 32, 64; parts > 0: lane-family slot partitions, parts < 0: the
    row-wave kernel with -parts rows per work-group); threshold 0 releases
    it."""
    _core.set_batch_plan(int(threshold), int(parts))


def last_plan() -> Tuple[int, int, bool]:
    """(class threshold (negative: block-row wave kernel), slot partitions or
    -(row-wave rows per work-group), tuned) of the last batched call."""
    return tuple(_core.kpm_last_plan())


def spectral_bounds(H: VBCSR, num_iter: int = 120, seed: int = 20250101) -> Tuple[float, float]:
    """Extreme eigenvalues of a Hermitian H by Lanczos on the device.

    Plain three-term Lanczos from a normalised Rademacher start vector (the
    matrix-vector products and dots run through the library; only the
    tridiagonal coefficients reach the host), then the extreme eigenvalues of
    the tridiagonal matrix.  Ritz values lie inside the spectrum and converge
    to its ends from within, so callers map them to a little less than +-1
    (KPM scales to [-0.99, 0.99]); the reference examples estimate the bounds
    by Gershgorin discs (examples/inp_nanocrystal_dos/run.py:156) or by hand
    (examples/ex08_graphene_dos.py:72-76)."""
    from scipy.linalg import eigh_tridiagonal

    n = H.shape[0]
    if n <= 1:
        d = H.to_dense() if n else np.zeros((0, 0))
        w = np.linalg.eigvalsh(d) if n else np.zeros(1)
        return float(w[0]), float(w[-1])
    v = H.create_vector()
    v.set_random(seed, normalize=True)
    v_prev = H.create_vector()
    v_prev.set_constant(0.0)
    w = H.create_vector()
    alpha, beta = [], []
    b = 0.0
    for _ in range(min(num_iter, n)):
        H.mult(v, w)
        a = float(np.real(w.dot(v)))
        w.axpby(-a, v, 1.0)  # w <- w - a v - b v_prev
        if b != 0.0:
            w.axpby(-b, v_prev, 1.0)
        alpha.append(a)
        b = float(np.sqrt(max(np.real(w.dot(w)), 0.0)))
        if b <= 1e-12 * max(1.0, abs(a)):
            break
        beta.append(b)
        w.scale(1.0 / b)
        v_prev, v, w = v, w, v_prev
    k = len(alpha)
    ev = eigh_tridiagonal(np.asarray(alpha), np.asarray(beta[: k - 1]), eigvals_only=True)
    return float(ev[0]), float(ev[-1])


def scale_to_unit(H: VBCSR, edge: float = 0.99, bounds=None, num_iter: int = 120) -> Tuple[float, float]:
    """Map H's spectrum onto [-edge, edge] in place (H <- (H - c)/s * edge via
    scale and shift on the device); returns the bounds used."""
    lo, hi = bounds if bounds is not None else spectral_bounds(H, num_iter=num_iter)
    center, half = 0.5 * (lo + hi), 0.5 * (hi - lo)
    H.shift(-center)
    H.scale(edge / half)
    return lo, hi


def chebyshev_nodes(num_points: int) -> np.ndarray:
    """Ascending Gauss-Chebyshev nodes x_k = -cos(pi (k + 1/2)/K); never +-1."""
    return -np.cos(np.pi * (np.arange(num_points) + 0.5) / num_points)


def dos_from_moments(moments: np.ndarray, num_points: int = 2000, kernel: str = "jackson",
                     x: np.ndarray = None) -> Tuple[np.ndarray, np.ndarray]:
    """rho(x) = [g_0 mu_0 + 2 sum_n g_n mu_n T_n(x)] / (pi sqrt(1 - x^2)) on
    Gauss-Chebyshev nodes (examples/ex08_graphene_dos.py:145-155 with the node
    choice of examples/inp_nanocrystal_dos/kpm.py:135-151)."""
    mu = np.asarray(moments, dtype=np.float64)
    M = mu.shape[0]
    g = jackson_kernel(M) if kernel == "jackson" else np.ones(M)
    if x is None:
        x = chebyshev_nodes(num_points)
    t0 = np.ones_like(x)
    rho = g[0] * mu[0] * t0
    if M > 1:
        t1 = x.copy()
        rho = rho + 2.0 * g[1] * mu[1] * t1
        for n in range(2, M):
            t0, t1 = t1, 2.0 * x * t1 - t0
            rho = rho + 2.0 * g[n] * mu[n] * t1
    return x, rho / (np.pi * np.sqrt(1.0 - x * x))


def dos_error(rho: np.ndarray, rho_ref: np.ndarray) -> float:
    """max_x |rho - rho_ref| / max_x rho_ref (both on the same interior nodes)."""
    return float(np.max(np.abs(rho - rho_ref)) / np.max(np.abs(rho_ref)))
