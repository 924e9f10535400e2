"""Chebyshev / KPM helpers with the reference's signatures (vbcsr/utils.py).

``jackson_kernel`` and ``compute_kpm_coeffs`` are host-side coefficient
routines; ``kpm_apply`` is the series apply y = sum_n c_n T_n(H~) v and runs
as one fused device recursion (``BlockSpMat.cheb_apply``) instead of the
reference's per-term mult + axpby + axpy composition (vbcsr/utils.py:71-126).
"""
from __future__ import annotations

import numpy as np

from .multivector import DistMultiVector
from .vector import DistVector


def jackson_kernel(N: int) -> np.ndarray:
    """Jackson damping g_n, n = 0..N-1:
    g_n = [(N - n + 1) cos(pi n/(N+1)) + sin(pi n/(N+1)) cot(pi/(N+1))] / (N+1)."""
    n = np.arange(N)
    q = np.pi / (N + 1)
    return ((N - n + 1) * np.cos(q * n) + np.sin(q * n) / np.tan(q)) / (N + 1)


def fermi_dirac(E, beta, mu):
    from scipy.special import expit

    return expit(-beta * (E - mu))


def gaussian(E, mu):
    return np.exp(-((E - mu) ** 2))


def compute_kpm_coeffs(n_kpm: int, func) -> np.ndarray:
    """Jackson-damped Chebyshev coefficients of ``func`` on [-1, 1] from a
    2n-point cosine sampling and an FFT; c_0 carries the usual 1/2."""
    k = np.arange(2 * n_kpm)
    x = np.cos(np.pi * k / n_kpm)
    c = np.fft.fft(func(x)).real[:n_kpm] / n_kpm
    c = c * jackson_kernel(n_kpm)
    c[0] /= 2.0
    return c


def kpm_apply(matrix, vector, coeffs, scale: float, center: float, precision: str = "fp64"):
    """y = sum_n coeffs[n] T_n((H - center)/scale) v for a DistVector or
    DistMultiVector ``vector`` (the reference's kpm_apply contract)."""
    coeffs = np.asarray(coeffs)
    if coeffs.shape[0] == 0:
        res = vector.duplicate()
        res.set_constant(0.0)
        return res
    prec = {"fp64": 0, "fp32": 1, "mixed": 2}[precision]
    core = matrix._core.cheb_apply(vector._core, np.ascontiguousarray(coeffs, np.complex128), float(center),
                                   float(scale), prec)
    if isinstance(vector, DistMultiVector):
        return DistMultiVector(core, vector.comm)
    return DistVector(core, vector.comm)
