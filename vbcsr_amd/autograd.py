"""Differentiable Chebyshev propagation: psi(T) = U^nsteps psi0 with
U = exp(-i H dt) as a Chebyshev series, as a torch.autograd.Function whose
backward is hand-written on the HIP library (no dense matrices, no torch
autograd through the sparse product).

The forward follows the reference propagator (examples/ex09_graphene_dynamics.py
:102-133, vbcsr/utils.py:71-126): H~ = (H - center) / half_width,
exp(-i H dt) = exp(-i center dt) * sum_n c_n T_n(H~), c_n = (2 - d_n0)
(-i)^n J_n(half_width dt).  The gradient with respect to the stored block values
and to psi0 uses:

* checkpointed recomputation -- psi is kept every ``checkpoint`` steps, and each
  segment is recomputed once during the backward pass;
* per step, the Chebyshev basis t_n (vbh_cheb_basis), the Clenshaw recurrence
  with H~^H for the cotangents b_n (vbh_cheb_adjoint), and the block-sampled
  outer products sum_n w_n b_n t_{n-1}^H (vbh_sddmm).

Gradients follow torch's complex convention (dL/dRe + i dL/dIm)."""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np
import torch

from . import _core  # noqa: F401  (loads the library with torch's HIP runtime)

_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        L = ctypes.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "libvbcsr_hip.so"))
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.vbh_last_error.restype = ctypes.c_char_p
        L.vbh_matrix_set_values.argtypes = [vp, vp, i, vp]
        L.vbh_cheb_apply.argtypes = [vp, i, vp, i, vp, i, i, vp, d, d, i, vp]
        L.vbh_cheb_basis.argtypes = [vp, i, vp, i, i, vp, i, d, d, i, vp]
        L.vbh_cheb_adjoint.argtypes = [vp, i, vp, i, i, vp, vp, i, d, d, i, vp]
        L.vbh_sddmm.argtypes = [vp, i, vp, i, vp, i, vp, vp, i, vp]
        _LIB = L
    return _LIB


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {_lib().vbh_last_error().decode()}")


def propagator_coefficients(dt: float, center: float, half_width: float, nterms: int) -> np.ndarray:
    """exp(-i H dt) = sum_n c_n T_n(H~), phase exp(-i center dt) folded into c_n."""
    from scipy.special import jv

    n = np.arange(nterms)
    c = (2.0 - (n == 0)) * (-1j) ** n * jv(n, half_width * dt)
    return (np.exp(-1j * center * dt) * c).astype(np.complex128)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class _Ctx:
    """Matrix handle, series and buffers shared by forward and backward."""

    def __init__(self, H, K, nterms, coef, center, half, precision):
        self.H = H
        self.K = K
        self.nterms = nterms
        self.coef = np.ascontiguousarray(coef.view(np.float64))
        self.center = float(center)
        self.half = float(half)
        self.precision = precision
        self.cdtype = torch.complex128 if precision == 0 else torch.complex64

    def load(self, values: torch.Tensor):
        h = self.H._core.handle
        _check(_lib().vbh_matrix_set_values(ctypes.c_void_p(h), ctypes.c_void_p(values.data_ptr()), 1, _stream()),
               "set_values")
        self.H._core.mark_values_written()

    def step(self, psi: torch.Tensor) -> torch.Tensor:
        out = torch.empty_like(psi)
        h = self.H._core.handle
        cp = self.coef.ctypes.data_as(ctypes.c_void_p)
        _check(_lib().vbh_cheb_apply(ctypes.c_void_p(h), self.K, ctypes.c_void_p(psi.data_ptr()), self.K,
                                     ctypes.c_void_p(out.data_ptr()), self.K, self.nterms, cp, self.center,
                                     self.half, self.precision, _stream()), "cheb_apply")
        return out

    def step_backward(self, psi: torch.Tensor, lam: torch.Tensor, G: torch.Tensor) -> torch.Tensor:
        """Accumulate dL/dvalues of one step into G; return dL/dpsi."""
        N, K, nt = psi.shape[0], self.K, self.nterms
        L = _lib()
        h = ctypes.c_void_p(self.H._core.handle)
        s = _stream()
        Tb = torch.empty((N, nt * K), dtype=self.cdtype, device=psi.device)
        B = torch.empty((N, nt * K), dtype=self.cdtype, device=psi.device)
        _check(L.vbh_cheb_basis(h, K, ctypes.c_void_p(psi.data_ptr()), K, nt, ctypes.c_void_p(Tb.data_ptr()),
                                nt * K, self.center, self.half, self.precision, s), "cheb_basis")
        cp = self.coef.ctypes.data_as(ctypes.c_void_p)
        _check(L.vbh_cheb_adjoint(h, K, ctypes.c_void_p(lam.data_ptr()), K, nt, cp, ctypes.c_void_p(B.data_ptr()),
                                  nt * K, self.center, self.half, self.precision, s), "cheb_adjoint")
        esz = Tb.element_size()
        gp = ctypes.c_void_p(G.data_ptr())
        if nt >= 2:  # b_1 t_0^H
            one = np.array([1.0, 0.0])
            _check(L.vbh_sddmm(h, K, ctypes.c_void_p(B.data_ptr() + K * esz), nt * K, ctypes.c_void_p(Tb.data_ptr()),
                               nt * K, one.ctypes.data_as(ctypes.c_void_p), gp, self.precision, s), "sddmm")
        if nt >= 3:  # 2 sum_{n>=2} b_n t_{n-1}^H
            two = np.array([2.0, 0.0])
            _check(L.vbh_sddmm(h, (nt - 2) * K, ctypes.c_void_p(B.data_ptr() + 2 * K * esz), nt * K,
                               ctypes.c_void_p(Tb.data_ptr() + K * esz), nt * K, two.ctypes.data_as(ctypes.c_void_p),
                               gp, self.precision, s), "sddmm")
        return B[:, :K].contiguous()


class ChebPropagation(torch.autograd.Function):
    """psi(T) = [exp(-i H dt)]^nsteps psi0 with H's stored block values taken
    from ``values`` (complex128, slot order, length nnz)."""

    @staticmethod
    def forward(ctx, values, psi0, H, dt, nsteps, nterms, center, half_width, checkpoint, precision):
        coef = propagator_coefficients(dt, center, half_width, nterms)
        c = _Ctx(H, psi0.shape[1], nterms, coef, center, half_width, precision)
        c.load(values)
        psi = psi0.contiguous()
        ckpt = {0: psi}
        for s in range(nsteps):
            psi = c.step(psi)
            if (s + 1) % checkpoint == 0 and s + 1 < nsteps:
                ckpt[s + 1] = psi
        ctx.c = c
        ctx.ckpt = ckpt
        ctx.nsteps = nsteps
        ctx.checkpoint = checkpoint
        ctx.save_for_backward(values)
        return psi

    @staticmethod
    def backward(ctx, grad_out):
        (values,) = ctx.saved_tensors
        c = ctx.c
        c.load(values)  # the matrix may have been changed since the forward pass
        lam = grad_out.to(c.cdtype).contiguous()
        G = torch.zeros(values.shape[0], dtype=torch.complex128, device=values.device)
        starts = sorted(ctx.ckpt)
        for si in range(len(starts) - 1, -1, -1):
            s0 = starts[si]
            s1 = starts[si + 1] if si + 1 < len(starts) else ctx.nsteps
            seg = [ctx.ckpt[s0]]
            for _ in range(s0, s1 - 1):  # recompute the segment's states
                seg.append(c.step(seg[-1]))
            for k in range(s1 - 1, s0 - 1, -1):
                lam = c.step_backward(seg[k - s0], lam, G)
        grad_values = G / c.half
        return grad_values, lam, None, None, None, None, None, None, None, None


def propagate(H, values: torch.Tensor, psi0: torch.Tensor, dt: float, nsteps: int, nterms: int,
              center: float = 0.0, half_width: float = 1.0, checkpoint: int = 10,
              precision: str = "fp64") -> torch.Tensor:
    """Differentiable psi(T) for a complex VBCSR ``H`` (the pattern; values
    come from ``values``).  psi0: (N,) or (N, K) device tensor, complex128 for
    precision "fp64", complex64 for "fp32"."""
    if not H.dtype == np.complex128:
        raise TypeError("propagate needs a complex (complex128) matrix")
    prec = {"fp64": 0, "fp32": 1}[precision]
    want = torch.complex128 if prec == 0 else torch.complex64
    if psi0.dtype != want:
        raise TypeError(f"psi0 must be {want} for precision {precision}")
    if values.dtype != torch.complex128:
        raise TypeError("values must be complex128 (the matrix's storage type)")
    squeeze = psi0.dim() == 1
    p = psi0.reshape(psi0.shape[0], -1)
    out = ChebPropagation.apply(values, p, H, float(dt), int(nsteps), int(nterms), float(center),
                                float(half_width), int(checkpoint), prec)
    return out.reshape(psi0.shape[0]) if squeeze else out
