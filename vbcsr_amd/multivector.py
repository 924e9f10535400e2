"""DistMultiVector: K device-resident columns, row-major with a padded ld.

API of the reference wrapper vbcsr/multivector.py; storage in HBM behind
vbcsr_amd._core.DistMultiVector_* (row-major, ld rounded up to 64 bytes, pad
lanes zero -- vbcsr/core/dist_multivector.hpp:24-67).
"""
from __future__ import annotations

from typing import Any, Union

import numpy as np

from ._device_backed import DeviceBacked
from .vector import DistVector

Scalar = Union[float, complex, int]


class DistMultiVector(DeviceBacked):
    def __init__(self, core_obj: Any, comm: Any = None):
        self._attach(core_obj, comm)

    def _local_buffer(self) -> np.ndarray:
        return self._host_view(self._core)[: self.local_rows, :]

    @property
    def local_rows(self) -> int:
        return self._core.local_rows

    @property
    def ghost_rows(self) -> int:
        return self._core.ghost_rows

    @property
    def num_vectors(self) -> int:
        return self._core.num_vectors

    @property
    def ndim(self) -> int:
        return 2

    @property
    def shape(self):
        return (self.local_rows, self.num_vectors)

    @property
    def size(self) -> int:
        s = self.shape
        return s[0] * s[1] if s[0] is not None else 0

    def __len__(self) -> int:
        s = self.shape[0]
        return s if s is not None else 0

    def __repr__(self) -> str:
        s = self.shape
        rows = s[0] if s[0] is not None else "Unknown"
        return f"<DistMultiVector of shape ({rows}, {s[1]}), dtype={self.dtype}>"

    def reset_comm_stats(self) -> None:
        self._core.reset_comm_stats()

    @property
    def comm_seconds(self) -> float:
        return self._core.comm_seconds

    @property
    def comm_calls(self) -> int:
        return self._core.comm_calls

    def to_numpy(self) -> np.ndarray:
        return self._local_buffer()

    def from_numpy(self, arr: np.ndarray) -> None:
        arr = np.asarray(arr)
        if arr.shape != (self.local_rows, self.num_vectors):
            raise ValueError(
                f"Array shape {arr.shape} mismatch. Expected ({self.local_rows}, {self.num_vectors})")
        self._local_buffer()[:] = arr

    def __array__(self, dtype=None, copy=None) -> np.ndarray:
        a = self.to_numpy()
        return a.astype(dtype) if dtype is not None else a

    def sync_ghosts(self) -> None:
        self._core.sync_ghosts()

    def reduce_ghosts(self) -> None:
        self._core.reduce_ghosts()

    def duplicate(self) -> "DistMultiVector":
        return DistMultiVector(self._core.duplicate(), self.comm)

    def set_constant(self, val: Scalar) -> None:
        self._core.set_constant(val)

    def set_random(self, seed: int, normalize: bool = True) -> None:
        self._core.set_random(seed, normalize)

    def scale(self, alpha: Scalar) -> None:
        self._core.scale(alpha)

    def axpy(self, alpha: Scalar, x: "DistMultiVector") -> None:
        if not isinstance(x, DistMultiVector):
            raise TypeError("axpy expects DistMultiVector")
        self._core.axpy(alpha, x._core)

    def axpby(self, alpha: Scalar, x: "DistMultiVector", beta: Scalar) -> None:
        if not isinstance(x, DistMultiVector):
            raise TypeError("axpby expects DistMultiVector")
        self._core.axpby(alpha, x._core, beta)

    def pointwise_mult(self, other) -> None:
        if isinstance(other, DistMultiVector):
            self._core.pointwise_mult(other._core)
        elif isinstance(other, DistVector):
            self._core.pointwise_mult_vec(other._core)
        else:
            raise TypeError("pointwise_mult expects DistMultiVector or DistVector")

    def bdot(self, other: "DistMultiVector") -> list:
        if not isinstance(other, DistMultiVector):
            raise TypeError("bdot expects DistMultiVector")
        return self._core.bdot(other._core)

    # -- operators
    def __neg__(self):
        res = self.duplicate()
        res._core.scale(-1.0)
        return res

    def __pos__(self):
        return self

    def __add__(self, other):
        res = self.duplicate()
        res += other
        return res

    def __iadd__(self, other):
        if isinstance(other, DistMultiVector):
            self._core.axpy(1.0, other._core)
        elif np.isscalar(other) or isinstance(other, np.ndarray):
            self._local_buffer()[:] += other
        else:
            return NotImplemented
        return self

    def __sub__(self, other):
        res = self.duplicate()
        res -= other
        return res

    def __isub__(self, other):
        if isinstance(other, DistMultiVector):
            self._core.axpy(-1.0, other._core)
        elif np.isscalar(other) or isinstance(other, np.ndarray):
            self._local_buffer()[:] -= other
        else:
            return NotImplemented
        return self

    def __mul__(self, other):
        res = self.duplicate()
        res *= other
        return res

    def __imul__(self, other):
        if np.isscalar(other):
            self._core.scale(other)
        elif isinstance(other, DistMultiVector):
            self._core.pointwise_mult(other._core)
        elif isinstance(other, DistVector):
            self._core.pointwise_mult_vec(other._core)
        elif isinstance(other, np.ndarray):
            self._local_buffer()[:] *= other
        else:
            return NotImplemented
        return self

    def __radd__(self, other):
        return self.__add__(other)

    def __rsub__(self, other):
        res = -self
        res += other
        return res

    def __rmul__(self, other):
        return self.__mul__(other)

    def __truediv__(self, other):
        res = self.duplicate()
        res /= other
        return res

    def __itruediv__(self, other):
        if np.isscalar(other):
            self._core.scale(1.0 / other)
        elif isinstance(other, DistMultiVector):
            self._local_buffer()[:] /= self._host_view(other._core)[: other.local_rows, :]
        elif isinstance(other, DistVector):
            self._local_buffer()[:] /= self._host_view(other._core)[: other.local_size, None]
        elif isinstance(other, np.ndarray):
            self._local_buffer()[:] /= other
        else:
            return NotImplemented
        return self
