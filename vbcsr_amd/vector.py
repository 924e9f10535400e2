"""DistVector: a device-resident vector over one rank's local rows.

API of the reference wrapper vbcsr/vector.py (owned entries + ghost tail,
numpy views of the owned part, arithmetic operators); the storage lives in HBM
behind vbcsr_amd._core.DistVector_* and is synchronised to the host only when
a numpy view is taken.
"""
from __future__ import annotations

from typing import Any, Union

import numpy as np

from ._device_backed import DeviceBacked

Scalar = Union[float, complex, int]


class DistVector(DeviceBacked):
    def __init__(self, core_obj: Any, comm: Any = None):
        self._attach(core_obj, comm)

    def _local_buffer(self) -> np.ndarray:
        return self._host_view(self._core)[: self.local_size]

    # -- sizes
    @property
    def local_size(self) -> int:
        return self._core.local_size

    @property
    def ghost_size(self) -> int:
        return self._core.ghost_size

    @property
    def full_size(self) -> int:
        return self._core.full_size

    @property
    def ndim(self) -> int:
        return 1

    @property
    def shape(self):
        return (self.local_size,)

    @property
    def size(self) -> int:
        s = self.shape[0]
        return s if s is not None else 0

    @property
    def T(self) -> "DistVector":
        return self

    def __len__(self) -> int:
        return self.size

    def __repr__(self) -> str:
        s = self.shape[0]
        return f"<DistVector of size {s if s is not None else 'Unknown'}, dtype={self.dtype}>"

    # -- diagnostics (no exchange happens at one rank)
    def reset_comm_stats(self) -> None:
        self._core.reset_comm_stats()

    @property
    def comm_seconds(self) -> float:
        return self._core.comm_seconds

    @property
    def comm_calls(self) -> int:
        return self._core.comm_calls

    # -- host interop
    def to_numpy(self) -> np.ndarray:
        return self._local_buffer()

    def from_numpy(self, arr: np.ndarray) -> None:
        arr = np.asarray(arr)
        if arr.size != self.local_size:
            raise ValueError(f"Array size {arr.size} does not match vector local size {self.local_size}")
        self._local_buffer()[:] = arr.reshape(-1)

    def __array__(self, dtype=None, copy=None) -> np.ndarray:
        a = self.to_numpy()
        return a.astype(dtype) if dtype is not None else a

    # -- ops
    def sync_ghosts(self) -> None:
        self._core.sync_ghosts()

    def reduce_ghosts(self) -> None:
        self._core.reduce_ghosts()

    def duplicate(self) -> "DistVector":
        return DistVector(self._core.duplicate(), self.comm)

    def set_constant(self, val: Scalar) -> None:
        self._core.set_constant(val)

    def set_random(self, seed: int, normalize: bool = True) -> None:
        self._core.set_random(seed, normalize)

    def scale(self, alpha: Scalar) -> None:
        self._core.scale(alpha)

    def axpy(self, alpha: Scalar, x: "DistVector") -> None:
        if not isinstance(x, DistVector):
            raise TypeError("axpy expects DistVector")
        self._core.axpy(alpha, x._core)

    def axpby(self, alpha: Scalar, x: "DistVector", beta: Scalar) -> None:
        if not isinstance(x, DistVector):
            raise TypeError("axpby expects DistVector")
        self._core.axpby(alpha, x._core, beta)

    def pointwise_mult(self, other: "DistVector") -> None:
        if not isinstance(other, DistVector):
            raise TypeError("pointwise_mult expects DistVector")
        self._core.pointwise_mult(other._core)

    def dot(self, other: "DistVector") -> Scalar:
        if not isinstance(other, DistVector):
            raise TypeError("Dot product requires DistVector")
        return self._core.dot(other._core)

    def __matmul__(self, other):
        if isinstance(other, DistVector):
            return self.dot(other)
        return NotImplemented

    # -- operators
    def __add__(self, other):
        res = self.duplicate()
        res += other
        return res

    def __iadd__(self, other):
        if isinstance(other, DistVector):
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
        if isinstance(other, DistVector):
            self._core.axpy(-1.0, other._core)
        elif np.isscalar(other) or isinstance(other, np.ndarray):
            self._local_buffer()[:] -= other
        else:
            return NotImplemented
        return self

    def __neg__(self):
        res = self.duplicate()
        res._core.scale(-1.0)
        return res

    def __pos__(self):
        return self

    def __mul__(self, other):
        res = self.duplicate()
        res *= other
        return res

    def __imul__(self, other):
        if np.isscalar(other):
            self._core.scale(other)
        elif isinstance(other, DistVector):
            self._core.pointwise_mult(other._core)
        elif isinstance(other, np.ndarray):
            self._local_buffer()[:] *= other
        else:
            return NotImplemented
        return self

    def __rmul__(self, other):
        return self.__mul__(other)

    def __truediv__(self, other):
        res = self.duplicate()
        res /= other
        return res

    def __itruediv__(self, other):
        if np.isscalar(other):
            self._core.scale(1.0 / other)
        elif isinstance(other, DistVector):
            self._local_buffer()[:] /= self._host_view(other._core)[: other.local_size]
        elif isinstance(other, np.ndarray):
            self._local_buffer()[:] /= other
        else:
            return NotImplemented
        return self
