"""Plumbing shared by the DistVector and DistMultiVector wrappers.

Both wrap a core object of vbcsr_amd._core whose buffer lives in HBM.  The
core reports its scalar type (``is_complex``), and the buffer protocol of the
core hands out a host view, synchronised from the device when it is taken.
This build holds one rank's share, so the global extent of a vector is its
owned (local) extent.
"""
from __future__ import annotations

from typing import Any

import numpy as np


class DeviceBacked:
    """Base of the vector wrappers: the core object, its communicator and dtype."""

    _core: Any
    comm: Any
    dtype: np.dtype

    def _attach(self, core_obj: Any, comm: Any) -> None:
        self._core = core_obj
        self.comm = comm
        self.dtype = np.dtype(np.complex128 if core_obj.is_complex else np.float64)

    @staticmethod
    def _host_view(core_obj: Any) -> np.ndarray:
        """All rows (owned and ghost) of a core vector as a host array."""
        return np.asarray(memoryview(core_obj))

    def copy(self):
        """A new wrapper of the same class over a deep copy of the data."""
        return type(self)(self._core.duplicate(), self.comm)
