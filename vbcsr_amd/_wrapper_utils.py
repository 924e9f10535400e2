"""Small shared helpers of the Python wrappers (vbcsr/_wrapper_utils.py)."""
from __future__ import annotations

from typing import Any, Optional

import numpy as np


def infer_dtype_from_core(core_obj: Any) -> np.dtype:
    return np.dtype(np.complex128) if "Complex" in type(core_obj).__name__ else np.dtype(np.float64)


def reduced_extent(comm: Any, local_extent: int) -> Optional[int]:
    """Global extent: the local one at a single rank, else an allreduce."""
    if comm is None:
        return int(local_extent)
    if hasattr(comm, "allreduce"):
        try:
            return int(comm.allreduce(int(local_extent)))
        except Exception:
            return None
    return None


def core_buffer(core_obj: Any) -> np.ndarray:
    """Host view of a core vector (synchronised from the device on export)."""
    return np.asarray(memoryview(core_obj))


def duplicate_wrapper(wrapper: Any) -> Any:
    dup = wrapper.duplicate()
    dup.comm = getattr(wrapper, "comm", None)
    return dup
