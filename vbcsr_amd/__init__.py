"""vbcsr_amd -- the vbcsr variable-block sparse hot path on AMD MI355X (gfx950).

Drop-in for the reference package's operator API on one GPU: ``VBCSR``,
``DistVector``, ``DistMultiVector``, ``DistGraph`` and ``AssemblyMode`` keep
the reference's names and semantics (vbcsr/__init__.py), while every product
(block SpMV/SpMM, thresholded SpGEMM, the KPM / Chebyshev recursion) runs as
hand-written HIP kernels on data resident in HBM.

Importing requires the compiled extension (``python -m vbcsr_amd.build``);
there is no CPU fallback.
"""
import os as _os

_HERE = _os.path.dirname(_os.path.abspath(__file__))
if not _os.path.exists(_os.path.join(_HERE, "libvbcsr_hip.so")):
    raise ImportError("vbcsr_amd: libvbcsr_hip.so is not built; run `python -m vbcsr_amd.build` first")

try:
    # Load torch's HIP runtime first when torch is installed: the library then
    # binds to that same runtime (same soname), so device pointers, streams
    # and events are shared with torch instead of living in a second runtime.
    import torch as _torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is optional for the core path
    _torch = None

from . import _core  # noqa: E402  (fails loudly if the HIP extension is missing)
from ._core import AssemblyMode, DistGraph, RedistOp  # noqa: E402
from .matrix import VBCSR  # noqa: E402
from .multivector import DistMultiVector  # noqa: E402
from .vector import DistVector  # noqa: E402

HAS_MPI = False


class _NoMPI:
    COMM_WORLD = None
    SUM = None


MPI = _NoMPI()

__version__ = "0.1.0"

__all__ = [
    "VBCSR",
    "DistVector",
    "DistMultiVector",
    "DistGraph",
    "AssemblyMode",
    "RedistOp",
    "HAS_MPI",
    "MPI",
]
