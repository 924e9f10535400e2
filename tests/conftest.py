import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="session")
def golden_real():
    return load_golden("vbcsr_real.npz")


@pytest.fixture(scope="session")
def golden_complex():
    return load_golden("bsr_complex.npz")


@pytest.fixture(scope="session")
def golden_kpm():
    return load_golden("kpm_graphene.npz")


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b)) if b.size else 1.0
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0)) if a.size else 0.0


def _ensure_built():
    """Build the extension and the CPU checkers once if a fresh checkout lacks them."""
    lib = os.path.join(ROOT, "vbcsr_amd", "libvbcsr_hip.so")
    orc = os.path.join(ROOT, "oracle", "liboracle.so")
    import glob

    if not os.path.exists(lib) or not glob.glob(os.path.join(ROOT, "vbcsr_amd", "_core*.so")):
        import importlib.util

        # load build.py by path: importing the package itself requires the built library
        spec = importlib.util.spec_from_file_location("_vbcsr_build", os.path.join(ROOT, "vbcsr_amd", "build.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build()
    if not os.path.exists(orc):
        import oracle

        oracle.build()


_ensure_built()
