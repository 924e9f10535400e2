"""Build the gfx950 library and the pybind host module in-tree.

    python -m vbcsr_amd.build        (or __graft_entry__.build())

Produces, next to this file:
  libvbcsr_hip.so   HIP kernels + the C ABI of include/vbcsr_hip.h (hipcc, gfx950)
  _core*.so         the pybind11 host module (g++), linked against libvbcsr_hip.so
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
HIP_SOURCES = ["matrix.hip", "apply.hip", "vector.hip", "kpm.hip", "spgemm.hip", "adjoint.hip",
               "lanes_f64.hip", "lanes_f32.hip", "lanes_c128.hip", "lanes_c64.hip", "lanes_f64m.hip",
               "lanes_c128m.hip"]
LIB = os.path.join(HERE, "libvbcsr_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # code-object v5: loadable by this ROCm's runtime and by the one bundled with torch
    "-mcode-object-version=5",
    # no packed-FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in VGPRs: the matrix-core row loop branches on the
    # block width between MFMAs, and AGPR accumulators then cost a copy to
    # and from VGPRs around every k-step
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wno-unused-result",
]


def core_path() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "_core" + suffix)


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose: bool = False, jobs: int = 8) -> None:
    os.makedirs(OBJ, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)

    import pybind11

    inc = [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}"]
    core = core_path()
    tmp = core + ".tmp"
    _run(["g++", "-O2", "-shared", "-fPIC", "-std=c++17", *inc, os.path.join(CSRC, "pybind_core.cpp"),
          "-o", tmp, f"-L{HERE}", "-lvbcsr_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, core)
    if verbose:
        print("built", LIB, "and", core)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
