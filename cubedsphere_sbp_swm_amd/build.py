"""Build the device library libswm_gpu.so in-tree (gfx950).

Host set-up code (geometry.cpp) is compiled by the host C++ compiler, the HIP sources by
hipcc for gfx950 only; both with -ffp-contract=off so that FP64 results follow the
reference's operation order without fused multiply-adds.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libswm_gpu.so")
ARCH = os.environ.get("SWM_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-function"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    geo_o = os.path.join(BUILD, "geometry.o")
    capi_o = os.path.join(BUILD, "swm_capi.o")
    _run(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, "geometry.cpp"), "-o", geo_o], verbose)
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "swm_capi.hip"), "-o", capi_o], verbose)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, geo_o, capi_o, "-lpthread"], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
