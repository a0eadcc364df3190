"""Build the device library libswm_gpu.so in-tree (gfx950) and the Fortran host driver.

Host set-up code (geometry.cpp) is compiled by the host C++ compiler, the HIP sources by
hipcc for gfx950 only; both with -ffp-contract=off so that FP64 results follow the
reference's operation order without fused multiply-adds.  The Fortran binding module
(fortran/swm_gpu_mod.f90) and the namelist driver (fortran/swm_gpu_driver.f90) are compiled
with amdflang into _build/swm_gpu_driver, linked against libswm_gpu.so by rpath.
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
HIP_SOURCES = ["swm_capi.hip", "stage_o21.hip"] + [f"{k}_o{o}{s}.hip" for o in (42, 43, 63)
                                                for k, s in (("stage", ""), ("stage", "_f32"), ("ens", ""))]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(BUILD, "swm_gpu_driver")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    geo_o = os.path.join(BUILD, "geometry.o")
    jobs = [["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, "geometry.cpp"), "-o", geo_o]]
    objs = [geo_o]
    for src in HIP_SOURCES:
        o = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    # independent translation units: compile in parallel
    procs = []
    for cmd in jobs:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
    for cmd, pr in procs:
        if pr.wait() != 0:
            raise subprocess.CalledProcessError(pr.returncode, cmd)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"], verbose)
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


def build_fortran(verbose: bool = False) -> str:
    """Compile the ISO_C_BINDING module and the namelist driver (needs libswm_gpu.so)."""
    os.makedirs(BUILD, exist_ok=True)
    mod_o = os.path.join(BUILD, "swm_gpu_mod.o")
    _run([FC, "-O2", "-J", BUILD, "-c", os.path.join(FORTRAN, "swm_gpu_mod.f90"), "-o", mod_o], verbose)
    _run([FC, "-O2", "-J", BUILD, os.path.join(FORTRAN, "swm_gpu_driver.f90"), mod_o, "-L" + HERE,
          "-lswm_gpu", "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/..", "-o", DRIVER], verbose)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
