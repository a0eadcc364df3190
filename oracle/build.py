"""Build recipe for the oracle (CPU FP64 restatement of the reference hot path).

Test infrastructure: produces oracle/_build/liboracle.so, loaded only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg.

The reference itself (Fortran 2008 + MPI) is not built here: it needs an
`mpi` module for the installed Fortran compiler, which this image lacks.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "liboracle.so")


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11", "-ffp-contract=off",
           "-fopenmp", "-Wall", "-Wno-unused-variable",
           os.path.join(HERE, "swm_oracle.c"), "-o", LIB + ".tmp", "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
