"""Build recipe for the oracle (test infrastructure only).

    python oracle/build.py

1. ``oracle/dg_oracle.c`` -> ``oracle/liboracle.so``: the C restatement of
   the reference's dense CPU algorithm, plain gcc with IEEE double semantics
   (no -ffast-math, no FMA contraction).
2. The same source with ``-DORC_EXTENDED`` -> ``oracle/liboracle_x.so``: the
   same algorithm in x87 extended precision, used to measure how far each
   double evaluation lies from the exact discrete residual.
3. Where the reference sources exist (``/root/reference/solver``, this build
   container only), ``oracle/ref_harness.cpp`` compiled against them, with
   the image's MPICH, into ``oracle/_ref/ref_harness``: the reference's own
   residual and RK steps, used to write the golden vectors under
   ``tests/golden/`` (``tests/golden/make_reference_residuals.py``); and the
   reference's CPU driver itself (solver/dg.cpp) into ``oracle/_ref/dg``,
   whose output files give the driver-level golden vectors
   (``tests/golden/make_reference_driver_outputs.py``).  The reference's
   build system is not used: one g++ line each on its unity sources, as its
   Makefile's ``dg`` target compiles solver/dg.cpp.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "dg_oracle.c")
OUT = os.path.join(HERE, "liboracle.so")
OUTX = os.path.join(HERE, "liboracle_x.so")

REF_SOLVER = "/root/reference/solver"
MPI_PREFIX = "/opt/conda"
REF_OUT_DIR = os.path.join(HERE, "_ref")
REF_HARNESS = os.path.join(REF_OUT_DIR, "ref_harness")
REF_DG = os.path.join(REF_OUT_DIR, "dg")


def _cc(out, defines, verbose):
    cmd = (["gcc", "-std=c99", "-O2", "-fPIC", "-shared", "-ffp-contract=off",
            "-Wall", "-Wno-unused-function"] + ["-D" + d for d in defines]
           + ["-o", out + ".tmp", SRC, "-lm"])
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)


def reference_available() -> bool:
    return (os.path.isfile(os.path.join(REF_SOLVER, "solve.cpp"))
            and os.path.isfile(os.path.join(MPI_PREFIX, "include", "mpi.h")))


def _ref_cxx(src, out, verbose):
    lib = os.path.join(MPI_PREFIX, "lib")
    cmd = ["g++", "-std=c++11", "-O2", "-I" + REF_SOLVER,
           "-I" + os.path.join(MPI_PREFIX, "include"),
           "-o", out + ".tmp", src,
           "-L" + lib, "-Wl,-rpath," + lib, "-lmpicxx", "-lmpi"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)


def build_reference(verbose: bool = False) -> str:
    """Compile the reference harness and the reference's CPU driver into
    oracle/_ref (only where the reference sources are present); returns the
    harness path."""
    os.makedirs(REF_OUT_DIR, exist_ok=True)
    _ref_cxx(os.path.join(HERE, "ref_harness.cpp"), REF_HARNESS, verbose)
    _ref_cxx(os.path.join(REF_SOLVER, "dg.cpp"), REF_DG, verbose)
    return REF_HARNESS


def build(verbose: bool = False) -> str:
    _cc(OUT, [], verbose)
    _cc(OUTX, ["ORC_EXTENDED"], verbose)
    if reference_available():
        build_reference(verbose)
    elif verbose:
        print("reference sources absent: oracle/_ref not built", flush=True)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
