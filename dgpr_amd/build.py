"""Build recipe of libdgx.so (gfx950).

    python -m dgpr_amd.build        (or python dgpr_amd/build.py)

Compiles the HIP kernels and the C ABI with hipcc for gfx950 into
``dgpr_amd/libdgx.so`` in-tree, and the ``dgpr_amd/dgx_drive`` program
(dgx_drive.cpp, host C++) against it.  Packed-FP32 instructions are disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``) for all device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libdgx.so")
DRIVER = os.path.join(HERE, "dgx_drive")
BUILD = os.path.join(HERE, "_build")

SOURCES = ["dgx_kernels.hip", "dgx_ensemble_kernels.hip", "dgx_general.hip", "dgx_api.cpp", "dgx_ensemble.cpp", "dgx_adjoint.cpp",
           "dgx_tables.cpp", "dgx_host.cpp"]
ARCH = os.environ.get("DGX_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(HERE, "..", "include")]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = False, jobs: int = 7, out: str = OUT, defines=()) -> str:
    """Compile every source and link ``out``; ``defines`` are extra -D flags
    (for experiments; the shipped library needs none)."""
    build_dir = BUILD if out == OUT else out + ".objs"
    os.makedirs(build_dir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = (["hipcc", "-x", "hip"] + COMMON + DEVICE + ["-D" + d for d in defines]
               + ["-c", path, "-o", obj])
        if src.endswith(".hip") and os.environ.get("DGX_RESOURCE_USAGE"):
            cmd += ["-Rpass-analysis=kernel-resource-usage"]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError("compile failed")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("compile failed")
    _run(["hipcc", "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", out + ".tmp"] + objs,
         verbose)
    os.replace(out + ".tmp", out)
    if out == OUT:
        build_driver(verbose)
    return out


def build_driver(verbose: bool = False) -> str:
    """Build the ``dgx_drive`` program (host C++ only) next to libdgx.so,
    linked to it with an $ORIGIN run path."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(HERE, "..", "include"),
           os.path.join(CSRC, "dgx_drive.cpp"), "-o", DRIVER + ".tmp", "-L" + HERE, "-ldgx",
           "-Wl,-rpath,$ORIGIN"]
    _run(cmd, verbose)
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
