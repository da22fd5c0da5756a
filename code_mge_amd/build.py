"""Build step for the HIP library (gfx950) and, for tests, the CPU oracle.

The library is built in-tree as code_mge_amd/lib/libmge.so so that it
travels with the repository snapshot to the GPU host.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmge.so")

HIP_FLAGS = [
    "-O3", "-fPIC", "-shared", "--offload-arch=gfx950", "-std=c++17",
    # the FP64 path reproduces the reference arithmetic: no contraction
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_library(verbose=True):
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, "mge_device.hip"), os.path.join(CSRC, "mge_geom.cpp"),
            os.path.join(CSRC, "mge_input.cpp")]
    tmp = LIB + ".tmp"
    _run(["hipcc", *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, *srcs])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-B", "liboracle.so"], cwd=odir)
    return os.path.join(odir, "liboracle.so")


def build_fortran_driver():
    """Fortran mgsol driver (amdflang) bound to libmge.so via iso_c_binding."""
    fdir = os.path.join(HERE, "fortran")
    if not os.path.exists(os.path.join(fdir, "Makefile")):
        return None
    _run(["make", "-s", "-B"], cwd=fdir)
    return os.path.join(fdir, "mgsol")


if __name__ == "__main__":
    what = sys.argv[1:] or ["lib", "oracle"]
    if "lib" in what:
        build_library()
    if "oracle" in what:
        build_oracle()
    if "fortran" in what:
        build_fortran_driver()
