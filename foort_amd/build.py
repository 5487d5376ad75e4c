"""Build the gfx950 library (libfoort_gpu.so, in-tree) and the CPU oracle.

The HIP sources are compiled with floating-point contraction off: the
finite-difference path must round exactly like the CPU oracle.  Packed FP32
instructions are disabled (-packed-fp32-ops).
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libfoort_gpu.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(REPO, "include"), "-I" + CSRC,
]

SOURCES = ["foort_gpu.hip"]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def build_gpu(extra=(), out=LIB, verbose=False):
    cmd = [hipcc()] + HIPCC_FLAGS + list(extra) + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


def build_oracle(verbose=False):
    cmd = ["make", "-s", "-C", os.path.join(REPO, "oracle"), "-j4"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_all(verbose=False):
    build_gpu(verbose=verbose)
    build_oracle(verbose=verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
