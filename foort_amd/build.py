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

COMMON = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(REPO, "include"), "-I" + CSRC,
]

# (source, contraction): the finite-difference unit must round exactly like the
# CPU oracle; the analytic unit fuses the multiply-adds written within one
# expression ("on": decided from the source, so every kernel that runs the same
# code rounds the same way -- "fast" lets each kernel's code generation decide).
SOURCES = [("foort_gpu.hip", "off"), ("foort_gpu_an.hip", "on")]
BUILD_DIR = os.path.join(PKG, "_build")


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_gpu(extra=(), out=LIB, verbose=False):
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(BUILD_DIR, exist_ok=True)
    jobs = []
    objs = []
    for src, contract in SOURCES:
        obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        jobs.append([hipcc()] + COMMON + ["-ffp-contract=" + contract] + list(extra)
                    + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(len(jobs)) as ex:
        for f in [ex.submit(_run, j, verbose) for j in jobs]:
            f.result()
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", out + ".tmp"], verbose)
    os.replace(out + ".tmp", out)
    return out


def build_host_path(verbose=False):
    """The path's templates compiled for the host (CPU tests of the analytic formulations)."""
    out = os.path.join(PKG, "libfoort_path_host.so")
    _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared",
          "-I" + os.path.join(REPO, "include"), "-I" + CSRC, os.path.join(CSRC, "path_host.cpp"),
          "-o", out], verbose)


def build_opcount(verbose=False):
    """Count the FLOPs of one step of each formulation (foort_amd/flops_per_step.json)."""
    exe = os.path.join(BUILD_DIR, "opcount")
    os.makedirs(BUILD_DIR, exist_ok=True)
    _run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(REPO, "include"), "-I" + CSRC,
          os.path.join(CSRC, "opcount.cpp"), "-o", exe], verbose)
    res = subprocess.run([exe], check=True, capture_output=True, text=True)
    with open(os.path.join(PKG, "flops_per_step.json"), "w") as f:
        f.write(res.stdout)


HOST_SOURCES = ["cfg.cpp", "config.cpp", "mesh.cpp", "output.cpp", "driver.cpp", "server.cpp"]
HOST_LIB = os.path.join(PKG, "libfoort_host.so")
CLI = os.path.join(PKG, "foort")


def build_host(verbose=False):
    """The host side of the drop-in (configuration, meshes, output files, driver):
    libfoort_host.so over libfoort_gpu.so, and the `foort` command line program."""
    hdir = os.path.join(PKG, "host")
    inc = ["-I" + os.path.join(REPO, "include"), "-I" + hdir]
    flags = ["-std=c++17", "-O2", "-fPIC", "-fopenmp", "-Wall", "-Wextra", "-Wno-unused-parameter"]
    _run(["g++"] + flags + inc + ["-shared"] + [os.path.join(hdir, s) for s in HOST_SOURCES]
         + ["-L" + PKG, "-lfoort_gpu", "-Wl,-rpath,$ORIGIN", "-o", HOST_LIB + ".tmp"], verbose)
    os.replace(HOST_LIB + ".tmp", HOST_LIB)
    _run(["g++"] + flags + inc + [os.path.join(hdir, "main.cpp"), "-L" + PKG, "-lfoort_host",
          "-lfoort_gpu", "-Wl,-rpath,$ORIGIN", "-o", CLI + ".tmp"], verbose)
    os.replace(CLI + ".tmp", CLI)


def build_oracle(verbose=False):
    cmd = ["make", "-s", "-C", os.path.join(REPO, "oracle"), "-j4"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_all(verbose=False):
    build_gpu(verbose=verbose)
    build_host(verbose=verbose)
    build_host_path(verbose=verbose)
    build_opcount(verbose=verbose)
    build_oracle(verbose=verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
