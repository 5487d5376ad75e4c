"""Build the gfx950 library (libfoort_gpu.so, in-tree) and the CPU oracle.

The HIP sources are compiled with floating-point contraction off: the
finite-difference path must round exactly like the CPU oracle.  Packed FP32
instructions are disabled (-packed-fp32-ops).
"""
import os
import re
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


def _run(cmd, verbose, capture=False):
    if verbose:
        print(" ".join(cmd), flush=True)
    if capture:
        r = subprocess.run(cmd, stderr=subprocess.PIPE, text=True)
        if r.returncode:
            sys.stderr.write(r.stderr)
            raise subprocess.CalledProcessError(r.returncode, cmd)
        return r.stderr
    subprocess.run(cmd, check=True)
    return ""


RESOURCES = os.path.join(REPO, "profiles", "kernel_resources.md")
_REMARK = re.compile(r"remark:\s+(Function Name|TotalSGPRs|VGPRs|AGPRs|ScratchSize \[bytes/lane\]|"
                     r"Occupancy \[waves/SIMD\]|SGPRs Spill|VGPRs Spill|LDS Size \[bytes/block\]):\s*(\S+)")


def _census(remarks):
    """Kernel resource usage from the compiler's kernel-resource-usage remarks: one dict per kernel."""
    out, cur = [], None
    for line in remarks.splitlines():
        m = _REMARK.search(line)
        if not m:
            continue
        key, val = m.groups()
        if key == "Function Name":
            cur = {"name": val}
            out.append(cur)
        elif cur is not None:
            cur[key] = val
    return out


def _demangle(names):
    try:
        r = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True, check=True)
        return r.stdout.splitlines()
    except (OSError, subprocess.CalledProcessError):
        return list(names)


def write_census(units, path=RESOURCES):
    """profiles/kernel_resources.md: registers, scratch and occupancy of every kernel, as the
    compiler reports them for this build (units: [(source, remarks text)])."""
    rows = []
    for src, text in units:
        ks = _census(text)
        for k, dn in zip(ks, _demangle([k["name"] for k in ks])):
            dn = re.sub(r"\(.*$", "", dn.replace("foort::", "").replace("void ", ""))
            dn = dn.replace("(anonymous namespace)::", "")
            rows.append((src, dn, k))
    spill = [r for r in rows if int(r[2].get("ScratchSize [bytes/lane]", "0")) > 0]
    lines = ["# Kernel resources of this build (every kernel of libfoort_gpu.so)", "",
             "Registers, scratch memory and occupancy as the compiler reports them, written by the build",
             "(`python foort_amd/build.py`, which compiles with `-Rpass-analysis=kernel-resource-usage`).",
             "Template arguments: right-hand side policy `<metric>` (0 Kerr, 1 FlatSpace, 2 Rasheed-Larsen,",
             "3 Johannsen, 4 Manko-Novikov), integrator (0 RK4, 1 Verlet), and `true` when every termination",
             "and diagnostic updates on every step (no update counters).", "",
             f"{len(rows)} kernels; {len(spill)} use scratch memory (register spills to memory):", ""]
    for src, dn, k in spill:
        lines.append(f"* `{dn}`: {k.get('ScratchSize [bytes/lane]')} B/lane "
                     f"({k.get('VGPRs Spill')} VGPR, {k.get('SGPRs Spill')} SGPR spills), "
                     f"{k.get('Occupancy [waves/SIMD]')} waves/SIMD")
    lines += ["", "| unit | kernel | VGPRs | AGPRs | SGPRs | scratch B/lane | waves/SIMD | VGPR spills | SGPR spills |",
              "|---|---|---|---|---|---|---|---|---|"]
    for src, dn, k in rows:
        lines.append(f"| {src} | `{dn}` | {k.get('VGPRs')} | {k.get('AGPRs')} | {k.get('TotalSGPRs')} | "
                     f"{k.get('ScratchSize [bytes/lane]')} | {k.get('Occupancy [waves/SIMD]')} | "
                     f"{k.get('VGPRs Spill')} | {k.get('SGPRs Spill')} |")
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return rows


def build_gpu(extra=(), out=LIB, verbose=False, census=True):
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(BUILD_DIR, exist_ok=True)
    jobs = []
    objs = []
    for src, contract in SOURCES:
        obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        jobs.append([hipcc()] + COMMON + ["-ffp-contract=" + contract, "-Rpass-analysis=kernel-resource-usage"]
                    + list(extra) + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(len(jobs)) as ex:
        remarks = [f.result() for f in [ex.submit(_run, j, verbose, True) for j in jobs]]
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", out + ".tmp"], verbose)
    os.replace(out + ".tmp", out)
    if census and out == LIB and not extra:
        write_census([(src, r) for (src, _), r in zip(SOURCES, remarks)])
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
