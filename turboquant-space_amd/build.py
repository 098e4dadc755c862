"""Build libtq_hip.so (gfx950 kernels + C ABI) and the pybind11 drop-in module.

Both land in-tree next to the Python package (turboquant/), so they travel with
the repository and are what `import turboquant` loads. Run directly or through
__graft_entry__.build().
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
PKG = os.path.join(HERE, "turboquant")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = [
    "tq_capi.hip", "tq_encode.hip", "tq_scan.hip", "tq_lut.hip", "tq_mfma.hip", "tq_search.hip",
    "tq_plan.hip", "tq_serve.hip", "tq_symfull.hip", "tq_noalloc.hip",
]
# -ffp-contract=off: the encoder reproduces the reference's fp32 rounding
# exactly, so every fused multiply-add in the kernels is written explicitly.
# Packed fp32 ops are disabled for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if proc.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd))


def _jobs():
    try:
        n = int(os.environ.get("MAX_JOBS", "0"))
    except ValueError:
        n = 0
    return max(1, min(n or os.cpu_count() or 1, 16, len(HIP_SOURCES)))


def build_hip():
    """Each source compiles to its own object in parallel (every kernel is
    launched from the file that defines it, so no relocatable device code is
    needed); one link makes the shared library."""
    out = os.path.join(PKG, "libtq_hip.so")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"] + ["-c"]
    objs, cmds = [], []
    for s in HIP_SOURCES:
        obj = os.path.join(objdir, os.path.splitext(s)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC] + compile_flags + ["-o", obj, os.path.join(CSRC, s)])
    with ThreadPoolExecutor(_jobs()) as pool:
        errs = [e for e in pool.map(_try_run, cmds) if e]
    if errs:
        raise RuntimeError("build step failed: " + "; ".join(errs))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out] + objs)
    return out


def _try_run(cmd):
    try:
        _run(cmd)
        return None
    except RuntimeError as e:
        return str(e)


def build_module():
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PKG, "_turboquant" + suffix)
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall",
           "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"],
           os.path.join(CSRC, "bindings.cpp"), "-L" + PKG, "-ltq_hip",
           "-Wl,-rpath,$ORIGIN", "-o", out]
    _run(cmd)
    return out


def main():
    build_hip()
    build_module()


if __name__ == "__main__":
    sys.exit(main())
