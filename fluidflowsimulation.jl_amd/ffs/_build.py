"""Builds libffs_hip.so (gfx950) in-tree with hipcc.

The library is compiled from the HIP sources under ../csrc; objects go to
../csrc/_build and the shared library next to this file, so it travels with
the source tree.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB = os.path.join(HERE, "libffs_hip.so")
# (source, defines, object): ffs_kernels.hip compiles in KERNEL_PARTS parts that run in parallel, each with its
# own share of the template instantiations (FFS_PART in that file)
KERNEL_PARTS = 8
UNITS = [("ffs_kernels.hip", [f"-DFFS_PART={p}"], f"ffs_kernels_p{p}.o") for p in range(KERNEL_PARTS)] + [
    ("ffs_sim.hip", [], "ffs_sim.o")
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FFS_OFFLOAD_ARCH", "gfx950")
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wno-unused-value",
    "-Wno-unused-result",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
    # packed FP32 instructions stay off: some of their forms have produced wrong results on gfx950
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose=False):
    """Compiles every HIP source and links libffs_hip.so; returns its path."""
    bdir = os.path.join(CSRC, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    cmds = []
    for src, defs, obj in UNITS:
        o = os.path.join(bdir, obj)
        objs.append(o)
        cmds.append([HIPCC, *CFLAGS, *defs, "-c", os.path.join(CSRC, src), "-o", o])
    with ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out, file=sys.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
