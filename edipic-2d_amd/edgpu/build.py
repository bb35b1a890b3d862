"""Build libedgpu.so (HIP for gfx950) in-tree with hipcc.

FP64 sources whose arithmetic must round exactly as written (the push, the field
kernels, the moments and the host tables) are compiled with -ffp-contract=off;
the mixed-precision, sort and CNN sources allow FMA contraction.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(PKG))
CSRC = os.path.join(os.path.dirname(PKG), "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build", "edgpu")
LIB = os.path.join(PKG, "libedgpu.so")

ARCH = "gfx950"
COMMON = [
    "-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
    "-I" + INCLUDE, "-I" + CSRC,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-command-line-argument",
]
EXACT = ["-ffp-contract=off"]
FAST = ["-ffp-contract=fast"]
SOURCES = [
    ("api.cpp", EXACT),
    ("particles_fp64.hip", EXACT),
    ("field.hip", EXACT),
    ("moments.hip", EXACT),
    ("particles_mixed.hip", FAST),
    ("sort.hip", FAST),
    ("krylov.hip", EXACT),  # every multigrid path rounds each operation the same way
    ("cnn.hip", FAST),
    ("emission.hip", EXACT),
    ("noalloc.hip", EXACT),
]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False, jobs=4):
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src, extra in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [cc] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = LIB + ".tmp"
    cmd = [cc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    return LIB


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    if p.returncode != 0:
        raise RuntimeError("compiling %s failed:\n%s" % (src, out.decode(errors="replace")))


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
