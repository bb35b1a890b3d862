"""Build libedgpu.so (HIP for gfx950) in-tree with hipcc.

Sources whose arithmetic must round exactly as written (the FP64 push, the field
kernels, the moments, the host tables, and the mixed-precision push, whose fused
multiply-adds are written out so that the plain and the sorting push give the same
bits) are compiled with -ffp-contract=off; the sort and CNN sources allow FMA
contraction.
"""
import os
import re
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
    ("particles_mixed.hip", EXACT),  # FMAs written out: every kernel inlining the push rounds alike
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


# sources whose kernels' register and scratch use is kept in BUILD/<name>.resources
# (the compiler's kernel-resource-usage remarks)
RESOURCE_REPORTS = ("particles_fp64.hip", "particles_mixed.hip", "sort.hip")


def build(verbose=False, jobs=4):
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src, extra in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [cc] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src in RESOURCE_REPORTS:
            cmd.append("-Rpass-analysis=kernel-resource-usage")
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
    if src in RESOURCE_REPORTS:
        with open(os.path.join(BUILD, os.path.splitext(src)[0] + ".resources"), "w") as f:
            f.write(out.decode(errors="replace"))


def kernel_resources(src):
    """{kernel symbol: {"VGPRs": .., "ScratchSize": .., "Occupancy": ..}} from the last build's
    report of src (one of RESOURCE_REPORTS)."""
    path = os.path.join(BUILD, os.path.splitext(src)[0] + ".resources")
    out = {}
    with open(path) as f:
        blocks = f.read().split("Function Name: ")[1:]
    for b in blocks:
        name = b.split()[0]
        rec = {}
        for key, pat in (("VGPRs", r"VGPRs: (\d+)"), ("ScratchSize", r"ScratchSize \[bytes/lane\]: (\d+)"),
                         ("Occupancy", r"Occupancy \[waves/SIMD\]: (\d+)")):
            m = re.search(pat, b)
            if m:
                rec[key] = int(m.group(1))
        out[name] = rec
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
