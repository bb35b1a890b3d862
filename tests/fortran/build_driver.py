"""Builds tests/fortran/libshim_step.so: the Fortran binding (edipic-2d_amd/fortran/edgpu_mod.F90)
and the test driver shim_step.f90, compiled with amdflang and linked with gcc against the Flang
runtime and libedgpu.so.  Returns the path, or None when amdflang or its runtime is absent."""
import glob
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BUILD = os.path.join(ROOT, "_build", "fortran")
LIB = os.path.join(HERE, "libshim_step.so")


def flang():
    return shutil.which("amdflang") or next((c for c in ("/opt/rocm/bin/amdflang",) if os.path.exists(c)), None)


def flang_runtime():
    hits = sorted(glob.glob("/opt/rocm*/lib/llvm/lib/clang/*/lib/x86_64-unknown-linux-gnu/libflang_rt.runtime.a"))
    return hits[-1] if hits else None


def build():
    fc, rt = flang(), flang_runtime()
    if fc is None or rt is None:
        return None
    os.makedirs(BUILD, exist_ok=True)
    shim = os.path.join(ROOT, "edipic-2d_amd", "fortran", "edgpu_mod.F90")
    o1, o2 = os.path.join(BUILD, "edgpu_mod.o"), os.path.join(BUILD, "shim_step.o")
    subprocess.run([fc, "-c", "-fPIC", "-cpp", "-module-dir", BUILD, shim, "-o", o1], check=True)
    subprocess.run([fc, "-c", "-fPIC", "-module-dir", BUILD, "-I", BUILD, os.path.join(HERE, "shim_step.f90"),
                    "-o", o2], check=True)
    libdir = os.path.join(ROOT, "edipic-2d_amd", "edgpu")
    tmp = LIB + ".tmp"
    subprocess.run(["gcc", "-shared", "-fPIC", "-o", tmp, o2, o1, rt, "-L" + libdir, "-ledgpu",
                    "-Wl,-rpath,$ORIGIN/../../edipic-2d_amd/edgpu", "-lm", "-lstdc++"], check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
