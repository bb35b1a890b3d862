"""Writes the golden vectors of tests/golden/ (run in a container that holds the
reference tree; the committed .npy files are what the tests read).

well19937a_seed<S>.npy: the first 4096 uniforms of the reference's own RNG path
(src/WELL19937a_new.c + well_random_seed/well_random_number of
src/pic2d_RandomNumberInterface.f90, compiled from the reference sources into
oracle/_ref/libwellref.so by `make -C oracle ref`) after seeding with S.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

SEEDS = (123456789, 7, 2026)
COUNT = 4096


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    import oracle_ffi
    ref = oracle_ffi.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libwellref.so was not built")
    for s in SEEDS:
        ref.ref_well_seed(s)
        out = np.empty(COUNT, np.float64)
        ref.ref_well_fill(COUNT, out.ctypes.data_as(C.POINTER(C.c_double)))
        np.save(os.path.join(HERE, "well19937a_seed%d.npy" % s), out)
        print("seed %d: first %r" % (s, out[:3].tolist()))


if __name__ == "__main__":
    main()
