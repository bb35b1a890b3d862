"""The Fortran binding run end to end: a Fortran driver (tests/fortran/shim_step.f90) keeps its
particles in arrays of the reference's TYPE particle, hands them over with
GPU_PARTICLES_UPLOAD_AOS (c_loc of the first element and the compiler's own stride) and runs
gather, FFT_X solve, field, both pushes and the appends through the edgpu_drop_in wrappers.
Its densities, potential, field, wall hits and particles equal those of the same chain called
through ctypes (to 1e-12: the deposit's atomic sums may add in another order)."""
import ctypes as C
import os

import numpy as np
import pytest

from edgpu import config
from helpers import oracle_particles, rel_err, sorted_rows, synthetic_field

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "libshim_step.so")


def test_fortran_drop_in_step_matches_ctypes(edgpu, oracle):
    if not os.path.exists(LIB):
        pytest.fail("tests/fortran/libshim_step.so is not built (run __graft_entry__.build())")
    prob = config.workload(1, 8)
    print("case Fortran drop-in step cfg1/8: %d particles/species" % prob.n_particles, flush=True)
    e, i = oracle_particles(oracle, prob)
    rng = np.random.default_rng(2)
    i[0] = np.clip(i[0] + rng.uniform(-0.3, 0.3, i[0].size), 0.0, prob.gmi - 1.0)
    params = edgpu.make_params(prob, fuse_collisions=0)
    nobj = prob.n_boundary_obj
    # the chain through ctypes
    rt = edgpu.Runtime(prob, params=params)
    try:
        rt.upload(0, *e)
        rt.upload(1, *i)
        ref = dict(rho_i=rt.gather_ion_charge_density(), rho_e=rt.gather_electron_charge_density(),
                   phi=rt.solve_poisson_fftx())
        ref["EX"], ref["EY"] = rt.calculate_electric_field()
        se = rt.advance_electrons()
        si = rt.advance_ions()
        rt.process_added_electrons()
        rt.process_added_ions()
        pe, pi = rt.download(0), rt.download(1)
    finally:
        rt.close()
    # the same chain through the Fortran wrappers
    L = C.CDLL(LIB)
    n = prob.n_particles
    E = np.asfortranarray(np.stack(e, axis=1))
    I = np.asfortranarray(np.stack(i, axis=1))
    ne, ni = C.c_int32(n), C.c_int32(n)
    grids = {k: np.zeros((prob.ny, prob.nx)) for k in ("rho_e", "rho_i", "phi", "EX", "EY")}
    hits = np.zeros((2, nobj), np.int32)
    stride = C.c_int64()
    D = C.POINTER(C.c_double)
    L.shim_one_step(C.byref(params), C.byref(ne), E.ctypes.data_as(D), C.byref(ni), I.ctypes.data_as(D),
                    prob.nx, prob.ny, *[grids[k].ctypes.data_as(D) for k in ("rho_e", "rho_i", "phi", "EX", "EY")],
                    nobj, hits.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(stride))
    print("TYPE particle stride %d B; electrons %d -> %d, ions -> %d" % (stride.value, n, ne.value, ni.value),
          flush=True)
    assert stride.value >= 44
    for k in grids:
        assert rel_err(grids[k].ravel(), ref[k]) < 1e-12, k
    assert list(hits[0]) == list(se["hit_count"][:nobj]) and list(hits[1]) == list(si["hit_count"][:nobj])
    assert ne.value == pe[0].size and ni.value == pi[0].size
    # the deposits sum with atomics, so the field (and the pushed particles) agree to rounding
    assert rel_err(sorted_rows(*E[:ne.value].T), sorted_rows(*pe[:5])) < 1e-12
    assert rel_err(sorted_rows(*I[:ni.value].T), sorted_rows(*pi[:5])) < 1e-12
