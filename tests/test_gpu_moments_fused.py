"""The pushes of snapshot steps, which sum the moments in the same pass over the particles
(ADVANCE_ELECTRONS_AND_CALCULATE_MOMENTS_2D, pic2d_ElectronDynamicsWithVDFMoments.f90:156-1010;
ADVANCE_IONS_AND_CALCULATE_MOMENTS_2D, pic2d_IonDynamicsWithVDFMoments.f90:180-1026): the 13
sums against the oracle's FP32 accumulation (orc_moments_raw) of the particles at their
positions before the step with the half-step velocity 0.5 (v + v') of the oracle's own push,
for electrons and ions, sorted (the tiles' LDS windows) and unsorted (the grid directly); the
particles the fused push leaves are those of the plain push, bit for bit."""
import ctypes as C

import numpy as np
import pytest

from edgpu import config
from helpers import oracle_particles, synthetic_field

pytestmark = pytest.mark.gpu


def _oracle_moments(O, g, prob, before, after):
    """orc_moments_raw at the positions before the step with the half-step velocity"""
    half = [0.5 * (before[q] + after[q]) for q in (2, 3, 4)]
    acc = np.zeros(13 * prob.nx * prob.ny, np.float32)
    O.lib().orc_moments_raw(C.byref(g), before[0].size, O.dp(before[0]), O.dp(before[1]), O.dp(half[0]),
                            O.dp(half[1]), O.dp(half[2]), acc.ctypes.data_as(C.POINTER(C.c_float)))
    return acc.reshape(13, -1)


def _compare(got, ref, tol):
    for a in range(13):
        scale = np.max(np.abs(ref[a])) + 1e-30
        err = np.max(np.abs(got[a] - ref[a])) / scale
        assert err < tol, "moment %d: %.3g" % (a, err)


@pytest.mark.parametrize("precision,sort", [(0, True), (0, False), (1, True)])
def test_fused_moments_match_oracle(edgpu, oracle, precision, sort):
    O = oracle
    prob = config.workload(2, 16)
    print("case fused moments, precision %d, sorted %s: %d particles/species" % (precision, sort, prob.n_particles),
          flush=True)
    rt = edgpu.Runtime(prob, precision=precision, fuse_collisions=0, sort_interval=4 if sort else 0)
    try:
        e, i = oracle_particles(O, prob)
        rng = np.random.default_rng(5)
        i[2], i[3], i[4] = [rng.uniform(-0.02, 0.02, i[0].size) for _ in range(3)]  # ions that move
        EX, EY = synthetic_field(prob, amp=0.05)
        aEX, aEY = synthetic_field(prob, amp=0.5, seed=3)
        for name, val in (("EX", EX), ("EY", EY), ("accEX", aEX), ("accEY", aEY)):
            rt.set_grid(name, val)
        g = O.make_geom(prob.geometry())
        B = (prob.Bx, prob.By, prob.Bz)
        sp_i = O.make_species(False, Qs=prob.Qs, QM2s=prob.QM2s, N_subcycles=prob.N_subcycles,
                              sense_B=prob.ions_sense_magnetic_field)
        tol = 1e-4 if precision == 0 else 1e-3
        for species in (0, 1):
            arrs = e if species == 0 else i
            tags = np.arange(1, arrs[0].size + 1, dtype=np.int32)
            rt.upload(species, *arrs, tag=tags)
            if sort:
                rt.sort_particles(species)
            down = rt.download(species)
            before = down[:5]  # the order the push sees
            if species == 0:
                st, mom = rt.advance_electrons_and_moments()
            else:
                st, mom = rt.advance_ions_and_moments()
                mom = mom[0]
            fused = rt.download(species)
            # the plain push of the same particles (a sort may order a cell differently: by tag)
            rt.upload(species, *arrs, tag=tags)
            if sort:
                rt.sort_particles(species)
            plain_st = rt.advance_electrons() if species == 0 else rt.advance_ions()
            plain = rt.download(species)
            assert st["n_removed"] == plain_st["n_removed"] and list(st["hit_count"]) == list(plain_st["hit_count"])
            fo, po = np.argsort(fused[5]), np.argsort(plain[5])
            for q in range(6):
                assert np.array_equal(fused[q][fo], plain[q][po])
            # the oracle: its push gives v', its accumulation the moments
            after = [a.copy() for a in before]
            sp = O.make_species(True) if species == 0 else sp_i
            O.advance(g, sp, EX if species == 0 else aEX, EY if species == 0 else aEY, B, prob.Ez, *after,
                      prob.n_boundary_obj)
            ref = _oracle_moments(O, g, prob, before, after)
            print("species %d: %d particles, %d removed, N sum %.6g (oracle %.6g)" %
                  (species, before[0].size, st["n_removed"], mom[0].sum(), ref[0].sum()), flush=True)
            _compare(mom, ref, tol)
            assert abs(float(mom[0].sum(dtype=np.float64)) - before[0].size) < 1e-4 * before[0].size
    finally:
        rt.close()


def test_step_with_moments(edgpu, oracle):
    """A snapshot step through edgpu_step_with_moments: the N moment of each species sums to
    the particles pushed, the other sums are finite, and the steps around it run as usual."""
    prob = config.workload(1, 8)
    rt = edgpu.Runtime(prob, sort_interval=4)
    try:
        rt.initiate_electron_neutral_collisions()
        rt.initiate_ion_neutral_collisions()
        rt.load_initial_particles()
        rt.step()
        ne, ni = rt.count(0), rt.count(1)
        se, si, me, mi = rt.step_with_moments()
        assert se["n_before"] == ne and si["n_before"] == ni and mi is not None
        assert abs(float(me[0].sum(dtype=np.float64)) - ne) < 1e-4 * ne
        assert abs(float(mi[0, 0].sum(dtype=np.float64)) - ni) < 1e-4 * ni
        assert np.all(np.isfinite(me)) and np.all(np.isfinite(mi))
        # mean velocity of the thermal plasma is small next to its spread
        assert abs(me[1].sum()) < 0.05 * np.sqrt(me[4].sum() * me[0].sum())
        rt.step()
    finally:
        rt.close()
