"""Collision paths the step runs, fed the same uniforms as the oracle.

- The fused electron push (candidacy with the probability-bound prefilter inside the push,
  then the selection kernel, then the kinematics kernel) against the oracle's push followed
  by PERFORM_ELECTRON_NEUTRAL_COLLISIONS in its per-particle form (orc_mcc_en_stream), in
  FP64 (same selections, positions bit for bit, velocities to 1e-12) and in mixed precision (same selections
  for uniforms away from the selection thresholds, velocities to FP32 accuracy).
- The stand-alone mixed-precision e-n MCC against the same oracle.
- The order of one step (pic2d_MainProgram.f90:229-363): electrons emitted at a wall during
  the push are collided in the step, ionisation electrons and the constant emission are not.
- Ionisation products go to the ion species of their process.
"""
import ctypes as C

import numpy as np
import pytest

from edgpu import config
from helpers import oracle_particles, rel_err, sorted_rows, synthetic_field

pytestmark = pytest.mark.gpu

STRIDE = 11  # uniforms per slot of an injected e-n stream


def _table(O, prob, procs=None):
    procs = procs if procs is not None else prob.cross_sections()
    procs = [dict(p, ion_velocity_factor=prob.ion_velocity_factor()) for p in procs]
    sv = [config.ENERGY_SEGMENTS[0]] + [s[0] for s in config.ENERGY_SEGMENTS[1]]
    ss = [s[1] for s in config.ENERGY_SEGMENTS[1]]
    return O.MccTable(sv, ss, procs, prob.M_i_amu, prob.neutral_N_m3, prob.N_subcycles, prob.dt, prob.V_scale)


def _thresholds(O, tab, vx, vy, vz):
    """Cumulative selection probabilities c_k(E) of every process at each electron's energy."""
    en = (vx * vx + vy * vy + vz * vz) * tab.t.energy_factor * config.M_E_KG * 0.5 / config.E_CL
    idx, coll = C.c_int32(), C.c_int32()
    a0, a1 = C.c_double(), C.c_double()
    P = tab.prob2d
    out = np.empty((en.size, P.shape[1]))
    for k, e in enumerate(en):
        O.lib().orc_mcc_select(C.byref(tab.t), float(e), 0.5, C.byref(idx), C.byref(a0), C.byref(a1), C.byref(coll))
        out[k] = a0.value * P[idx.value] + a1.value * P[idx.value + 1]
    return out


def _energetic(prob, e, rng, emax=200.0):
    """Velocities with energies spread over 0..emax eV, isotropic, so every process occurs."""
    n = e[0].size
    vv = rng.standard_normal((3, n))
    vv /= np.linalg.norm(vv, axis=0)
    speed = np.sqrt(rng.uniform(0.0, emax, n) / prob.energy_factor_eV)
    e[2], e[3], e[4] = [np.ascontiguousarray(vv[q] * speed) for q in range(3)]
    return e


@pytest.mark.parametrize("precision", [0, 1])
def test_fused_push_collisions_stream_parity(edgpu, oracle, precision):
    """The production collision path of ADVANCE_ELECTRONS with fuse_collisions=1: the push,
    its candidacy test with the prefilter, k_select and k_candidates, fed uniforms slot by slot
    (words 0-1 candidacy and selection, 2-10 density draw and kinematics), against orc_advance
    then orc_mcc_en_stream on the pushed survivors
    (pic2d_enCollisionsGeneralProc.f90:449-485, pic2d_enCollisionsSpecificProc.f90:61-639)."""
    O = oracle
    prob = config.workload(1, 8)
    print("case fused e-n MCC precision %d: %d electrons, grid %dx%d" % (precision, prob.n_particles, prob.gmi,
                                                                        prob.gmj), flush=True)
    rt = edgpu.Runtime(prob, precision=precision, fuse_collisions=1)
    try:
        rt.initiate_electron_neutral_collisions()
        tab = _table(O, prob)
        rng = np.random.default_rng(31)
        e, _ = oracle_particles(O, prob)
        e = _energetic(prob, e, rng)
        n = e[0].size
        EX, EY = synthetic_field(prob, amp=0.05)
        ref = [a.copy() for a in e]
        g = O.make_geom(prob.geometry())
        rc, rem, _, _ = O.advance(g, O.make_species(True), EX, EY, (prob.Bx, prob.By, prob.Bz), prob.Ez, *ref,
                                  prob.n_boundary_obj)
        assert rc == 0
        keep = rem == 0
        u = rng.uniform(0.0, 1.0, (n, STRIDE))
        # about half of the electrons are candidates; the selection draw spans the whole range
        u[:, 0] = rng.uniform(0.0, 2.0 * tab.t.max_colliding_fraction, n)
        if precision == 1:
            # keep the selection draw 1e-5 away from every threshold of the pushed energy, so
            # that the FP32 push (relative error ~1e-7) cannot change which process is picked
            c = _thresholds(O, tab, ref[2], ref[3], ref[4])
            for _ in range(8):
                near = np.any(np.abs(u[:, 1:2] - c) < 1e-5, axis=1)
                if not near.any():
                    break
                u[near, 1] = rng.uniform(0.0, 1.0, int(near.sum()))
            assert not np.any(np.abs(u[:, 1:2] - c) < 1e-5, axis=1).any()
        survivors = [np.ascontiguousarray(a[keep]) for a in ref]
        u_s = np.ascontiguousarray(u[keep].ravel())
        m = int(keep.sum())
        kind = np.zeros(m, np.int32)
        cnt = np.zeros(8, np.int64)
        nnew = C.c_int64(0)
        enew, inew = np.zeros(5 * m), np.zeros(5 * m)
        v, _ = O.maxwell_tables()
        assert O.lib().orc_mcc_en_stream(C.byref(tab.t), O.dp(v), m, *[O.dp(a) for a in survivors], O.dp(u_s),
                                         O.ip(kind), O.lp(cnt), m, C.byref(nnew), O.dp(enew), O.dp(inew)) == 0
        # the GPU: tags carry the slot so the compacted order can be matched
        rt.upload(0, *e, tag=np.arange(1, n + 1, dtype=np.int32))
        rt.set_grid("EX", EX)
        rt.set_grid("EY", EY)
        rt.set_uniform_stream(u.ravel())
        st = rt.advance_electrons()
        rt.set_uniform_stream(None)
        print("events per process: oracle %s gpu %s, ionisations %d" % (list(cnt[:3]), st["collisions"][:3],
                                                                         nnew.value), flush=True)
        assert list(st["collisions"][:3]) == list(cnt[:3]) and min(cnt[:3]) > 0
        got = rt.download(0)
        assert got[0].size == m
        order = np.argsort(got[5])
        slots = got[5][order] - 1
        assert np.array_equal(slots, np.nonzero(keep)[0])
        # the oracle survivors in slot order against the GPU particles in slot order
        A = np.stack(survivors, axis=1)
        B = np.stack([a[order] for a in got[:5]], axis=1)
        pushed = np.stack([a[keep] for a in ref[2:]], axis=1)
        hit_ref = np.any(A[:, 2:] != pushed, axis=1)
        hit_gpu = np.any(B[:, 2:] != pushed, axis=1)
        assert hit_ref.sum() > 0
        if precision == 0:
            # the same electrons scattered, positions and untouched velocities bit for bit;
            # the scattered velocities to 1e-12 (the device's sin/cos/acos/sqrt/tan/atan and
            # glibc's may differ in the last bit)
            assert np.array_equal(hit_gpu, hit_ref)
            assert np.array_equal(A[:, :2], B[:, :2]) and np.array_equal(A[~hit_ref], B[~hit_ref])
            assert rel_err(A[:, 2:], B[:, 2:]) < 1e-12
        else:
            assert rel_err(A[:, :2], B[:, :2]) < 1e-6
            assert rel_err(A[:, 2:], B[:, 2:]) < 1e-5
        # ionisation products: electrons and ions at the collided positions
        rt.process_added_electrons()
        rt.process_added_ions()
        ge = rt.download(0)
        gi = rt.download(1)
        k = nnew.value
        assert ge[0].size == m + k and gi[0].size == k
        E_ref = sorted_rows(*enew[:5 * k].reshape(k, 5).T)
        I_ref = sorted_rows(*inew[:5 * k].reshape(k, 5).T)
        E_gpu = sorted_rows(*[a[m:] for a in ge[:5]])
        I_gpu = sorted_rows(*gi[:5])
        tol = 1e-12 if precision == 0 else 1e-5
        assert rel_err(E_ref, E_gpu) < tol and rel_err(I_ref, I_gpu) < tol
    finally:
        rt.close()


def test_mixed_en_collisions_stream_parity(edgpu, oracle):
    """PERFORM_ELECTRON_NEUTRAL_COLLISIONS in mixed precision (FP32 velocities, FP64 kinematics)
    fed the oracle's uniforms: the same processes for selection draws away from the thresholds,
    velocities and products to FP32 accuracy."""
    O = oracle
    prob = config.workload(1, 8)
    rt = edgpu.Runtime(prob, precision=1, fuse_collisions=0)
    try:
        rt.initiate_electron_neutral_collisions()
        tab = _table(O, prob)
        rng = np.random.default_rng(37)
        e, _ = oracle_particles(O, prob)
        e = _energetic(prob, e, rng)
        n = e[0].size
        # the FP32 values the device holds, so both sides start from the same velocities
        rt.upload(0, *e, tag=np.arange(1, n + 1, dtype=np.int32))
        start = rt.download(0)
        order0 = np.argsort(start[5])
        e32 = [np.ascontiguousarray(a[order0]) for a in start[:5]]
        u = rng.uniform(0.0, 1.0, (n, STRIDE))
        u[:, 0] = rng.uniform(0.0, 2.0 * tab.t.max_colliding_fraction, n)
        c = _thresholds(O, tab, e32[2], e32[3], e32[4])
        for _ in range(8):
            near = np.any(np.abs(u[:, 1:2] - c) < 1e-5, axis=1)
            if not near.any():
                break
            u[near, 1] = rng.uniform(0.0, 1.0, int(near.sum()))
        u = np.ascontiguousarray(u.ravel())
        rt.set_uniform_stream(u)
        c_gpu = rt.perform_electron_neutral_collisions()
        rt.set_uniform_stream(None)
        ref = [a.copy() for a in e32]
        kind = np.zeros(n, np.int32)
        cnt = np.zeros(8, np.int64)
        nnew = C.c_int64(0)
        enew, inew = np.zeros(5 * n), np.zeros(5 * n)
        v, _ = O.maxwell_tables()
        assert O.lib().orc_mcc_en_stream(C.byref(tab.t), O.dp(v), n, *[O.dp(a) for a in ref], O.dp(u), O.ip(kind),
                                         O.lp(cnt), n, C.byref(nnew), O.dp(enew), O.dp(inew)) == 0
        print("mixed MCC events per process: oracle %s gpu %s" % (list(cnt[:3]), list(c_gpu[:3])), flush=True)
        assert list(c_gpu[:3]) == list(cnt[:3]) and min(cnt[:3]) > 0
        got = rt.download(0)
        order = np.argsort(got[5])
        for q in (2, 3, 4):
            assert rel_err(got[q][order], ref[q]) < 1e-5
        rt.process_added_electrons()
        rt.process_added_ions()
        ge, gi = rt.download(0), rt.download(1)
        k = nnew.value
        assert ge[0].size == n + k and gi[0].size == k
        assert rel_err(sorted_rows(*enew[:5 * k].reshape(k, 5).T), sorted_rows(*[a[n:] for a in ge[:5]])) < 1e-5
        assert rel_err(sorted_rows(*inew[:5 * k].reshape(k, 5).T), sorted_rows(*gi[:5])) < 1e-5
    finally:
        rt.close()


OBJ_ABOVE, OBJ_BELOW = 1, 3


@pytest.mark.parametrize("fuse", [1, 0])
def test_step_collides_wall_emitted_electrons_only(edgpu, oracle, fuse):
    """One whole step with a uniform stream that makes every tested electron collide: the
    electrons of the push and the true secondaries emitted at the wall above during the push
    are collided (PROCESS_ADDED_ELECTRONS precedes PERFORM_ELECTRON_NEUTRAL_COLLISIONS,
    pic2d_MainProgram.f90:264-275), the ionisation electrons and the beam from the wall below
    are not (they are appended after it, :353-363).  Slots below n/2 draw ionisation, the
    others elastic scattering."""
    O = oracle
    prob = config.workload(1, 8)
    see = dict(model_true=1, setD_true=0.5, minE_true_eV=0.0, maxE_true_eV=1000.0, T_true_eV=2.0)
    prob.wall_emission = {"see": {OBJ_ABOVE: see},
                          "constant": {OBJ_BELOW: dict(n_per_step=40.3, model=1, Te_normal_eV=1.0,
                                                       Te_parallel_eV=0.1, We_beam_eV=100.0)}}
    print("case step order fuse %d: %d particles/species" % (fuse, prob.n_particles), flush=True)
    rt = edgpu.Runtime(prob, fuse_collisions=fuse, capacity_factor=1.5)
    try:
        rt.setup_wall_emission()
        rt.initiate_electron_neutral_collisions()
        tab = _table(O, prob)
        rng = np.random.default_rng(41)
        e, i = oracle_particles(O, prob)
        n = e[0].size
        # 101.3 eV electrons moving in every direction: they ionise when selected, and many
        # reach the walls in one step.  (Not 100 eV: an energy within rounding below a point
        # of the energy grid, 60 + 8 x 5 eV here, makes the reference's segment index land one
        # point high and its double check abort, pic2d_enCollisionsGeneralProc.f90:466-474; the
        # device reproduces that check and stops the same way.)
        vv = rng.standard_normal((3, n))
        vv /= np.linalg.norm(vv, axis=0)
        speed = np.sqrt(101.3 / prob.energy_factor_eV)
        e[2], e[3], e[4] = [np.ascontiguousarray(vv[q] * speed) for q in range(3)]
        rt.upload(0, *e)
        rt.upload(1, *i)
        c = _thresholds(O, tab, e[2], e[3], e[4])[0]
        cap = int(rt.params.max_electrons)
        u = rng.uniform(0.05, 0.95, (cap, STRIDE))
        u[:, 0] = 1e-9      # every electron is a candidate
        u[:, 1] = 1e-9      # elastic: below every cumulative probability
        u[:n // 2, 1] = 0.5 * (c[1] + c[2])  # ionisation for the first half of the slots
        u[:, 2] = 0.5       # the density draw
        rt.set_uniform_stream(u.ravel())
        se, si = rt.step()
        rt.set_uniform_stream(None)
        n_see, n_beam = se["emit_count"][OBJ_ABOVE], se["emit_count"][OBJ_BELOW]
        coll = se["collisions"]
        n_after_push = se["n_after"]
        print("pushed %d kept %d, secondaries %d, beam %d, collisions %s" % (n, n_after_push, n_see, n_beam,
                                                                             coll[:3]), flush=True)
        assert n_see > 0 and n_beam >= 40 and coll[2] > 0
        # exactly the pushed survivors and the wall-emitted electrons were collided
        assert sum(coll[:3]) == n_after_push + n_see
        assert rt.count(0) == n_after_push + n_see + n_beam + coll[2]
        assert rt.count(1) == si["n_after"] + coll[2]
    finally:
        rt.close()


def test_ionisation_products_join_their_species(edgpu, oracle):
    """Two ion species; the ionising process produces the second one
    (ion_species_produced, pic2d_enCollisionsSpecificProc.f90:629): its ions are appended to
    that species only."""
    O = oracle
    prob = config.workload(1, 8)
    prob = config.Problem(T_e_eV=2.0, N_plasma_m3=1.0e15, N_of_cells_debye=2, N_max_vel=5, N_of_particles_cell=50,
                          gmi=prob.gmi, gmj=prob.gmj, periodic_x=True,
                          extra_ion_species=[dict(M_i_amu=40.0, Qs=2)], name="two ion species")
    rt = edgpu.Runtime(prob, fuse_collisions=0)
    try:
        procs = prob.cross_sections()
        procs[2] = dict(procs[2], ion_species=1)
        rt.initiate_electron_neutral_collisions(processes=procs)
        tab = _table(O, prob, procs)
        rng = np.random.default_rng(43)
        e, i = oracle_particles(O, prob)
        e = _energetic(prob, e, rng)
        n = e[0].size
        rt.upload(0, *e)
        rt.upload(1, *i)
        rt.upload(2, *[a[: n // 4] for a in i])
        u = rng.uniform(0.0, 1.0, (n, STRIDE))
        u[:, 0] = rng.uniform(0.0, 2.0 * tab.t.max_colliding_fraction, n)
        rt.set_uniform_stream(u.ravel())
        c = rt.perform_electron_neutral_collisions()
        rt.set_uniform_stream(None)
        rt.process_added_ions()
        print("ionisations %d; species 1 %d, species 2 %d ions" % (c[2], rt.count(1), rt.count(2)), flush=True)
        assert c[2] > 0
        assert rt.count(1) == n and rt.count(2) == n // 4 + c[2]
    finally:
        rt.close()
