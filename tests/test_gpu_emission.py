"""GPU checks of the wall emission against the oracle: electron-induced emission of the
electrons a push absorbs (fed the same uniforms per absorbed slot), the constant (beam)
emission of a wall, and the emission statistics of a multi-step run in both precisions."""
import math

import numpy as np
import pytest

import oracle_ffi as OF
from edgpu import config
from helpers import rel_err, sorted_rows

pytestmark = pytest.mark.gpu

# below wall: model-1 yields, specular reflection; wall above: model-2 yields, random angles
BELOW = dict(model_elastic=1, setD_elastic=0.15, minE_elastic_eV=0.0, maxE_elastic_eV=500.0,
             model_inelastic=1, setD_inelastic=0.1, minE_inelastic_eV=0.0, maxE_inelastic_eV=500.0,
             model_true=1, setD_true=0.5, minE_true_eV=5.0, maxE_true_eV=400.0, T_true_eV=2.0)
ABOVE = dict(model_elastic=2, elastic_reflection=1, E_elast_0_eV=0.0, E_elast_max_eV=5.0, maxD_elast=0.5,
             dE_elast_eV=10.0, frac_elast_highenergy=0.05, model_inelastic=2, frac_inelastic=0.07,
             model_true=2, E_see_0_eV=15.0, E_see_max_eV=150.0, maxD_see_classic=2.5, k_smooth=1.0, T_true_eV=3.0)
OBJ_ABOVE, OBJ_BELOW = 1, 3


def _wall_electrons(prob, n, seed):
    """Electrons within half a cell of either wall, moving into it at 1..300 eV."""
    rng = np.random.default_rng(seed)
    below = rng.uniform(0, 1, n) < 0.5
    x = rng.uniform(0.0, prob.gmi - 1, n)
    y = np.where(below, rng.uniform(0.0, 0.5, n), prob.gmj - rng.uniform(0.0, 0.5, n))
    s = np.sqrt(rng.uniform(1.0, 300.0, n) / prob.energy_factor_eV)
    th = rng.uniform(0.0, 1.3, n)
    ph = rng.uniform(0.0, 2 * np.pi, n)
    vy = s * np.cos(th) * np.where(below, -1.0, 1.0)
    return [x, y, s * np.sin(th) * np.cos(ph), vy, s * np.sin(th) * np.sin(ph)]


@pytest.mark.parametrize("sort", [False, True])
def test_electron_induced_emission_matches_oracle(edgpu, oracle, sort):
    """Same absorbed set, same uniforms per absorbed slot: the same number of electrons per
    wall and the same emitted electrons (positions, velocities, tags) to 1e-12."""
    O = oracle
    prob = config.workload(1, 8)
    rt = edgpu.Runtime(prob, fuse_collisions=0, sort_interval=4 if sort else 0)
    try:
        rt.setup_electron_induced_emission(OBJ_BELOW, **BELOW)
        rt.setup_electron_induced_emission(OBJ_ABOVE, **ABOVE)
        e = _wall_electrons(prob, 20000, 5)
        rt.upload(0, *e)
        if sort:
            rt.sort_particles(0)
            e = rt.download(0)[:5]
        zero = np.zeros(prob.nx * prob.ny)
        rt.set_grid("EX", zero)
        rt.set_grid("EY", zero)
        n = e[0].size
        u = np.random.default_rng(9).uniform(0.0, 1.0, n * 32)
        rt.set_uniform_stream(u)
        st = rt.advance_electrons()
        rt.set_uniform_stream(None)
        n_after = rt.count(0)
        rt.process_added_electrons()
        got = rt.download(0)
        emitted = np.stack(got[:5], axis=1)[n_after:]
        tags = got[5][n_after:]
        # oracle: the same push, then the emission of every absorbed electron in slot order
        g = O.make_geom(prob.geometry())
        ref = [np.array(a, np.float64) for a in e]
        rc, rem, hit_obj, hits = O.advance(g, O.make_species(True), zero, zero, (prob.Bx, prob.By, prob.Bz),
                                           prob.Ez, *ref, prob.n_boundary_obj)
        assert rc == 0
        assert list(st["hit_count"][:4]) == list(hits[:4])
        maxw, inj = O.maxwell_tables()
        mats = {OBJ_BELOW: OF.see_from_material(BELOW, prob.energy_factor_eV, prob.T_e_eV, prob.N_max_vel),
                OBJ_ABOVE: OF.see_from_material(ABOVE, prob.energy_factor_eV, prob.T_e_eV, prob.N_max_vel)}
        rows, rtags, per_obj = [], [], np.zeros(4, np.int64)
        for k in np.nonzero(rem)[0]:
            o = int(hit_obj[k])
            dirflag = 4 if o == OBJ_BELOW else 2
            rc, out = OF.see_emit(mats[o], maxw, inj, dirflag, float(prob.gmi), float(prob.gmj), ref[0][k], ref[1][k],
                                  ref[2][k], ref[3][k], ref[4][k], u[32 * k:32 * k + 32])
            assert rc == 0
            rows.append(out)
            rtags += [o + 1] * out.shape[0]
            per_obj[o] += out.shape[0]
        rows = np.concatenate(rows)
        print("absorbed %d, emitted below %d above %d" % (rem.sum(), per_obj[OBJ_BELOW], per_obj[OBJ_ABOVE]),
              flush=True)
        assert list(st["emit_count"][:4]) == list(per_obj)
        assert per_obj[OBJ_BELOW] > 1000 and per_obj[OBJ_ABOVE] > 1000
        assert emitted.shape == rows.shape
        A = sorted_rows(*rows.T)
        B = sorted_rows(*emitted.T)
        assert rel_err(A, B) < 1e-12
        assert sorted(tags.tolist()) == sorted(rtags)
    finally:
        rt.close()


@pytest.mark.parametrize("model,side", [(1, OBJ_BELOW), (0, OBJ_ABOVE)])
def test_constant_emission_matches_oracle(edgpu, oracle, model, side):
    O = oracle
    prob = config.workload(1, 8)
    rt = edgpu.Runtime(prob, fuse_collisions=0)
    try:
        N = 977.3
        rt.setup_constant_emission(side, N, model=model, Te_normal_eV=1.0, Te_parallel_eV=0.1,
                                   We_beam_eV=100.0)
        rng = np.random.default_rng(11)
        u = rng.uniform(0.0, 1.0, 4 * 1000)
        rt.set_uniform_stream(u)
        emitted = rt.perform_electron_emission()
        rt.set_uniform_stream(None)
        n0 = rt.count(0)
        rt.process_added_electrons()
        got = rt.download(0)
        rows = np.stack(got[:5], axis=1)[n0:]
        fn, fp, vb = OF.emission_factors(model, 1.0, 0.1, 100.0, prob.T_e_eV, prob.N_max_vel)
        maxw, inj = O.maxwell_tables()
        L = prob.gmi - 1
        ref = OF.constant_emission(4 if side == OBJ_BELOW else 2, 1, prob.gmi, prob.gmj, 0, prob.gmi, L, N, model,
                                   fn, fp, vb, maxw, inj, u)
        print("emitted", emitted[side], "oracle", ref.shape[0], flush=True)
        assert emitted[side] == ref.shape[0] and rows.shape == ref.shape
        assert rel_err(sorted_rows(*ref.T), sorted_rows(*rows.T)) < 1e-13
        assert np.all(got[5][n0:] == side + 1)
    finally:
        rt.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_emission_statistics_over_steps(edgpu, precision):
    """Constant yield 0.5 on both walls (true secondaries, all energies) and a 100 eV beam
    from the wall below over 40 steps: secondaries per absorbed electron within 4 standard
    deviations of 0.5, beam electrons per step within 4 standard deviations of N."""
    prob = config.workload(1, 8)
    rt = edgpu.Runtime(prob, precision=precision, seed=77, sort_interval=4)
    try:
        see = dict(model_true=1, setD_true=0.5, minE_true_eV=0.0, maxE_true_eV=1e5, T_true_eV=2.0)
        rt.setup_electron_induced_emission(OBJ_BELOW, **see)
        rt.setup_electron_induced_emission(OBJ_ABOVE, **see)
        N = 12.6
        rt.setup_constant_emission(OBJ_BELOW, N, model=1, Te_normal_eV=1.0, Te_parallel_eV=0.1, We_beam_eV=100.0)
        rt.initiate_electron_neutral_collisions()
        rt.initiate_ion_neutral_collisions()
        rt.load_initial_particles()
        hits = emits = 0
        steps = 40
        for _ in range(steps):
            se, si = rt.step()
            hits += se["hit_count"][OBJ_BELOW] + se["hit_count"][OBJ_ABOVE]
            emits += se["emit_count"][OBJ_BELOW] + se["emit_count"][OBJ_ABOVE]
        beam = steps * N
        see_emits = emits - beam
        print("precision %d: hits %d, emitted %d (beam about %.0f)" % (precision, hits, emits, beam), flush=True)
        # beam: INT(N) per step plus a Bernoulli of the fraction
        sd_beam = math.sqrt(steps * 0.6 * 0.4)
        assert hits > 2000
        ratio = (emits - beam) / hits
        assert abs(ratio - 0.5) < 4 * math.sqrt(0.25 / hits) + 4 * sd_beam / hits
    finally:
        rt.close()


def test_cfg5_wall_physics_steps(edgpu, tmp_path):
    """The combined case reduced 16x per direction (256x128 cells, 100 particles/cell, one
    grounded inner box, Krylov solve): y walls with yield 0.5, the 100 eV beam from the wall
    below.  Every step's wall records go to history_bo_NN.dat; the inner box absorbs without
    emission; the beam injects INT(N) or INT(N)+1 electrons per step."""
    from edgpu import diagnostics as D
    prob = config.workload(5, 16)
    rt = edgpu.Runtime(prob, precision=1, seed=5, sort_interval=4)
    try:
        rt.setup_wall_emission()
        rt.initiate_electron_neutral_collisions()
        rt.initiate_ion_neutral_collisions()
        rt.load_initial_particles()
        nobj = prob.n_boundary_obj + len(prob.inner_boxes)
        D.initiate_history(str(tmp_path), nobj)
        n_beam = prob.wall_emission["constant"][OBJ_BELOW]["n_per_step"]
        tot_hits = tot_emit_see = 0
        for step in range(1, 7):
            se, si = rt.step()
            D.append_history(str(tmp_path), step, se["hit_count"][:nobj], [si["hit_count"][:nobj]],
                             se["emit_count"][:nobj])
            assert se["emit_count"][nobj - 1] == 0  # the inner box does not emit
            assert se["emit_count"][0] == 0 and se["emit_count"][2] == 0  # periodic x: no side walls
            tot_hits += se["hit_count"][OBJ_BELOW] + se["hit_count"][OBJ_ABOVE]
            tot_emit_see += se["emit_count"][OBJ_ABOVE] + se["emit_count"][OBJ_BELOW]
        tot_emit_see -= 6 * n_beam
        # an inner object switches the x-periodic system to the Krylov solve (the reference's
        # PERIODICITY_X_PETSC): the grounded box's nodes hold its potential exactly
        phi = rt.get_grid("phi").reshape(prob.ny, prob.nx)
        b = prob.inner_boxes[0]
        box = phi[b[2]:b[3] + 1, b[0]:b[1] + 1]
        print("inner box potential: max |phi| %.3g, plasma potential max %.3g" % (np.abs(box).max(), phi.max()),
              flush=True)
        assert np.all(box == prob.object_phi[nobj - 1]) and np.abs(phi).max() > 0.0
        recs = D.read_history(str(tmp_path), OBJ_BELOW)
        assert [r[0] for r in recs] == list(range(1, 7))
        assert all(len(r) == 4 for r in recs)
        print("cfg5/16: wall hits %d, secondaries %.0f (beam %.1f per step)" % (tot_hits, tot_emit_see, n_beam),
              flush=True)
        assert tot_hits > 500
        assert abs(tot_emit_see / tot_hits - 0.5) < 4 * math.sqrt(0.25 / tot_hits) + 6 * 0.5 / tot_hits
    finally:
        rt.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_ion_induced_emission_matches_oracle(edgpu, oracle, precision):
    """Ions absorbed by both walls with the coefficient 1.4 inside an energy window that holds
    about two thirds of them, the same uniforms per absorbed slot: the same electrons per wall
    and the same emitted electrons as PROCESS_ION_INDUCED_ELECTRON_EMISSION in the oracle
    (FP64 to 1e-12; mixed precision from FP32 velocities, to 1e-6)."""
    O = oracle
    prob = config.workload(1, 8)
    rt = edgpu.Runtime(prob, precision=precision, fuse_collisions=0)
    try:
        rng = np.random.default_rng(21)
        n = 20000
        below = rng.uniform(0, 1, n) < 0.5
        x = rng.uniform(0.0, prob.gmi - 1, n)
        y = np.where(below, rng.uniform(0.0, 0.3, n), prob.gmj - rng.uniform(0.0, 0.3, n))
        vy = np.where(below, -1.0, 1.0) * rng.uniform(0.4, 1.0, n) / prob.N_subcycles
        vx = rng.uniform(-0.3, 0.3, n) / prob.N_subcycles
        vz = rng.uniform(-0.3, 0.3, n) / prob.N_subcycles
        i = [x, y, vx, vy, vz]
        if precision == 1:
            # the velocities the mixed path stores
            i[2:] = [v.astype(np.float32).astype(np.float64) for v in i[2:]]
        E_eV = prob.Ms * (i[2] ** 2 + i[3] ** 2 + i[4] ** 2) * prob.energy_factor_eV
        lo, hi = np.quantile(E_eV, [0.2, 0.9])
        mat = dict(setD=1.4, minE_eV=float(lo), maxE_eV=float(hi), T_eV=2.0)
        for o in (OBJ_BELOW, OBJ_ABOVE):
            rt.setup_ion_induced_emission(o, 0, **mat)
        rt.upload(1, *i)
        zero = np.zeros(prob.nx * prob.ny)
        for name in ("EX", "EY", "accEX", "accEY"):
            rt.set_grid(name, zero)
        u = np.random.default_rng(23).uniform(0.0, 1.0, n * 32)
        rt.set_uniform_stream(u)
        st = rt.advance_ions()
        rt.set_uniform_stream(None)
        assert rt.count(0) == 0
        rt.process_added_electrons()
        got = rt.download(0)
        emitted = np.stack(got[:5], axis=1)
        g = O.make_geom(prob.geometry())
        ref = [np.array(a, np.float64) for a in i]
        sp = O.make_species(False, Qs=prob.Qs, QM2s=prob.QM2s, N_subcycles=prob.N_subcycles,
                            sense_B=prob.ions_sense_magnetic_field)
        rc, rem, hit_obj, hits = O.advance(g, sp, zero, zero, (prob.Bx, prob.By, prob.Bz), prob.Ez, *ref,
                                           prob.n_boundary_obj)
        assert rc == 0
        assert list(st["hit_count"][:4]) == list(hits[:4])
        maxw, inj = O.maxwell_tables()
        efac = prob.energy_factor_eV
        vinj = np.sqrt(2.0 / prob.T_e_eV) / prob.N_max_vel
        rows, rtags, per_obj = [], [], np.zeros(4, np.int64)
        for k in np.nonzero(rem)[0]:
            o = int(hit_obj[k])
            dirflag = 4 if o == OBJ_BELOW else 2
            # the ion's velocity before its last step is what the push leaves in the slot
            rc, out = OF.iiee_emit(1.4, lo / efac, hi / efac, vinj, prob.Ms, maxw, inj, dirflag, float(prob.gmi),
                                   float(prob.gmj), ref[0][k], ref[1][k], ref[2][k], ref[3][k], ref[4][k],
                                   u[32 * k:32 * k + 32])
            assert rc == 0
            rows.append(out)
            rtags += [o + 1] * out.shape[0]
            per_obj[o] += out.shape[0]
        rows = np.concatenate(rows)
        print("precision %d: absorbed %d, emitted below %d above %d" % (precision, rem.sum(), per_obj[OBJ_BELOW],
                                                                         per_obj[OBJ_ABOVE]), flush=True)
        assert list(st["emit_count"][:4]) == list(per_obj)
        assert per_obj[OBJ_BELOW] > 3000 and per_obj[OBJ_ABOVE] > 3000
        assert emitted.shape == rows.shape
        if precision == 0:
            A, B = sorted_rows(*rows.T), sorted_rows(*emitted.T)
        else:
            # FP32 storage moves the values in their eighth digit: order on rounded keys
            def _ordered(M):
                R = np.round(M, 4)
                return M[np.lexsort((R[:, 4], R[:, 3], R[:, 2], R[:, 0], R[:, 1]))]
            A, B = _ordered(rows), _ordered(emitted)
        assert rel_err(A, B) < (1e-12 if precision == 0 else 1e-6)
        assert sorted(got[5].tolist()) == sorted(rtags)
    finally:
        rt.close()


def test_ion_reflection_matches_oracle(edgpu, oracle):
    """The wall below reflects every ion (ion-wall model 1), the wall above absorbs them: after
    the push and PROCESS_ADDED_IONS the ions that hit the wall below are back, as
    INJECT_REFLECTED_ION in the oracle places them, bit for bit; the hit counts include them and
    no electron is emitted."""
    O = oracle
    prob = config.workload(1, 8)
    rt = edgpu.Runtime(prob, fuse_collisions=0)
    try:
        rng = np.random.default_rng(31)
        n = 20000
        below = rng.uniform(0, 1, n) < 0.5
        x = rng.uniform(0.0, prob.gmi - 1, n)
        y = np.where(below, rng.uniform(0.0, 0.3, n), prob.gmj - rng.uniform(0.0, 0.3, n))
        vy = np.where(below, -1.0, 1.0) * rng.uniform(0.4, 1.0, n) / prob.N_subcycles
        vx = rng.uniform(-0.3, 0.3, n) / prob.N_subcycles
        vz = rng.uniform(-0.3, 0.3, n) / prob.N_subcycles
        i = [x, y, vx, vy, vz]
        rt.setup_ion_reflection(OBJ_BELOW)
        rt.upload(1, *i)
        zero = np.zeros(prob.nx * prob.ny)
        for name in ("EX", "EY", "accEX", "accEY"):
            rt.set_grid(name, zero)
        st = rt.advance_ions()
        rt.process_added_ions()
        got = np.stack(rt.download(1)[:5], axis=1)
        ref = [np.array(a, np.float64) for a in i]
        sp = O.make_species(False, Qs=prob.Qs, QM2s=prob.QM2s, N_subcycles=prob.N_subcycles,
                            sense_B=prob.ions_sense_magnetic_field)
        rc, rem, hit_obj, hits = O.advance(O.make_geom(prob.geometry()), sp, zero, zero, (prob.Bx, prob.By, prob.Bz),
                                           prob.Ez, *ref, prob.n_boundary_obj)
        assert rc == 0 and list(st["hit_count"][:4]) == list(hits[:4])
        keep = [np.stack([a[rem == 0] for a in ref], axis=1)]
        back = [O.ion_reflect(4, float(prob.gmi), float(prob.gmj), ref[0][k], ref[1][k], ref[2][k], ref[3][k],
                              ref[4][k]) for k in np.nonzero(rem)[0] if hit_obj[k] == OBJ_BELOW]
        want = np.concatenate(keep + [np.array(back)])
        print("ions: %d hit below (reflected), %d above (absorbed), %d kept" % (hits[OBJ_BELOW], hits[OBJ_ABOVE],
                                                                              (rem == 0).sum()), flush=True)
        assert len(back) == hits[OBJ_BELOW] > 5000 and hits[OBJ_ABOVE] > 5000
        assert sum(st["emit_count"]) == 0 and rt.count(0) == 0
        assert got.shape == want.shape
        assert rel_err(sorted_rows(*got.T), sorted_rows(*want.T)) < 1e-14
    finally:
        rt.close()
