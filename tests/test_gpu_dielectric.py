"""Dielectric walls below and above an x-periodic domain on the device, against the oracle on
the same inputs:

- SOLVE_POISSON_FFTX_LINSYSY with a dielectric wall below, above or on both sides
  (pic2d_ElectricFieldCalc_FFT_X.f90:100-145, 676-759; tables pic2d_Prepare_FFTX_SYSY.f90:
  776-897) at 1e-12 of the potential's scale, with the surface charge set through the C ABI;
- the surface charge left by the electrons and the ions that one push absorbs
  (pic2d_ElectronWallCollisions.f90:168-195, pic2d_IonWallCollisions.f90:200-219), the periodic
  images exchanged (pic2d_IonWallCollisions.f90:717-766);
- the reference's Hall-thruster deck input_data_sample_periodic_SEE (two dielectric walls with
  electron-induced emission) read by the deck reader and stepped: every absorbed particle and
  every emitted electron is accounted for in the surface charge;
- a time-dependent electrode (UPDATE_WALL_POTENTIALS, pic2d_WallPotentials.f90:1-54) applied by
  Runtime.step before each solve.
"""
import os

import numpy as np
import pytest

from edgpu import config
from helpers import oracle_particles, rel_err, synthetic_field

pytestmark = pytest.mark.gpu

OBJ_ABOVE, OBJ_BELOW = 1, 3
SIDE_DIRFLAG = {0: 1, 1: 2, 2: 3, 3: 4}  # EDGPU_SIDE_* -> the reference's dirflag
DECK = os.path.join(os.path.dirname(__file__), "golden", "deck_periodic_see")


def _problem(eps_below, eps_above, gmi=129, gmj=96, ppc=50):
    prob = config.Problem(T_e_eV=3.0, N_plasma_m3=5.0e15, N_of_cells_debye=2, N_max_vel=5, N_of_particles_cell=ppc,
                          gmi=gmi, gmj=gmj, periodic_x=True, name="dielectric walls")
    eps = [None] * prob.n_boundary_obj
    eps[OBJ_BELOW] = eps_below
    eps[OBJ_ABOVE] = eps_above
    prob.dielectric_eps = eps
    print("case dielectric below %s above %s, grid %dx%d" % (eps_below, eps_above, gmi, gmj), flush=True)
    return prob


def _mask(prob):
    return sum(1 << o for o, e in enumerate(prob.dielectric_eps) if e is not None)


@pytest.mark.parametrize("eps_below,eps_above", [(4.0, None), (None, 2.5), (3.0, 1.5), (1.0, 1.0)])
def test_fftx_dielectric_walls(edgpu, oracle, eps_below, eps_above):
    O = oracle
    prob = _problem(eps_below, eps_above)
    prob.object_phi[OBJ_ABOVE] = 0.2
    prob.object_phi[OBJ_BELOW] = -0.3
    below, above = eps_below is not None, eps_above is not None
    rt = edgpu.Runtime(prob, fuse_collisions=0)
    try:
        e, i = oracle_particles(O, prob)
        rng = np.random.default_rng(3)
        i[0] = np.clip(i[0] + rng.uniform(-1.0, 1.0, i[0].size), prob.x_limits[0], prob.x_limits[1])
        rt.upload(0, *e)
        rt.upload(1, *i)
        rho_e = rt.gather_electron_charge_density()
        rho_i = rt.gather_ion_charge_density()
        g = O.make_geom(prob.geometry())
        a_eq = O.prepare_sys_y_walls(prob.gmi - 1, prob.gmj, below, above, eps_below or 1.0, eps_above or 1.0)
        for trial in range(2):
            sb = rng.standard_normal(prob.nx) * 40.0
            sa = rng.standard_normal(prob.nx) * 40.0
            rt.set_surface_charge(sb if below else None, sa if above else None)
            gb, ga = rt.get_surface_charge()
            # the device keeps the end nodes folded onto their periodic images
            for got, s, on in ((gb, sb, below), (ga, sa, above)):
                want = s.copy() if on else np.zeros(prob.nx)
                if on:
                    O.surface_charge_fold(g, want)
                assert np.array_equal(got, want)
            phi = rt.solve_poisson_fftx()
            ref = O.solve_poisson_fftx_walls(g, prob.N_of_particles_cell, rho_e, rho_i, a_eq, _mask(prob), gb, ga,
                                             below and above)
            err = rel_err(phi, ref)
            print("trial %d: error vs oracle %.2e (|phi| max %.3g)" % (trial, err, np.abs(ref).max()), flush=True)
            assert err < 1e-12
            P = phi.reshape(prob.ny, prob.nx)
            if not below:
                assert np.allclose(P[0, 1:-1], -0.3, rtol=0, atol=1e-12 * np.abs(P).max())
            if not above:
                assert np.allclose(P[-1, 1:-1], 0.2, rtol=0, atol=1e-12 * np.abs(P).max())
    finally:
        rt.close()


def _near_walls(prob, n, seed, vmax=1.4):
    """Particles within 1.5 cells of the walls below and above, moving up to vmax cells a step."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(prob.x_limits[0], prob.x_limits[1], n)
    d = rng.uniform(0.0, 1.5, n)
    y = np.where(rng.random(n) < 0.5, d, prob.gmj - d)
    v = rng.uniform(-vmax, vmax, (3, n))
    return [np.ascontiguousarray(a) for a in (x, y, v[0], v[1], v[2])]


@pytest.mark.parametrize("precision", [0, 1])
def test_surface_charge_from_wall_hits(edgpu, oracle, precision):
    """Two pushes per species; after each the device's surface charge equals the oracle's,
    accumulated from the oracle's own push of the same particles (FP64: 1e-12 of the charge;
    mixed precision: positions in FP32, so 1e-5 per absorbed particle)."""
    O = oracle
    prob = _problem(2.0, 1.0)
    rt = edgpu.Runtime(prob, precision=precision, fuse_collisions=0)
    try:
        g = O.make_geom(prob.geometry())
        nobj = prob.n_boundary_obj
        EX, EY = synthetic_field(prob, amp=0.02)
        for name in ("EX", "accEX"):
            rt.set_grid(name, EX)
        for name in ("EY", "accEY"):
            rt.set_grid(name, EY)
        sb, sa = np.zeros(prob.nx), np.zeros(prob.nx)
        B = (prob.Bx, prob.By, prob.Bz)
        n_abs = 0
        for species, seed in ((0, 1), (1, 2), (0, 3)):
            p = _near_walls(prob, 200000, seed)
            if species == 1:
                p[2:] = [v / prob.N_subcycles for v in p[2:]]
            sp = O.make_species(True) if species == 0 else O.make_species(
                False, Qs=prob.Qs, QM2s=prob.QM2s, N_subcycles=prob.N_subcycles,
                sense_B=prob.ions_sense_magnetic_field)
            rt.upload(species, *p)
            st = rt.advance_electrons() if species == 0 else rt.advance_ions()
            ref = [a.copy() for a in p]
            rc, rem, hit_obj, hits, rx, ry, rd = O.advance_records(g, sp, EX, EY, B, prob.Ez, *ref, nobj)
            assert rc == 0
            assert list(st["hit_count"][:nobj]) == list(hits[:nobj])
            k = np.nonzero(rem)[0]
            n_abs += k.size
            O.surface_charge(g, _mask(prob), hit_obj[k], rd[k], rx[k], -1 if species == 0 else prob.Qs, sb, sa)
            O.surface_charge_fold(g, sb)
            O.surface_charge_fold(g, sa)
            gb, ga = rt.get_surface_charge()
            tol = (1e-12 if precision == 0 else 1e-5) * max(1, n_abs)
            print("species %d: %d absorbed (below %d, above %d); max |sigma| %.1f, differences %.2e %.2e"
                  % (species, k.size, hits[OBJ_BELOW], hits[OBJ_ABOVE], np.abs(sb).max(),
                     np.abs(gb - sb).max(), np.abs(ga - sa).max()), flush=True)
            assert hits[OBJ_BELOW] > 10000 and hits[OBJ_ABOVE] > 10000
            assert np.abs(gb - sb).max() <= tol and np.abs(ga - sa).max() <= tol
            assert gb[0] == 0.0 and gb[-1] == 0.0 and ga[0] == 0.0 and ga[-1] == 0.0
    finally:
        rt.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_periodic_see_deck_charge_balance(edgpu, oracle, precision):
    """The Hall-thruster deck (dielectric walls of BN ceramic, electron-induced emission, argon
    collisions) read from its files and stepped: the surface charge holds exactly the ions
    (+Qs) and electrons (-1) absorbed by the walls and the electrons they emitted (+1); the
    wall potential responds (the rows are not pinned to a metal potential)."""
    from edgpu import deck
    d = deck.read_deck(DECK)
    prob = d.problem()
    assert prob.dielectric_eps[OBJ_ABOVE] == 1.0 and prob.dielectric_eps[OBJ_BELOW] == 1.0
    print("case %s precision %d: grid %dx%d, %d particles/species" % (DECK, precision, prob.gmi, prob.gmj,
                                                                      prob.n_particles), flush=True)
    rt = edgpu.Runtime(prob, precision=precision, seed=prob.seed, sort_interval=4)
    try:
        d.apply(rt)
        rt.load_initial_particles()
        expect = 0.0
        emitted = 0
        for step in range(12):
            se, si = rt.step()
            for o in (OBJ_BELOW, OBJ_ABOVE):
                expect += -se["hit_count"][o] + prob.Qs * si["hit_count"][o] + se["emit_count"][o]
                emitted += se["emit_count"][o]
        gb, ga = rt.get_surface_charge()
        total = gb.sum() + ga.sum()
        print("surface charge %.6f, expected %.0f; electrons emitted %d" % (total, expect, emitted), flush=True)
        assert emitted > 0
        assert abs(total - expect) <= 1e-9 * max(1.0, np.abs(gb).sum() + np.abs(ga).sum())
        assert gb[0] == 0.0 and gb[-1] == 0.0 and ga[0] == 0.0 and ga[-1] == 0.0
        phi = rt.get_grid("phi").reshape(prob.ny, prob.nx)
        assert np.all(np.isfinite(phi))
        assert np.ptp(phi[0, 1:-1]) > 0.0 and np.ptp(phi[-1, 1:-1]) > 0.0
    finally:
        rt.close()


def test_step_applies_electrode_schedule(edgpu):
    """A deck problem with a time-dependent electrode: Runtime.step hands the schedule's
    potential of each step to the device before the solve, so the wall row of phi follows
    phi_const + phi_var sin(omega T_cntr + phase) step by step."""
    import math
    from edgpu import deck as D
    from edgpu import wallpotentials as W
    d = D.read_deck(os.path.join(os.path.dirname(__file__), "golden", "deck_periodic_ebeam"))
    prob = d.problem()
    prob.n_particles = 200_000
    wp = W.WallPotentials(d.dir, [], set(), prob.object_phi, prob.dt, prob.F_scale)
    wp.el[1] = W.Electrode(0.2, 0.5, 2.0 * math.pi / 8.0, 0.3)   # the wall above, an 8-step period
    prob.wall_potentials = wp
    print("case electrode schedule on the wall above, grid %dx%d" % (prob.gmi, prob.gmj), flush=True)
    rt = edgpu.Runtime(prob, seed=prob.seed)
    try:
        d.apply(rt)
        rt.load_initial_particles()
        for T in range(6):
            rt.step()
            top = rt.get_grid("phi").reshape(prob.ny, prob.nx)[-1, 1:-1]
            want = 0.2 + 0.5 * math.sin(2.0 * math.pi / 8.0 * T + 0.3)
            assert np.max(np.abs(top - want)) < 1e-12, (T, top[:3], want)
    finally:
        rt.close()
