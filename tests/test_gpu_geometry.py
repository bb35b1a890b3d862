"""Field-solve geometries and ion species the headline workload does not exercise, against the
oracle on the same inputs:

- CALCULATE_ELECTRIC_FIELD of a walled domain with inner boxes
  (pic2d_ElectricFieldCalc_PETSc.F90:360-440), bit for bit;
- the Krylov solve of an X-periodic domain with an inner box and biased walls (the reference's
  PERIODICITY_X_PETSC, the combined case's geometry) against the oracle's CG;
- FFT_X with biased walls, and UPDATE_WALL_POTENTIALS (pic2d_WallPotentials.f90:1-54) between
  solves on both solver paths;
- ions of charge 2, 3 and -1 beside singly charged ones: push and charge deposit
  (pic2d_IonDynamics.f90:1653-1686).
"""
import ctypes as C

import numpy as np
import pytest

from edgpu import config
from helpers import oracle_particles, rel_err, sorted_rows, synthetic_field

pytestmark = pytest.mark.gpu

OBJ_ABOVE, OBJ_BELOW = 1, 3


def _problem(n, scale):
    prob = config.workload(n, scale)
    print("case cfg%d scale %d: %d particles/species, grid %dx%d" % (n, scale, prob.n_particles, prob.gmi, prob.gmj),
          flush=True)
    return prob


def _charged(O, prob, seed):
    """Oracle particles with the ions displaced so that the charge density is not neutral."""
    e, i = oracle_particles(O, prob)
    rng = np.random.default_rng(seed)
    i[0] = np.clip(i[0] + rng.uniform(-1.0, 1.0, i[0].size), prob.x_limits[0], prob.x_limits[1])
    return e, i


def _oracle_cg(O, prob, rho_e, rho_i, rtol=1e-12):
    g = O.make_geom(prob.geometry())
    b = np.zeros_like(rho_e)
    O.lib().orc_krylov_rhs(C.byref(g), prob.N_of_particles_cell, O.dp(rho_e), O.dp(rho_i), O.dp(b))
    x = np.zeros_like(rho_e)
    res = np.zeros(1)
    it = O.lib().orc_krylov_cg(C.byref(g), O.dp(b), O.dp(x), rtol, 200000, O.dp(res))
    assert it >= 0 and res[0] <= rtol
    return x


def test_walled_electric_field(edgpu, oracle):
    """The field of a walled box with two inner boxes (one-sided differences at the walls),
    accumulated twice, equals the oracle's bit for bit."""
    O = oracle
    prob = _problem(4, 16)
    rt = edgpu.Runtime(prob, fuse_collisions=0)
    try:
        rng = np.random.default_rng(51)
        phi = rng.standard_normal(prob.nx * prob.ny)
        rt.set_grid("phi", phi)
        rt.clear_accumulated_fields()
        g = O.make_geom(prob.geometry())
        acc = (np.zeros_like(phi), np.zeros_like(phi))
        for _ in range(2):
            EX, EY = rt.calculate_electric_field()
            rEX, rEY = O.efield(g, phi, prob.factor1_E_from_F, prob.factor2_E_from_F, acc)
            assert np.array_equal(EX, rEX) and np.array_equal(EY, rEY)
        assert np.array_equal(rt.get_grid("accEX"), acc[0]) and np.array_equal(rt.get_grid("accEY"), acc[1])
    finally:
        rt.close()


@pytest.mark.parametrize("scale", [16, 8])
@pytest.mark.parametrize("precond", [1, 0])
def test_periodic_krylov_against_oracle(edgpu, oracle, scale, precond):
    """The combined case's X-periodic domain with an inner box (so the Krylov solve, as
    PERIODICITY_X_PETSC), here with biased walls and a biased box: the GPU PCG (multigrid or
    Jacobi) to 1e-12 against the oracle CG to 1e-12 on the same charge density."""
    O = oracle
    prob = _problem(5, scale)
    prob.object_phi[OBJ_ABOVE] = 0.3
    prob.object_phi[OBJ_BELOW] = -0.2
    prob.object_phi[-1] = 0.5
    rt = edgpu.Runtime(prob, fuse_collisions=0)
    try:
        rt.set_krylov_preconditioner(precond)
        e, i = _charged(O, prob, 53)
        rt.upload(0, *e)
        rt.upload(1, *i)
        rho_e = rt.gather_electron_charge_density()
        rho_i = rt.gather_ion_charge_density()
        phi, it, rr = rt.solve_potential_krylov(rtol=1e-12, max_iter=20000)
        ref = _oracle_cg(O, prob, rho_e, rho_i)
        err = rel_err(phi, ref)
        # the oracle's operator applied to the GPU potential: a wrong periodic stencil or a
        # wrong Dirichlet value shows as an O(1) residual whatever the grid size
        g = O.make_geom(prob.geometry())
        b = np.zeros_like(phi)
        O.lib().orc_krylov_rhs(C.byref(g), prob.N_of_particles_cell, O.dp(rho_e), O.dp(rho_i), O.dp(b))
        Ax = np.zeros_like(phi)
        O.lib().orc_krylov_apply(C.byref(g), O.dp(phi), O.dp(Ax))
        res = np.linalg.norm(b - Ax) / np.linalg.norm(b)
        print("periodic Krylov %dx%d precond %d: %d iterations, relres %.2e, oracle-operator residual %.2e, "
              "error vs oracle CG %.2e" % (prob.gmi, prob.gmj, precond, it, rr, res, err), flush=True)
        assert rr <= 1e-12 and res < 1e-10 and err < 1e-7
        P = phi.reshape(prob.ny, prob.nx)
        b = prob.inner_boxes[0]
        assert np.all(P[b[2]:b[3] + 1, b[0]:b[1] + 1] == 0.5)
        assert np.all(P[0, :] == -0.2) and np.all(P[-1, :] == 0.3)
        # the periodic images hold the values of the columns they repeat
        assert np.array_equal(P[:, prob.gmi - 1], P[:, 0]) and np.array_equal(P[:, prob.gmi], P[:, 1])
        # and the walled field of that potential (X-periodic differences, walls in y)
        rt.clear_accumulated_fields()
        EX, EY = rt.calculate_electric_field()
        rEX, rEY = O.efield(O.make_geom(prob.geometry()), phi, prob.factor1_E_from_F, prob.factor2_E_from_F)
        assert np.array_equal(EX, rEX) and np.array_equal(EY, rEY)
    finally:
        rt.close()


def test_fftx_biased_walls_and_update(edgpu, oracle):
    """FFT_X with the wall above at +0.7 and the wall below at -0.4 (normalised), then
    UPDATE_WALL_POTENTIALS to new values before the next solve; each solve against the
    oracle's on the same density, and whole steps run after the update."""
    O = oracle
    prob = _problem(1, 8)
    prob.object_phi[OBJ_ABOVE] = 0.7
    prob.object_phi[OBJ_BELOW] = -0.4
    rt = edgpu.Runtime(prob, fuse_collisions=0)
    try:
        e, i = _charged(O, prob, 57)
        rt.upload(0, *e)
        rt.upload(1, *i)
        rho_e = rt.gather_electron_charge_density()
        rho_i = rt.gather_ion_charge_density()
        a_eq = O.prepare_sys_y(prob.gmi - 1, prob.gmj)
        for phis in ([0.0, 0.7, 0.0, -0.4], [0.0, -0.25, 0.0, 1.1], [0.0, 0.0, 0.0, 0.0]):
            rt.update_wall_potentials(phis)
            prob.object_phi[:4] = phis
            phi = rt.solve_poisson_fftx()
            ref = O.solve_poisson_fftx(O.make_geom(prob.geometry()), prob.N_of_particles_cell, rho_e, rho_i, a_eq)
            err = rel_err(phi, ref)
            print("FFT_X walls above %.2f below %.2f: error vs oracle %.2e" % (phis[1], phis[3], err), flush=True)
            assert err < 1e-12
            P = phi.reshape(prob.ny, prob.nx)
            assert np.abs(P[0, 1:prob.gmi - 1] - phis[3]).max() < 1e-12 * max(1.0, np.abs(P).max())
            assert np.abs(P[-1, 1:prob.gmi - 1] - phis[1]).max() < 1e-12 * max(1.0, np.abs(P).max())
        # whole steps with an update in between: the wall rows follow the new potentials
        for phis in ([0.0, 0.5, 0.0, -0.5], [0.0, -0.3, 0.0, 0.2]):
            rt.update_wall_potentials(phis)
            rt.step()
            P = rt.get_grid("phi").reshape(prob.ny, prob.nx)
            tol = 1e-12 * max(1.0, np.abs(P).max())
            assert np.abs(P[0, 1:prob.gmi - 1] - phis[3]).max() < tol
            assert np.abs(P[-1, 1:prob.gmi - 1] - phis[1]).max() < tol
    finally:
        rt.close()


def test_krylov_wall_update(edgpu, oracle):
    """UPDATE_WALL_POTENTIALS on the Krylov path: new box and wall potentials take effect in
    the next solve, which matches the oracle CG with the new Dirichlet values."""
    O = oracle
    prob = _problem(4, 16)
    rt = edgpu.Runtime(prob, fuse_collisions=0)
    try:
        e, i = _charged(O, prob, 59)
        rt.upload(0, *e)
        rt.upload(1, *i)
        rho_e = rt.gather_electron_charge_density()
        rho_i = rt.gather_ion_charge_density()
        rt.solve_potential_krylov(rtol=1e-12)
        new = [0.1, -0.2, 0.3, 0.0, -5.0, 4.0]
        rt.update_wall_potentials(new)
        prob.object_phi[:] = new
        phi, it, rr = rt.solve_potential_krylov(rtol=1e-12, warm_start=1)
        ref = _oracle_cg(O, prob, rho_e, rho_i)
        print("Krylov after the update: %d iterations, error vs oracle %.2e" % (it, rel_err(phi, ref)), flush=True)
        assert rel_err(phi, ref) < 1e-8
        P = phi.reshape(prob.ny, prob.nx)
        for m, (i0, i1, j0, j1) in enumerate(prob.inner_boxes):
            assert np.all(P[j0:j1 + 1, i0:i1 + 1] == new[4 + m])
        assert np.all(P[1:-1, 0] == 0.1) and np.all(P[1:-1, -1] == 0.3)
    finally:
        rt.close()


# ion species: Ar+ (the first), Ar++, Ar+++ and a negative ion of 35.45 amu
EXTRA_IONS = [dict(M_i_amu=40.0, Qs=2), dict(M_i_amu=40.0, Qs=3), dict(M_i_amu=35.45, Qs=-1)]


@pytest.mark.parametrize("precision", [0, 1])
def test_multi_charge_ion_species(edgpu, oracle, precision):
    """ADVANCE_IONS over four species of charge 1, 2, 3 and -1 in the accumulated field
    (each with its own QM2s and rotation), then GATHER_ION_CHARGE_DENSITY, which weights each
    species by Qs: the pushes bit-exact (FP64; FP32 accuracy in mixed precision) and the
    density to 1e-12 against the oracle."""
    O = oracle
    base = config.workload(2, 32)
    prob = config.Problem(T_e_eV=2.0, N_plasma_m3=1.0e15, N_of_cells_debye=2, N_max_vel=5, N_of_particles_cell=50,
                          gmi=base.gmi, gmj=base.gmj, periodic_x=True, Bz_T=0.01, extra_ion_species=EXTRA_IONS,
                          name="four ion species")
    print("case four ion species, precision %d: %d per species" % (precision, prob.n_particles), flush=True)
    rt = edgpu.Runtime(prob, precision=precision, fuse_collisions=0)
    try:
        _, i = oracle_particles(O, prob)
        rng = np.random.default_rng(61)
        species = []
        for s, (Qs, QM2s, Ms) in enumerate(prob.ion_species):
            arrs = [a.copy() for a in i]
            for q in (2, 3, 4):
                arrs[q] = arrs[q] * rng.uniform(5.0, 60.0, arrs[q].size)  # a few cells per hundred steps
            rt.upload(1 + s, *arrs, tag=np.arange(1, arrs[0].size + 1, dtype=np.int32))
            species.append(arrs)
        aEX, aEY = synthetic_field(prob, amp=0.5, seed=3)
        rt.set_grid("accEX", aEX)
        rt.set_grid("accEY", aEY)
        rt.advance_ions()
        g = O.make_geom(prob.geometry())
        B = (prob.Bx, prob.By, prob.Bz)
        rho_ref = np.zeros(prob.nx * prob.ny)
        for s, (Qs, QM2s, Ms) in enumerate(prob.ion_species):
            arrs = species[s]
            sp = O.make_species(False, Qs=Qs, QM2s=QM2s, N_subcycles=prob.N_subcycles,
                                sense_B=prob.ions_sense_magnetic_field)
            rc, rem, _, _ = O.advance(g, sp, aEX, aEY, B, prob.Ez, *arrs, prob.n_boundary_obj)
            assert rc == 0
            keep = rem == 0
            got = rt.download(1 + s)
            # tags carry the slot, so the compacted order is matched exactly
            order = np.argsort(got[5])
            assert np.array_equal(got[5][order] - 1, np.nonzero(keep)[0])
            A = np.stack([a[keep] for a in arrs], axis=1)
            Bm = np.stack([a[order] for a in got[:5]], axis=1)
            if precision == 0:
                assert np.array_equal(A, Bm), "Qs %d: max diff %g" % (Qs, np.max(np.abs(A - Bm)))
            else:
                assert rel_err(A[:, :2], Bm[:, :2]) < 1e-6 and rel_err(A[:, 2:], Bm[:, 2:]) < 1e-5
            rho_ref += O.deposit(g, Qs, arrs[0], arrs[1], skip=rem)
            print("Qs %+d: %d pushed, %d removed" % (Qs, keep.size, int(rem.sum())), flush=True)
        O.fold_walls(g, rho_ref)
        rho = rt.gather_ion_charge_density()
        assert rel_err(rho, rho_ref) < (1e-12 if precision == 0 else 1e-6)
    finally:
        rt.close()


def test_krylov_vacuum_gaps(edgpu, oracle):
    """A walled box whose walls above and below are vacuum gaps between a left electrode at
    -0.3 and a right one at +0.6: the gap rows hold the linear profile exactly and the Krylov
    solve of a charged load equals the oracle's CG."""
    O = oracle
    prob = config.Problem(T_e_eV=3.0, N_plasma_m3=5.0e15, N_of_cells_debye=2, N_max_vel=5, N_of_particles_cell=100,
                          gmi=64, gmj=48, periodic_x=False, name="vacuum gaps")
    prob.vacuum_gaps = [1, 3]
    prob.object_phi[:4] = [-0.3, 5.0, 0.6, -5.0]   # the gaps' own potentials are never used
    print("case vacuum gaps %dx%d" % (prob.gmi, prob.gmj), flush=True)
    rt = edgpu.Runtime(prob, fuse_collisions=0)
    try:
        e, i = _charged(O, prob, 61)
        rt.upload(0, *e)
        rt.upload(1, *i)
        rho_e = rt.gather_electron_charge_density()
        rho_i = rt.gather_ion_charge_density()
        phi, it, rr = rt.solve_potential_krylov(rtol=1e-12)
        ref = _oracle_cg(O, prob, rho_e, rho_i)
        err = rel_err(phi, ref)
        print("Krylov with vacuum gaps: %d iterations, error vs oracle %.2e" % (it, err), flush=True)
        assert err < 1e-9
        P = phi.reshape(prob.ny, prob.nx)
        ii = np.arange(prob.nx)
        prof = -0.3 + (0.6 - (-0.3)) * ii / prob.gmi
        prof[0], prof[-1] = -0.3, 0.6
        assert np.max(np.abs(P[0] - prof)) < 1e-15 and np.max(np.abs(P[-1] - prof)) < 1e-15
        assert np.all(P[1:-1, 0] == -0.3) and np.all(P[1:-1, -1] == 0.6)
    finally:
        rt.close()


def test_fftx_vacuum_gap_in_the_wall_below(edgpu, oracle):
    """The periodic beam deck's wall below (objects 4, 5, 6 from right to left), its middle
    part a vacuum gap between the electrodes at +0.5 and -0.2: FFT_X against the oracle."""
    import os
    from edgpu import deck as D
    O = oracle
    d = D.read_deck(os.path.join(os.path.dirname(__file__), "golden", "deck_periodic_ebeam"))
    prob = d.problem()
    prob.n_particles = 300_000
    prob.vacuum_gaps = [4]
    prob.object_phi[3], prob.object_phi[5] = 0.5, -0.2
    print("case FFT_X vacuum gap, grid %dx%d" % (prob.gmi, prob.gmj), flush=True)
    rt = edgpu.Runtime(prob, fuse_collisions=0)
    try:
        e, i = _charged(O, prob, 67)
        rt.upload(0, *e)
        rt.upload(1, *i)
        rho_e = rt.gather_electron_charge_density()
        rho_i = rt.gather_ion_charge_density()
        phi = rt.solve_poisson_fftx()
        g = O.make_geom(prob.geometry())
        assert (g.seg_gap_from[2], g.seg_gap_to[2]) == (5, 3)   # the gap runs from x = 200 to 313
        ref = O.solve_poisson_fftx(g, prob.N_of_particles_cell, rho_e, rho_i, O.prepare_sys_y(prob.gmi - 1, prob.gmj))
        err = rel_err(phi, ref)
        print("FFT_X with a vacuum gap: error vs oracle %.2e" % err, flush=True)
        assert err < 1e-12
        row = phi.reshape(prob.ny, prob.nx)[0]
        # the wall row passes through the transform pair, so its values hold to rounding
        assert abs(row[200] + 0.2) < 1e-13 and abs(row[313] - 0.5) < 1e-13
        assert abs(row[256] - (-0.2 + 0.7 * 56 / 113)) < 1e-13
    finally:
        rt.close()
