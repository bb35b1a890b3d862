"""GPU parity of the FP64 hot path against the CPU oracle, stage by stage.

Each stage is fed identical inputs on both sides, so the push (pure +,-,*,/ in the
reference's order, FMA contraction off) must agree bit for bit, including cell indices
and the set of wall-crossing particles; sums whose order differs (atomic deposits, FFT)
must agree to 1e-12 relative.
"""
import numpy as np
import pytest

from edgpu import config
from helpers import oracle_particles, rel_err, sorted_rows, synthetic_field

pytestmark = pytest.mark.gpu

# (workload, scale): cfg1 at 64x64 cells and cfg2 (Bz=0.01 T) at 64x32 cells
CASES = [(1, 4), (2, 32)]


def _setup(edgpu, O, n, scale, precision=0, oblique=False):
    prob = config.workload(n, scale)
    if oblique:
        # a magnetic field with x and y components and a field along z: the push then
        # changes vz, and the kernels take the path that stores it
        prob.Bx, prob.By, prob.Ez = 0.3 * prob.Bz, -0.2 * prob.Bz, 0.004
    print("case cfg%d scale %d: %d particles/species, grid %dx%d" % (n, scale, prob.n_particles, prob.gmi, prob.gmj),
          flush=True)
    rt = edgpu.Runtime(prob, precision=precision, fuse_collisions=0)
    e, i = oracle_particles(O, prob)
    rt.upload(0, *e)
    rt.upload(1, *i)
    g = O.make_geom(prob.geometry())
    return prob, rt, e, i, g


@pytest.mark.parametrize("n,scale", CASES)
def test_gather_charge_density(edgpu, oracle, n, scale):
    O = oracle
    prob, rt, e, i, g = _setup(edgpu, O, n, scale)
    try:
        rho_e = rt.gather_electron_charge_density()
        rho_i = rt.gather_ion_charge_density()
        ref_e = O.fold_walls(g, O.deposit(g, 1, e[0], e[1]))
        ref_i = O.fold_walls(g, O.deposit(g, 1, i[0], i[1]))
        assert rel_err(rho_e, ref_e) < 1e-12
        assert rel_err(rho_i, ref_i) < 1e-12
        # every particle contributes 1 before the fold/wall doubling: check the raw total too
        raw = O.deposit(g, 1, e[0], e[1])
        assert abs(raw.sum() - e[0].size) < 1e-6
    finally:
        rt.close()


@pytest.mark.parametrize("n,scale", CASES)
def test_poisson_fftx(edgpu, oracle, n, scale):
    O = oracle
    prob, rt, e, i, g = _setup(edgpu, O, n, scale)
    try:
        rng = np.random.default_rng(3)
        rho_e = O.fold_walls(g, O.deposit(g, 1, e[0], e[1]))
        # perturb the ions so the charge density is not neutral
        rho_i = rho_e * (1.0 + 0.02 * rng.standard_normal(rho_e.size))
        rt.set_grid("rho_e", rho_e)
        rt.set_grid("rho_i", rho_i)
        phi = rt.solve_poisson_fftx()
        a_eq = O.prepare_sys_y(prob.gmi - 1, prob.gmj)
        ref = O.solve_poisson_fftx(g, prob.N_of_particles_cell, rho_e, rho_i, a_eq)
        assert rel_err(phi, ref) < 1e-12
    finally:
        rt.close()


def test_poisson_fftx_full_grid(edgpu, oracle):
    """The FFT_X solve at the benchmark grid (2048 x 1024 cells), where the chunked Thomas
    sweep runs 16 rows per chunk, against the oracle's sequential sweep."""
    O = oracle
    prob = config.workload(2)
    prob.n_particles = 1 << 16
    rt = edgpu.Runtime(prob)
    try:
        rng = np.random.default_rng(4)
        n = prob.nx * prob.ny
        base = 200.0 + rng.standard_normal(n)
        rho_i = base * (1.0 + 0.01 * rng.standard_normal(n))
        rt.set_grid("rho_e", base)
        rt.set_grid("rho_i", rho_i)
        phi = rt.solve_poisson_fftx()
        g = O.make_geom(prob.geometry())
        a_eq = O.prepare_sys_y(prob.gmi - 1, prob.gmj)
        ref = O.solve_poisson_fftx(g, prob.N_of_particles_cell, base, rho_i, a_eq)
        err = rel_err(phi, ref)
        print("full-grid FFT_X Poisson: relative error %.3g" % err, flush=True)
        assert err < 1e-12
    finally:
        rt.close()


@pytest.mark.parametrize("n,scale", CASES)
def test_electric_field(edgpu, oracle, n, scale):
    O = oracle
    prob, rt, e, i, g = _setup(edgpu, O, n, scale)
    try:
        rng = np.random.default_rng(5)
        phi = rng.standard_normal(prob.nx * prob.ny)
        # periodic images as the solver writes them
        P = phi.reshape(prob.ny, prob.nx)
        P[:, prob.gmi - 1] = P[:, 0]
        P[:, prob.gmi] = P[:, 1]
        phi = P.ravel().copy()
        rt.set_grid("phi", phi)
        rt.clear_accumulated_fields()
        EX, EY = rt.calculate_electric_field()
        acc = (np.zeros_like(phi), np.zeros_like(phi))
        rEX, rEY = O.efield(g, phi, prob.factor1_E_from_F, prob.factor2_E_from_F, acc)
        assert np.array_equal(EX, rEX)
        assert np.array_equal(EY, rEY)
        assert np.array_equal(rt.get_grid("accEX"), acc[0])
    finally:
        rt.close()


def _edge_particles(prob, rng):
    """Particles on the exact edges, on integer nodes and next to the walls, moving out."""
    X = np.array([0.0, float(prob.gmi), float(prob.gmi), 1.0, 0.0, 5.0, 0.02, prob.gmi - 0.01, 3.0, 7.0])
    Y = np.array([0.0, float(prob.gmj), 3.0, float(prob.gmj), 4.0, 0.01, 2.0, 2.0, prob.gmj - 0.02, 1.0])
    VX = np.array([-0.1, 0.2, 0.05, 0.0, -0.3, 0.01, -0.05, 0.04, 0.0, 0.0])
    VY = np.array([-0.1, 0.1, 0.0, 0.3, 0.0, -0.2, 0.0, 0.0, 0.1, -0.99])
    VZ = rng.standard_normal(X.size) * 0.1
    return X, Y, VX, VY, VZ


@pytest.mark.parametrize("n,scale,oblique", [c + (False,) for c in CASES] + [(2, 32, True)])
def test_advance_bit_exact(edgpu, oracle, n, scale, oblique):
    """ADVANCE_ELECTRONS / ADVANCE_IONS on the same field: positions, velocities, cell indices,
    removed set and per-object hit counts identical.  With Bz only and no Ez the kernels do
    not store vz (it cannot change); the oblique case has the push change it."""
    O = oracle
    prob, rt, e, i, g = _setup(edgpu, O, n, scale, oblique=oblique)
    try:
        rng = np.random.default_rng(11)
        ex = _edge_particles(prob, rng)
        e = [np.concatenate([a, b]) for a, b in zip(e, ex)]
        rt.upload(0, *e)
        EX, EY = synthetic_field(prob, amp=0.08)
        rt.set_grid("EX", EX)
        rt.set_grid("EY", EY)
        aEX, aEY = synthetic_field(prob, amp=0.08, seed=9)
        rt.set_grid("accEX", aEX)
        rt.set_grid("accEY", aEY)
        B = (prob.Bx, prob.By, prob.Bz)
        nobj = prob.n_boundary_obj + len(prob.inner_boxes)
        for species, arrs, is_e in ((0, e, True), (1, i, False)):
            arrs = [a.copy() for a in arrs]
            if is_e:
                se = rt.advance_electrons()
                sp = O.make_species(True)
                rc, removed, hit_obj, hits = O.advance(g, sp, EX, EY, B, prob.Ez, *arrs, nobj)
            else:
                se = rt.advance_ions()
                sp = O.make_species(False, Qs=prob.Qs, QM2s=prob.QM2s, N_subcycles=prob.N_subcycles,
                                    sense_B=prob.ions_sense_magnetic_field)
                rc, removed, hit_obj, hits = O.advance(g, sp, aEX, aEY, B, prob.Ez, *arrs, nobj)
            assert rc == 0
            keep = removed == 0
            got = rt.download(species)
            assert got[0].size == keep.sum()
            A = sorted_rows(*[a[keep] for a in arrs])
            Bm = sorted_rows(*got[:5])
            assert np.array_equal(A, Bm), "species %d: max diff %g" % (species, np.max(np.abs(A - Bm)))
            vz_before = (e if is_e else i)[4][keep]
            assert oblique == bool(np.any(arrs[4][keep] != vz_before))
            # cell indices (INT(X), INT(Y) with the upper-edge rule) follow from bit-equal positions
            idx, obj = rt.removed(species)
            ref_idx = np.nonzero(removed)[0]
            assert np.array_equal(np.sort(idx), ref_idx)
            order = np.argsort(idx)
            assert np.array_equal(obj[order], hit_obj[ref_idx])
            assert list(se["hit_count"][:nobj]) == list(hits[:nobj])
            print("species %d: %d pushed, %d removed, hits %s" % (species, keep.size, ref_idx.size,
                                                                 list(hits[:nobj])), flush=True)
            if is_e:
                assert ref_idx.size > 0
    finally:
        rt.close()


@pytest.mark.parametrize("n,scale", CASES + [(1, 1)])
def test_one_step_chain(edgpu, oracle, n, scale):
    """gather -> solve -> field -> push through the GPU chain against the oracle chain:
    positions, velocities, density and potential within 1e-12 relative; also at the full
    256x256-cell, 50 particles/cell size of the plain FP64 reference case (3.3 M per species)."""
    O = oracle
    prob, rt, e, i, g = _setup(edgpu, O, n, scale)
    try:
        # displace the ions a little so the first field is not zero
        rng = np.random.default_rng(2)
        i[0] = np.clip(i[0] + rng.uniform(-0.3, 0.3, i[0].size), 0.0, prob.gmi - 1.0)
        rt.upload(1, *i)
        rho_e = rt.gather_electron_charge_density()
        rho_i = rt.gather_ion_charge_density()
        phi = rt.solve_poisson_fftx()
        rt.clear_accumulated_fields()
        EX, EY = rt.calculate_electric_field()
        se = rt.advance_electrons()
        si = rt.advance_ions()
        r_e = O.fold_walls(g, O.deposit(g, 1, e[0], e[1]))
        r_i = O.fold_walls(g, O.deposit(g, 1, i[0], i[1]))
        a_eq = O.prepare_sys_y(prob.gmi - 1, prob.gmj)
        r_phi = O.solve_poisson_fftx(g, prob.N_of_particles_cell, r_e, r_i, a_eq)
        acc = (np.zeros_like(r_phi), np.zeros_like(r_phi))
        rEX, rEY = O.efield(g, r_phi, prob.factor1_E_from_F, prob.factor2_E_from_F, acc)
        assert rel_err(rho_e, r_e) < 1e-12 and rel_err(rho_i, r_i) < 1e-12
        assert rel_err(phi, r_phi) < 1e-12
        assert rel_err(EX, rEX) < 1e-12 and rel_err(EY, rEY) < 1e-12
        nobj = prob.n_boundary_obj
        B = (prob.Bx, prob.By, prob.Bz)
        sp = O.make_species(True)
        rc, rem, _, hits = O.advance(g, sp, rEX, rEY, B, prob.Ez, *e, nobj)
        got = rt.download(0)
        A = sorted_rows(*[a[rem == 0] for a in e])
        Bm = sorted_rows(*got[:5])
        assert A.shape == Bm.shape
        assert rel_err(A, Bm) < 1e-12
        # density after the step (deposited inside the push) against the oracle deposit
        rho_next = rt.gather_electron_charge_density()
        ref_next = O.fold_walls(g, O.deposit(g, 1, e[0], e[1], skip=rem))
        assert rel_err(rho_next, ref_next) < 1e-12
    finally:
        rt.close()
