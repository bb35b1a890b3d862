"""Electrons and ions that end a step inside a metal inner box, against the oracle on the same
inputs: the object charged and the crossing of the step's trajectory with the face hit
(TRY_ELECTRON_COLL_WITH_INNER_OBJECT, pic2d_ElectronWallCollisions.f90:470-557, 648-921),
including a box attached to another one (covered cells), and the electrons that a box with a
secondary-emission material sends back from that crossing (the inner-object branches of
PROCESS_ELECTRON_INDUCED_ELECTRON_EMISSION, pic2d_SEEmission.f90:372-549, 696-790); and the
initial loading, which removes the electrons placed inside the boxes before the ions take their
positions (DISTRIBUTE_PARTICLES, pic2d_CurProblemValues.f90:3245-3258, 3270-3299)."""
import math

import numpy as np
import pytest

import oracle_ffi as OF
from edgpu import config
from helpers import rel_err, sorted_rows, synthetic_field

pytestmark = pytest.mark.gpu

# the walled 64 x 64 box of the learned warm-start case with a third box attached to the
# right face of the first one: that face is covered where the two meet
BOXES = [(16, 24, 30, 34), (24, 28, 30, 34), (40, 48, 30, 34)]
SIDE_DIRFLAG = {0: 1, 1: 2, 2: 3, 3: 4}  # EDGPU_SIDE_* -> the reference's dirflag


def _problem():
    return config.Problem(T_e_eV=3.0, N_plasma_m3=5.0e15, N_of_cells_debye=2, N_max_vel=5, N_of_particles_cell=100,
                          gmi=64, gmj=64, periodic_x=False, inner_boxes=BOXES, inner_phi_V=(50.0, 50.0, -50.0),
                          name="three boxes")


def _around_boxes(prob, n, seed, vmax=1.4):
    """Particles within 2 cells of the boxes, outside them, moving up to vmax cells per step."""
    rng = np.random.default_rng(seed)
    out = []
    while sum(len(a) for a in out) < n:
        b = BOXES[rng.integers(0, len(BOXES))]
        x = rng.uniform(b[0] - 2.0, b[1] + 2.0, 4096)
        y = rng.uniform(b[2] - 2.0, b[3] + 2.0, 4096)
        inside = np.zeros(x.size, bool)
        for (i0, i1, j0, j1) in BOXES:
            inside |= (x >= i0) & (x <= i1) & (y >= j0) & (y <= j1)
        out.append(np.stack([x[~inside], y[~inside]], axis=1))
    p = np.concatenate(out)[:n]
    v = rng.uniform(-vmax, vmax, (n, 3))
    return [np.ascontiguousarray(p[:, 0]), np.ascontiguousarray(p[:, 1])] + [np.ascontiguousarray(v[:, q])
                                                                          for q in range(3)]


def _away_from_faces(O, g, sp, EX, EY, arrs, nobj, tol=1e-4):
    """The particles whose pushed position (oracle, FP64) is more than tol from every box face
    line, so that FP32 rounding cannot move one across a face."""
    ref = [a.copy() for a in arrs]
    rc, _, _, _ = O.advance(g, sp, EX, EY, (0.0, 0.0, 0.0), 0.0, *ref, nobj)
    assert rc == 0
    ok = np.ones(ref[0].size, bool)
    for (i0, i1, j0, j1) in BOXES:
        for q in (i0, i1):
            ok &= np.abs(ref[0] - q) > tol
        for q in (j0, j1):
            ok &= np.abs(ref[1] - q) > tol
    return [np.ascontiguousarray(a[ok]) for a in arrs]


@pytest.mark.parametrize("precision", [0, 1])
def test_inner_box_hits_match_oracle(edgpu, oracle, precision):
    """The same particles removed, charged to the same boxes, with the same face and (FP64) the
    same crossing bit for bit; mixed precision: the same objects and faces away from the box
    corners and crossings to FP32 accuracy."""
    O = oracle
    prob = _problem()
    print("case inner-box hits, precision %d" % precision, flush=True)
    rt = edgpu.Runtime(prob, precision=precision, fuse_collisions=0)
    try:
        g = O.make_geom(prob.geometry())
        nobj = prob.n_boundary_obj + len(BOXES)
        EX, EY = synthetic_field(prob, amp=0.02)
        rt.set_grid("EX", EX)
        rt.set_grid("EY", EY)
        rt.set_grid("accEX", EX)
        rt.set_grid("accEY", EY)
        for species, seed in ((0, 1), (1, 2)):
            e = _around_boxes(prob, 60000, seed)
            if species == 1:
                e[2:] = [v * 0.5 for v in e[2:]]  # ions move N_subcycles v with N_subcycles = 1
            sp = O.make_species(True) if species == 0 else O.make_species(
                False, Qs=prob.Qs, QM2s=prob.QM2s, N_subcycles=prob.N_subcycles,
                sense_B=prob.ions_sense_magnetic_field)
            e = _away_from_faces(O, g, sp, EX, EY, e, nobj)
            rt.upload(species, *e)
            st = rt.advance_electrons() if species == 0 else rt.advance_ions()
            ref = [a.copy() for a in e]
            rc, rem, hit_obj, hits, rx, ry, rd = O.advance_records(g, sp, EX, EY, (0.0, 0.0, 0.0), 0.0, *ref, nobj)
            assert rc == 0
            idx, obj = rt.removed(species)
            x, y, side = rt.removed_records(species)
            order = np.argsort(idx)
            idx, obj, x, y, side = idx[order], obj[order], x[order], y[order], side[order]
            inner = obj >= prob.n_boundary_obj
            print("species %d: %d removed, %d by the boxes %s" % (species, idx.size, inner.sum(),
                                                                   list(hits[4:nobj])), flush=True)
            assert inner.sum() > 1000
            assert np.array_equal(idx, np.nonzero(rem)[0])
            assert list(st["hit_count"][:nobj]) == list(hits[:nobj])
            assert np.array_equal(obj, hit_obj[idx])
            dirflag = np.array([SIDE_DIRFLAG[s] for s in side])
            if precision == 0:
                assert np.array_equal(dirflag, rd[idx])
                assert np.array_equal(x[inner], rx[idx][inner]) and np.array_equal(y[inner], ry[idx][inner])
            else:
                # a crossing within 1e-5 of a box corner may take the other face in FP32
                cx, cy = rx[idx][inner], ry[idx][inner]
                near = np.zeros(cx.size, bool)
                for (i0, i1, j0, j1) in BOXES:
                    for (qx, qy) in ((i0, j0), (i0, j1), (i1, j0), (i1, j1)):
                        near |= (np.abs(cx - qx) < 1e-4) & (np.abs(cy - qy) < 1e-4)
                assert np.array_equal(dirflag[inner][~near], rd[idx][inner][~near])
                assert np.max(np.abs(x[inner] - cx)) < 1e-4 and np.max(np.abs(y[inner] - cy)) < 1e-4
    finally:
        rt.close()


def test_inner_box_emission_matches_oracle(edgpu, oracle):
    """A box with an elastic, inelastic and true-secondary material: the electrons it emits
    for the electrons that hit it (fed the same uniforms per absorbed slot) equal the oracle's,
    placed 1e-6 outside the face hit, to 1e-12."""
    O = oracle
    prob = _problem()
    mat = dict(model_elastic=1, setD_elastic=0.15, minE_elastic_eV=0.0, maxE_elastic_eV=500.0, elastic_reflection=1,
               model_inelastic=1, setD_inelastic=0.1, minE_inelastic_eV=0.0, maxE_inelastic_eV=500.0,
               model_true=1, setD_true=0.9, minE_true_eV=1.0, maxE_true_eV=400.0, T_true_eV=2.0)
    rt = edgpu.Runtime(prob, fuse_collisions=0)
    try:
        box_obj = prob.n_boundary_obj  # the first box
        rt.setup_electron_induced_emission(box_obj, **mat)
        g = O.make_geom(prob.geometry())
        nobj = prob.n_boundary_obj + len(BOXES)
        zero = np.zeros(prob.nx * prob.ny)
        rt.set_grid("EX", zero)
        rt.set_grid("EY", zero)
        e = _around_boxes(prob, 60000, 7)
        n = e[0].size
        u = np.random.default_rng(9).uniform(0.0, 1.0, n * 32)
        rt.upload(0, *e)
        rt.set_uniform_stream(u)
        st = rt.advance_electrons()
        rt.set_uniform_stream(None)
        n_after = rt.count(0)
        rt.process_added_electrons()
        got = rt.download(0)
        emitted = np.stack(got[:5], axis=1)[n_after:]
        ref = [a.copy() for a in e]
        rc, rem, hit_obj, hits, rx, ry, rd = O.advance_records(g, O.make_species(True), zero, zero, (0.0, 0.0, 0.0),
                                                                0.0, *ref, nobj)
        assert rc == 0
        maxw, inj = O.maxwell_tables()
        see = OF.see_from_material(mat, prob.energy_factor_eV, prob.T_e_eV, prob.N_max_vel)
        rows = []
        for k in np.nonzero(rem & (hit_obj == box_obj))[0]:
            rc, out = OF.see_emit_inner(see, maxw, inj, int(rd[k]), BOXES[0], rx[k], ry[k], ref[2][k], ref[3][k],
                                        ref[4][k], u[32 * k:32 * k + 32])
            assert rc == 0
            rows.append(out)
        rows = np.concatenate(rows)
        print("box hits %d, emitted %d (gpu %d)" % (hits[box_obj], rows.shape[0], emitted.shape[0]), flush=True)
        assert st["emit_count"][box_obj] == rows.shape[0] > 500
        assert emitted.shape == rows.shape
        assert rel_err(sorted_rows(*rows.T), sorted_rows(*emitted.T)) < 1e-12
        assert np.all(got[5][n_after:] == box_obj + 1)
    finally:
        rt.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_loading_removes_particles_inside_boxes(edgpu, precision):
    """DISTRIBUTE_PARTICLES on the device with three boxes (one attached to another): no
    electron or ion is left strictly inside a box, the count removed is the binomial share of
    the boxes' area (within 6 sigma), the ions sit on the first electrons' positions, and a
    push right after the loading finds a crossing for every particle that enters a box."""
    prob = _problem()
    prob.n_particles = 1 << 20
    print("case loading with inner boxes, precision %d" % precision, flush=True)
    rt = edgpu.Runtime(prob, precision=precision, fuse_collisions=0)
    try:
        rt.load_initial_particles()
        ne, ni = rt.count(0), rt.count(1)
        area = sum((b[1] - b[0]) * (b[3] - b[2]) for b in BOXES) / float(prob.gmi * prob.gmj)
        mean = prob.n_particles * (1.0 - area)
        sigma = math.sqrt(prob.n_particles * area * (1.0 - area))
        print("loaded %d electrons, %d ions, expected %.0f +- %.0f" % (ne, ni, mean, sigma), flush=True)
        assert ne == ni and abs(ne - mean) < 6 * sigma
        e = rt.download(0)
        i = rt.download(1)
        for (i0, i1, j0, j1) in BOXES:
            for (X, Y) in ((e[0], e[1]), (i[0], i[1])):
                assert not np.any((X > i0) & (X < i1) & (Y > j0) & (Y < j1))
        assert np.array_equal(i[0], e[0]) and np.array_equal(i[1], e[1])
        rt.set_grid("EX", np.zeros(prob.nx * prob.ny))
        rt.set_grid("EY", np.zeros(prob.nx * prob.ny))
        st = rt.advance_electrons()
        assert sum(st["hit_count"][prob.n_boundary_obj:prob.n_boundary_obj + len(BOXES)]) > 0
    finally:
        rt.close()
