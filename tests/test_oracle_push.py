"""Pinning of the oracle's RNG, Maxwell sampler, push, routing and deposit (CPU only).

Each check is independent of the oracle's own formulas: the reference's RNG compiled
from its sources (golden vectors), closed forms of the Boris push, exactness of bilinear
interpolation on linear fields, hand-worked routing cases and conservation of charge.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------------------------------
# WELL19937a (src/WELL19937a_new.c, pic2d_RandomNumberInterface.f90:39-57)

@pytest.mark.parametrize("seed", [123456789, 7, 2026])
def test_well_matches_reference_golden(oracle, seed):
    ref = np.load(os.path.join(GOLDEN, "well19937a_seed%d.npy" % seed))
    got = oracle.well_fill(oracle.well(seed), ref.size)
    assert np.array_equal(got, ref)


def test_well_matches_reference_build(oracle):
    """When the reference's RNG was compiled here (oracle/_ref), compare far into the stream."""
    r = oracle.ref_lib()
    if r is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    r.ref_well_seed(123456789)
    ref = np.empty(200000)
    r.ref_well_fill(ref.size, ref.ctypes.data_as(C.POINTER(C.c_double)))
    got = oracle.well_fill(oracle.well(123456789), ref.size)
    assert np.array_equal(got, ref)


# face m of an inner box (1 left, 2 top, 3 right, 4 bottom) from the dirflag the oracle records
_SEGMENT_OF_DIRFLAG = {3: 1, 4: 2, 1: 3, 2: 4}


def _push_cases():
    from edgpu import config
    oblique = config.Problem(T_e_eV=2.0, N_plasma_m3=1.0e15, N_of_cells_debye=2, N_max_vel=5, N_of_particles_cell=40,
                             gmi=40, gmj=30, periodic_x=False, Bx_T=0.004, By_T=-0.007, Bz_T=0.01, Ez_Vm=3.0e3,
                             name="walled box, oblique B and Ez")
    return [("cfg1/8", config.workload(1, 8)), ("cfg2/32 (Bz)", config.workload(2, 32)),
            ("cfg4/16 (inner boxes)", config.workload(4, 16)), (oblique.name, oblique)]


@pytest.mark.parametrize("case", range(4))
def test_push_matches_reference_build(oracle, case):
    """The oracle's electron push against the reference's own ADVANCE_ELECTRONS compiled from
    its sources (oracle/_ref/libpushref.so): the same electrons left, bit for bit (matched by
    tag), the same hit count per object, and per object the same multiset of wall-hit records
    (tag + 100 * segment, REAL coordinate): for inner boxes the face crossed and the crossing
    of the trajectory, for walls the clamped position along the wall."""
    from helpers import oracle_particles, synthetic_field
    O = oracle
    if O.ref_push_lib() is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    name, prob = _push_cases()[case]
    e, _ = oracle_particles(O, prob)
    e[2:] = [v * 8.0 for v in e[2:]]  # a few cells per step: many wall and box hits
    EX, EY = synthetic_field(prob, amp=0.05)
    B = (prob.Bx, prob.By, prob.Bz)
    cfg = prob.geometry()
    nobj = prob.n_boundary_obj + len(prob.inner_boxes)
    left, tag, hits, (r_obj, r_tok, r_coord) = O.ref_advance_electrons(cfg, B, prob.Ez, *e, EX, EY)
    ref = [a.copy() for a in e]
    rc, rem, hit_obj, ohits, rx, ry, rd = O.advance_records(O.make_geom(cfg), O.make_species(True), EX, EY, B,
                                                            prob.Ez, *ref, nobj)
    assert rc == 0
    keep = np.nonzero(rem == 0)[0]
    order = np.argsort(tag)
    print("%s: %d electrons, %d left, hits %s" % (name, e[0].size, keep.size, list(hits)))
    assert np.array_equal(tag[order] - 1, keep)
    for q in range(5):
        assert np.array_equal(left[q][order], ref[q][keep])
    assert list(hits) == list(ohits[:nobj]) and hits.sum() == rem.sum() > 1000
    _check_records(prob, nobj, rem, hit_obj, rx, ry, rd, r_obj, r_tok, r_coord, 0)


def _check_records(prob, nobj, rem, hit_obj, rx, ry, rd, r_obj, r_tok, r_coord, token_base):
    """Per object, the multiset of the reference's wall-hit records (token, REAL coordinate)
    equals the one the oracle's removals imply: for inner boxes the face crossed and the
    coordinate of the crossing along it, for walls segment 1 and the clamped position."""
    for o in range(nobj):
        ks = np.nonzero((rem != 0) & (hit_obj == o))[0]
        if o >= prob.n_boundary_obj:
            seg = np.array([_SEGMENT_OF_DIRFLAG[d] for d in rd[ks]], np.int64)
            coord = np.where(np.isin(seg, (1, 3)), ry[ks], rx[ks])
        else:
            seg = np.ones(ks.size, np.int64)
            along_x = np.isin(rd[ks], (2, 4))
            coord = np.where(along_x, np.clip(rx[ks], 0.0, prob.gmi), np.clip(ry[ks], 0.0, prob.gmj))
        want = sorted(zip((ks + 1 + 100 * seg + token_base).tolist(), coord.astype(np.float32).tolist()))
        sel = r_obj == o + 1
        got = sorted(zip(r_tok[sel].astype(np.int64).tolist(), r_coord[sel].tolist()))
        assert got == want, "object %d" % (o + 1)


@pytest.mark.parametrize("case", range(4))
def test_ion_push_matches_reference_build(oracle, case):
    """The oracle's ion push against the reference's own ADVANCE_IONS compiled from its
    sources: charges 1, 2 and -1 beside other species, N_subcycles 1, 3 and 5, with and
    without the magnetic response and Ez, walls and inner boxes (ions go back N_subcycles
    steps for the crossing).  The same ions left bit for bit, hit counts and hit records."""
    from helpers import oracle_particles, synthetic_field
    O = oracle
    if O.ref_push_lib() is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    name, prob = _push_cases()[case]
    # (species [(Qs, mass in amu)], pushed species, N_subcycles, sense_B, sense_Ez)
    setup = [([(1, 40.0)], 1, 1, True, False),
             ([(1, 40.0), (2, 40.0), (-1, 16.0)], 2, 3, True, True),
             ([(1, 40.0)], 1, 1, False, False),
             ([(1, 40.0), (-1, 16.0)], 2, 5, True, True)][case]
    species = [(q, 0.5 * q / (O.AMU_KG * m / O.M_E_KG)) for q, m in setup[0]]
    s, nsub, sense_b, sense_ez = setup[1:]
    _, i = oracle_particles(O, prob)
    rng = np.random.default_rng(100 + case)
    vmax = 1.6 / nsub  # up to 1.6 cells per ion step
    i[2], i[3], i[4] = [rng.uniform(-vmax, vmax, i[0].size) for _ in range(3)]
    aEX, aEY = synthetic_field(prob, amp=2.0, seed=11)
    B = (prob.Bx, prob.By, prob.Bz)
    cfg = prob.geometry()
    nobj = prob.n_boundary_obj + len(prob.inner_boxes)
    left, tag, hits, (r_obj, r_tok, r_coord) = O.ref_advance_ions(cfg, B, prob.Ez, species, s, nsub, sense_b,
                                                                  sense_ez, *i, aEX, aEY)
    Qs, QM2s = species[s - 1]
    sp = O.make_species(False, Qs=Qs, QM2s=QM2s, N_subcycles=nsub, sense_B=1 if sense_b else 0,
                        sense_Ez=1 if sense_ez else 0)
    ref = [a.copy() for a in i]
    rc, rem, hit_obj, ohits, rx, ry, rd = O.advance_records(O.make_geom(cfg), sp, aEX, aEY, B, prob.Ez, *ref, nobj)
    assert rc == 0
    keep = np.nonzero(rem == 0)[0]
    order = np.argsort(tag)
    print("%s, ions Qs %+d, N_subcycles %d: %d ions, %d left, hits %s" % (name, Qs, nsub, i[0].size, keep.size,
                                                                        list(hits)))
    assert np.array_equal(tag[order] - 1, keep)
    for q in range(5):
        assert np.array_equal(left[q][order], ref[q][keep])
    assert list(hits) == list(ohits[:nobj]) and hits.sum() == rem.sum() > 1000
    _check_records(prob, nobj, rem, hit_obj, rx, ry, rd, r_obj, r_tok, r_coord, 10000 * s)


def test_well_uniform_statistics(oracle):
    u = oracle.well_fill(oracle.well(99), 200000)
    assert u.min() >= 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 5 * math.sqrt(1 / 12 / u.size)
    h, _ = np.histogram(u, bins=20, range=(0, 1))
    chi2 = ((h - u.size / 20) ** 2 / (u.size / 20)).sum()
    assert chi2 < 60  # 19 degrees of freedom


# ---------------------------------------------------------------------------------------
# Maxwell tables (PrepareMaxwellDistribIntegral, pic2d_ElectronWallCollisions.f90:835-929)

def _erf(x):
    return np.vectorize(math.erf)(x)


def test_maxwell_table_is_inverse_cdf(oracle):
    v, vi = oracle.maxwell_tables()
    # exp(-v^2) on [-3, 3]: F(v) = (erf v + erf 3) / (2 erf 3)
    F = (_erf(v) + math.erf(3.0)) / (2 * math.erf(3.0))
    k = np.arange(v.size) / 9000.0
    assert np.max(np.abs(F - k)) < 1e-4
    assert v[0] == -3.0 and np.all(np.diff(v) > 0)
    # v exp(-v^2) on [0, 3]: F(v) = (1 - exp(-v^2)) / (1 - exp(-9))
    Fi = (1 - np.exp(-vi ** 2)) / (1 - math.exp(-9.0))
    ki = np.arange(vi.size) / 4500.0
    assert np.max(np.abs(Fi - ki)) < 1e-4


def test_maxwell_sampler_moments(oracle):
    v, vi = oracle.maxwell_tables()
    L = oracle.lib()
    u = (np.arange(200000) + 0.5) / 200000
    s = np.array([L.orc_maxwell_sample(oracle.dp(v), x) for x in u[::10]])
    # truncated exp(-v^2): mean 0, <v^2> = 1/2 - 3 exp(-9) / (sqrt(pi) erf 3)
    var = 0.5 - 3 * math.exp(-9) / (math.sqrt(math.pi) * math.erf(3))
    assert abs(s.mean()) < 1e-3
    assert abs((s ** 2).mean() - var) < 2e-3


# ---------------------------------------------------------------------------------------
# push (ADVANCE_ELECTRONS / ADVANCE_IONS)

def _box(oracle, gmi=17, gmj=12, periodic=True):
    if periodic:
        segs = [(oracle.SIDE_ABOVE, 0, gmi, 1), (oracle.SIDE_BELOW, 0, gmi, 3)]
    else:
        segs = [(oracle.SIDE_LEFT, 0, gmj, 0), (oracle.SIDE_ABOVE, 0, gmi, 1), (oracle.SIDE_RIGHT, 0, gmj, 2),
                (oracle.SIDE_BELOW, 0, gmi, 3)]
    return oracle.make_geom(dict(gmi=gmi, gmj=gmj, periodic_x=periodic, segments=segs, object_phi=[0.0] * 4,
                                 n_boundary_obj=4, inner_boxes=[]))


def _interior(rng, n, gmi, gmj, margin=2.0):
    return rng.uniform(margin, gmi - margin, n), rng.uniform(margin, gmj - margin, n)


def _linear_field(gmi, gmj, a, b, c):
    i = np.arange(gmi + 1)[None, :]
    j = np.arange(gmj + 1)[:, None]
    return np.ascontiguousarray((a + b * i + c * j + 0 * i * j).ravel())


@pytest.mark.parametrize("periodic", [True, False])
def test_push_e_only_closed_form(oracle, periodic):
    """B = 0: v' = v - E(x), x' = x + v' with E bilinear-exact for a linear field."""
    gmi, gmj = 17, 12
    g = _box(oracle, gmi, gmj, periodic)
    rng = np.random.default_rng(1)
    n = 2000
    X, Y = _interior(rng, n, gmi, gmj)
    V = [rng.normal(0, 0.05, n) for _ in range(3)]
    EX = _linear_field(gmi, gmj, 0.01, 0.002, -0.001)
    EY = _linear_field(gmi, gmj, -0.02, 0.0005, 0.003)
    x0, y0, v0 = X.copy(), Y.copy(), [a.copy() for a in V]
    rc, rem, _, _ = oracle.advance(g, oracle.make_species(True), EX, EY, (0.0, 0.0, 0.0), 0.0, X, Y, *V, 4)
    assert rc == 0 and rem.sum() == 0
    ex = 0.01 + 0.002 * x0 - 0.001 * y0
    ey = -0.02 + 0.0005 * x0 + 0.003 * y0
    assert np.allclose(V[0], v0[0] - ex, rtol=0, atol=1e-15)
    assert np.allclose(V[1], v0[1] - ey, rtol=0, atol=1e-15)
    assert np.array_equal(V[2], v0[2])
    assert np.allclose(X, x0 + V[0], rtol=0, atol=1e-13)
    assert np.allclose(Y, y0 + V[1], rtol=0, atol=1e-13)


def test_push_b_only_rotation(oracle):
    """E = 0, B = Bz: the Boris step rotates v_perp counter-clockwise (electron) by
    2 atan(Bz/2), keeps |v| and vz."""
    gmi, gmj = 17, 12
    g = _box(oracle, gmi, gmj)
    rng = np.random.default_rng(2)
    n = 1000
    X, Y = _interior(rng, n, gmi, gmj, 3.0)
    V = [rng.normal(0, 0.1, n) for _ in range(3)]
    Z = np.zeros((gmj + 1) * (gmi + 1))
    bz = 0.0986
    v0 = [a.copy() for a in V]
    rc, rem, _, _ = oracle.advance(g, oracle.make_species(True), Z, Z, (0.0, 0.0, bz), 0.0, X, Y, *V, 4)
    assert rc == 0 and rem.sum() == 0
    ang = np.arctan2(v0[0] * V[1] - v0[1] * V[0], v0[0] * V[0] + v0[1] * V[1])
    assert np.allclose(ang, 2 * math.atan(bz / 2), atol=1e-13)
    assert np.allclose(np.hypot(V[0], V[1]), np.hypot(v0[0], v0[1]), rtol=1e-14)
    assert np.array_equal(V[2], v0[2])
    # ions (positive charge, alpha = +QM2s*Nsub*B) turn the other way
    Xi, Yi = _interior(rng, n, gmi, gmj, 3.0)
    Vi = [rng.normal(0, 0.01, n) for _ in range(3)]
    vi0 = [a.copy() for a in Vi]
    xi0 = Xi.copy()
    qm2 = 0.5 / (40 * 1822.888)
    sp = oracle.make_species(False, QM2s=qm2, N_subcycles=3)
    rc, rem, _, _ = oracle.advance(g, sp, Z, Z, (0.0, 0.0, bz), 0.0, Xi, Yi, *Vi, 4)
    assert rc == 0
    ang = np.arctan2(vi0[0] * Vi[1] - vi0[1] * Vi[0], vi0[0] * Vi[0] + vi0[1] * Vi[1])
    assert np.allclose(ang, -2 * math.atan(qm2 * 3 * bz), atol=1e-13)
    assert np.allclose(Xi, xi0 + 3 * Vi[0], rtol=0, atol=1e-13)


def test_push_ions_without_magnetic_response(oracle):
    """ions_sense_magnetic_field = 0: v' = v + 2 QM2s E, x' = x + N_sub v'."""
    gmi, gmj = 17, 12
    g = _box(oracle, gmi, gmj)
    rng = np.random.default_rng(3)
    n = 500
    X, Y = _interior(rng, n, gmi, gmj, 3.0)
    V = [rng.normal(0, 0.01, n) for _ in range(3)]
    EX = _linear_field(gmi, gmj, 0.3, 0.0, 0.0)
    EY = _linear_field(gmi, gmj, -0.2, 0.0, 0.0)
    x0, v0 = X.copy(), [a.copy() for a in V]
    q = 1e-4
    sp = oracle.make_species(False, QM2s=q, N_subcycles=3, sense_B=0, sense_Ez=0)
    rc, _, _, _ = oracle.advance(g, sp, EX, EY, (0.0, 0.0, 0.5), 0.7, X, Y, *V, 4)
    assert rc == 0
    assert np.allclose(V[0], v0[0] + 2 * q * 0.3, atol=1e-16)
    assert np.allclose(V[1], v0[1] - 2 * q * 0.2, atol=1e-16)
    assert np.array_equal(V[2], v0[2])  # Ez not sensed
    assert np.allclose(X, x0 + 3 * V[0], atol=1e-13)


def test_push_edge_cells_at_upper_boundary(oracle):
    """x == xmax uses cell gmi-1 with weight 1 on the last column (pic2d_ElectronDynamics.f90:62-66)."""
    gmi, gmj = 9, 6
    g = _box(oracle, gmi, gmj, periodic=False)
    EX = _linear_field(gmi, gmj, 0.0, 0.001, 0.0)
    EY = _linear_field(gmi, gmj, 0.0, 0.0, 0.002)
    X = np.array([float(gmi), 3.25])
    Y = np.array([2.5, float(gmj)])
    V = [np.array([-0.5, 0.0]), np.array([0.0, -0.5]), np.zeros(2)]
    rc, rem, _, _ = oracle.advance(g, oracle.make_species(True), EX, EY, (0.0, 0.0, 0.0), 0.0, X, Y, *V, 4)
    assert rc == 0 and rem.sum() == 0
    # v - E/2 - E/2 rounds within an ulp of v - E
    assert V[0][0] == pytest.approx(-0.5 - 0.001 * gmi, abs=1e-15)
    assert V[1][1] == pytest.approx(-0.5 - 0.002 * gmj, abs=1e-15)


def test_routing_hand_cases(oracle):
    """Periodic wrap by +-L, wall hits charged to the segment's object, corner rule of
    SURROUNDED_BY_WALL corners in a walled box."""
    gmi, gmj = 17, 12
    L = gmi - 1
    Z = np.zeros((gmj + 1) * (gmi + 1))
    g = _box(oracle, gmi, gmj, periodic=True)
    X = np.array([0.2, 16.9, 5.0, 5.0])
    Y = np.array([5.0, 5.0, 0.3, 11.8])
    V = [np.array([-0.5, 0.3, 0.0, 0.0]), np.array([0.0, 0.0, -0.5, 0.4]), np.zeros(4)]
    rc, rem, obj, hits = oracle.advance(g, oracle.make_species(True), Z, Z, (0, 0, 0), 0.0, X, Y, *V, 4)
    assert rc == 0
    assert list(rem) == [0, 0, 1, 1]
    assert X[0] == pytest.approx(0.2 - 0.5 + L) and X[1] == pytest.approx(16.9 + 0.3 - L)
    assert list(obj[2:]) == [3, 1] and list(hits) == [0, 1, 0, 1]
    # walled box: x<0 near the bottom-left corner goes to the closer wall
    g = _box(oracle, gmi, gmj, periodic=False)
    X = np.array([0.3, 0.1, 5.0, 16.5])
    Y = np.array([0.4, 0.8, 6.0, 11.5])
    V = [np.array([-0.5, -0.3, -5.5, 0.6]), np.array([-0.6, -0.9, 0.0, 0.9]), np.zeros(4)]
    rc, rem, obj, hits = oracle.advance(g, oracle.make_species(True), Z, Z, (0, 0, 0), 0.0, X, Y, *V, 4)
    assert rc == 0 and list(rem) == [1, 1, 1, 1]
    # case 0: x=-0.2, y=-0.2 -> not (0.2 < 0.2) -> LEFT (object 0)
    # case 1: x=-0.2, y=-0.1 -> not (0.2 < 0.1) -> LEFT (object 0)
    # case 2: x=-0.5, y=6 -> LEFT
    # case 3: x=17.1, y=12.4 -> (x-xmax)=0.1 < (y-ymax)=0.4 -> ABOVE (object 1)
    assert list(obj) == [0, 0, 0, 1]
    X = np.array([0.3])
    Y = np.array([0.4])
    V = [np.array([-0.35]), np.array([-0.9]), np.zeros(1)]
    rc, rem, obj, hits = oracle.advance(g, oracle.make_species(True), Z, Z, (0, 0, 0), 0.0, X, Y, *V, 4)
    # x=-0.05, y=-0.5 -> 0.05 < 0.5 -> BELOW (object 3)
    assert list(obj) == [3]


# ---------------------------------------------------------------------------------------
# deposit and fold (GATHER_*_CHARGE_DENSITY)

@pytest.mark.parametrize("Qs", [1, 2, 3, -1])
def test_single_particle_weights(oracle, Qs):
    gmi, gmj = 9, 6
    g = _box(oracle, gmi, gmj, periodic=False)
    x, y = 3.25, 2.625
    rho = oracle.deposit(g, Qs, np.array([x]), np.array([y])).reshape(gmj + 1, gmi + 1)
    fx, fy = 0.25, 0.625
    want = {(3, 2): (1 - fx) * (1 - fy), (4, 2): fx * (1 - fy), (3, 3): (1 - fx) * fy, (4, 3): fx * fy}
    for (i, j), w in want.items():
        assert rho[j, i] == pytest.approx(Qs * w, abs=1e-15)
    assert np.count_nonzero(rho) == 4
    assert rho.sum() == pytest.approx(Qs, abs=1e-14)


def test_deposit_total_and_periodic_fold(oracle):
    """Raw total = N; after the periodic fold the N unique columns hold the whole
    charge (node 0 = node N, node 1 = node N+1) and the wall rows are doubled."""
    gmi, gmj = 17, 12
    N = gmi - 1
    g = _box(oracle, gmi, gmj, periodic=True)
    rng = np.random.default_rng(4)
    n = 5000
    X = rng.uniform(0, N, n)
    Y = rng.uniform(0, gmj, n)
    raw = oracle.deposit(g, 1, X, Y)
    assert raw.sum() == pytest.approx(n, rel=1e-13)
    f = oracle.fold_walls(g, raw.copy()).reshape(gmj + 1, gmi + 1)
    r = raw.reshape(gmj + 1, gmi + 1)
    assert np.array_equal(f[:, 0], f[:, N]) and np.array_equal(f[:, gmi], f[:, 1])
    interior = f[1:gmj, 1:N + 1].sum() + 0.5 * (f[0, 1:N + 1].sum() + f[gmj, 1:N + 1].sum())
    assert interior == pytest.approx(n, rel=1e-13)
    # periodic images: column 1 got column N+1, column N got column 0
    assert np.allclose(f[1:gmj, 1], r[1:gmj, 1] + r[1:gmj, gmi])
    assert np.allclose(f[1:gmj, N], r[1:gmj, N] + r[1:gmj, 0])


def test_deposit_walled_box_fold(oracle):
    """Walled box: edges doubled, corners x4: a uniform particle density deposits a
    uniform folded density (the half/quarter cells at the walls are compensated)."""
    gmi, gmj = 9, 7
    g = _box(oracle, gmi, gmj, periodic=False)
    # one particle at every sub-cell centre of a fine lattice: uniform density
    s = 8
    xs = (np.arange(gmi * s) + 0.5) / s
    ys = (np.arange(gmj * s) + 0.5) / s
    X, Y = [a.ravel() for a in np.meshgrid(xs, ys)]
    f = oracle.fold_walls(g, oracle.deposit(g, 1, X, Y)).reshape(gmj + 1, gmi + 1)
    assert np.allclose(f, s * s, rtol=1e-12)


def test_distribute_draw_order_and_clamps(oracle):
    """DISTRIBUTE_PARTICLES (pic2d_CurProblemValues.f90:3197-3240, 3280-3299): each electron
    takes five consecutive draws of the one stream, in the order X, Y, VX, VY, VZ; positions
    are clamped to the area; ions copy the electron positions, take two draws (VX, VY) each
    and VZ = 0.  Checked against a second generator of the same seed drawn by hand, with
    limits beyond the area so that the clamps act."""
    import ctypes as C
    O = oracle
    v, _ = O.maxwell_tables()
    n = 500
    fc = 0.37
    w = O.well(20261203)
    e = [np.empty(n) for _ in range(5)]
    O.lib().orc_distribute(C.byref(w), O.dp(v), n, -2.0, 10.0, 1.0, 3.0, 0.0, 8.0, 0.0, 4.0, fc,
                           *[O.dp(a) for a in e])
    ions = [np.empty(n) for _ in range(5)]
    O.lib().orc_distribute_ions(C.byref(w), O.dp(v), n, O.dp(e[0]), O.dp(e[1]), 0.5, *[O.dp(a) for a in ions])
    u = O.well_fill(O.well(20261203), 7 * n + 1)
    ue, ui = u[:5 * n].reshape(n, 5), u[5 * n:7 * n].reshape(n, 2)
    ms = np.vectorize(lambda r: O.lib().orc_maxwell_sample(O.dp(v), r))
    assert np.array_equal(e[0], np.clip(-2.0 + 12.0 * ue[:, 0], 0.0, 8.0))
    assert np.array_equal(e[1], np.clip(1.0 + 2.0 * ue[:, 1], 0.0, 4.0))
    for q in (2, 3, 4):
        assert np.array_equal(e[q], ms(ue[:, q]) * fc)
    assert (e[0] == 0.0).any() and (e[0] == 8.0).any()  # both clamps acted
    assert np.array_equal(ions[0], e[0]) and np.array_equal(ions[1], e[1])
    assert np.array_equal(ions[2], ms(ui[:, 0]) * 0.5) and np.array_equal(ions[3], ms(ui[:, 1]) * 0.5)
    assert not ions[4].any()
    # the generator moved on by exactly 7 draws per particle pair
    assert O.lib().orc_well_next(C.byref(w)) == u[7 * n]


def test_remove_inside_inner_hand_case(oracle):
    """The loading's removal of electrons placed inside inner objects
    (pic2d_CurProblemValues.f90:3245-3258 with REMOVE_ELECTRON, pic2d_ElectronDynamics.f90:517-553)
    on the box [2,4]x[2,4], worked by hand: slot 1 is inside and takes the last electron (inside
    too), which takes the next last (inside too), which takes (1,1); faces and corners stay."""
    g = _walled(oracle, [(2, 4, 2, 4)], gmi=8, gmj=8)
    X = [3.0, 2.0, 4.0, 1.0, 3.5, 2.1]
    Y = [3.0, 3.0, 4.0, 1.0, 2.5, 3.9]
    VX = [0.0, 1.0, 2.0, 3.0, 4.0, 5.0]
    out = oracle.remove_inside_inner(g, X, Y, VX, VX, VX)
    assert out[0].tolist() == [1.0, 2.0, 4.0] and out[1].tolist() == [1.0, 3.0, 4.0]
    assert out[2].tolist() == [3.0, 1.0, 2.0] and out[4].tolist() == [3.0, 1.0, 2.0]
    # everything inside: nothing left; nothing inside: nothing moves
    assert oracle.remove_inside_inner(g, [3.0] * 4, [3.0] * 4, *([np.zeros(4)] * 3))[0].size == 0
    keep = oracle.remove_inside_inner(g, [1.0, 5.0], [1.0, 7.0], [1.0, 2.0], [0.0, 0.0], [0.0, 0.0])
    assert keep[0].tolist() == [1.0, 5.0] and keep[2].tolist() == [1.0, 2.0]


def test_remove_inside_inner_random_loads(oracle):
    """On random loads over two boxes (one attached to the wall below) the electrons kept are
    exactly the rows outside every open box, each row intact."""
    boxes = [(10, 14, 10, 13), (20, 26, 0, 5)]
    g = _walled(oracle, boxes)
    rng = np.random.default_rng(3)
    n = 20000
    X = rng.uniform(0.0, 40.0, n)
    Y = rng.uniform(0.0, 30.0, n)
    X[:500] = rng.integers(9, 15, 500)  # on and near the faces of the first box
    Y[:500] = rng.integers(9, 14, 500)
    V = [np.arange(n, dtype=np.float64) + q for q in (0.0, 0.25, 0.5)]
    inside = np.zeros(n, bool)
    for (i0, i1, j0, j1) in boxes:
        inside |= (X > i0) & (X < i1) & (Y > j0) & (Y < j1)
    out = oracle.remove_inside_inner(g, X, Y, *V)
    assert out[0].size == n - inside.sum() and inside.sum() > 500
    got = np.sort(out[2])
    assert np.array_equal(got, np.nonzero(~inside)[0].astype(np.float64))
    k = out[2].astype(np.int64)
    assert np.array_equal(out[0], X[k]) and np.array_equal(out[1], Y[k])
    assert np.array_equal(out[3], V[1][k]) and np.array_equal(out[4], V[2][k])


# ---------------------------------------------------------------------------------------
# TRY_ELECTRON_COLL_WITH_INNER_OBJECT (pic2d_ElectronWallCollisions.f90:470-557, 648-921)

def _walled(oracle, boxes, gmi=40, gmj=30):
    segs = [(oracle.SIDE_LEFT, 0, gmj, 0), (oracle.SIDE_ABOVE, 0, gmi, 1), (oracle.SIDE_RIGHT, 0, gmj, 2),
            (oracle.SIDE_BELOW, 0, gmi, 3)]
    return oracle.make_geom(dict(gmi=gmi, gmj=gmj, periodic_x=False, segments=segs,
                                 object_phi=[0.0] * (4 + len(boxes)), n_boundary_obj=4, inner_boxes=list(boxes)))


def test_inner_crossing_hand_cases(oracle):
    """Crossings of straight trajectories with the box [10,14]x[10,13] worked by hand: the face
    crossed (dirflag: left face 3, top 4, right 1, bottom 2), the crossing point, the tie at a
    corner going to the first face in the order left, top, right, bottom, and the nearer of two
    boxes."""
    A = (10, 14, 10, 13)
    g = _walled(oracle, [A])
    cases = [((8.0, 11.0, 11.0, 12.5), (3, 10.0, 12.0)),        # left face at t = 2/3
             ((12.0, 15.0, 12.5, 12.0), (4, 12.0 + 0.5 * 2 / 3, 13.0)),  # top face at t = 2/3
             ((16.0, 11.0, 13.0, 11.5), (1, 14.0, 11.0 + 0.5 * 2 / 3)),  # right face at t = 2/3
             ((12.0, 8.0, 12.4, 10.2), (2, 12.0 + 0.4 * 2 / 2.2, 10.0)),  # bottom face at t = 2/2.2
             ((8.0, 14.0, 11.0, 12.0), (3, 10.0, 14.0 - 2.0 * 2 / 3)),  # misses the top face's span
             ((9.0, 14.0, 11.0, 12.0), (3, 10.0, 13.0))]        # through the corner: left face first
    for (xo, yo, x, y), (d, xc, yc) in cases:
        rc, obj, dirflag, gx, gy = oracle.inner_crossing(g, xo, yo, x, y)
        assert rc == 0 and obj == 4 and dirflag == d
        assert gx == pytest.approx(xc, abs=1e-13) and gy == pytest.approx(yc, abs=1e-13)
    # two boxes on the way: the crossing nearer the origin wins
    g2 = _walled(oracle, [(13, 15, 10, 12), (10, 12, 10, 12)])
    rc, obj, dirflag, gx, gy = oracle.inner_crossing(g2, 9.0, 11.0, 14.0, 11.0)
    assert rc == 0 and obj == 5 and dirflag == 3 and gx == 10.0 and gy == 11.0


def test_inner_crossing_skips_covered_cells(oracle):
    """Faces covered by an attached box or a boundary wall on the same line are not crossed
    (pic2d_CurProblemValues.f90:600-680): between two boxes sharing the face x = 14 there is
    no crossing (the reference aborts); a box standing on the wall below is entered through its
    top face only; uncovered cells of a partly covered face still count."""
    g = _walled(oracle, [(10, 14, 10, 13), (14, 18, 10, 13)])
    rc, *_ = oracle.inner_crossing(g, 15.0, 11.5, 13.0, 11.5)
    assert rc != 0
    g = _walled(oracle, [(10, 14, 0, 3)])
    rc, obj, dirflag, gx, gy = oracle.inner_crossing(g, 12.0, -0.5, 12.0, 3.5)
    assert rc == 0 and dirflag == 4 and gy == 3.0  # the bottom face on the wall is skipped
    # box (14,18,11,12) covers only cells 11 of the face x = 14 of the box (10,14,10,13)
    g = _walled(oracle, [(10, 14, 10, 13), (14, 18, 11, 12)])
    rc, obj, dirflag, gx, gy = oracle.inner_crossing(g, 15.0, 10.5, 13.0, 10.5)
    assert rc == 0 and obj == 4 and dirflag == 1 and gx == 14.0 and gy == 10.5
    rc, *_ = oracle.inner_crossing(g, 15.0, 11.5, 13.0, 11.5)
    assert rc != 0


def test_advance_records_inner_hits(oracle):
    """orc_advance_ex charges an electron that ends inside a box to the box whose face its step
    crosses, at the crossing; ions go back N_subcycles steps; boundary-wall removals record the
    position after the push and the wall's dirflag."""
    box = (10, 14, 10, 13)
    g = _walled(oracle, [box])
    nn = 41 * 31
    zero = np.zeros(nn)
    X = np.array([9.5, 12.0, 20.0])
    Y = np.array([11.0, 13.6, 0.3])
    VX = np.array([1.0, 0.2, 0.1])
    VY = np.array([0.5, -1.2, -0.5])
    VZ = np.zeros(3)
    rc, rem, obj, hits, rx, ry, rd = oracle.advance_records(g, oracle.make_species(True), zero, zero,
                                                            (0.0, 0.0, 0.0), 0.0, X, Y, VX, VY, VZ, 5)
    assert rc == 0 and list(rem) == [1, 1, 1] and list(obj) == [4, 4, 3]
    assert rd[0] == 3 and rx[0] == 10.0 and ry[0] == pytest.approx(11.25, abs=1e-14)
    assert rd[1] == 4 and ry[1] == 13.0 and rx[1] == pytest.approx(12.0 + 0.2 * 0.6 / 1.2, abs=1e-14)
    assert rd[2] == 4 and rx[2] == X[2] and ry[2] == Y[2]
    # an ion with N_subcycles = 3 moves 3 v and goes back as far
    sp = oracle.make_species(False, QM2s=1e-6, N_subcycles=3, sense_B=0)
    X, Y = np.array([8.6]), np.array([11.0])
    VX, VY, VZ = np.array([0.5]), np.array([0.25]), np.zeros(1)
    rc, rem, obj, hits, rx, ry, rd = oracle.advance_records(g, sp, zero, zero, (0.0, 0.0, 0.0), 0.0, X, Y, VX, VY,
                                                            VZ, 5)
    assert rc == 0 and rem[0] == 1 and rd[0] == 3 and rx[0] == 10.0
    assert ry[0] == pytest.approx(11.0 + 0.75 * 1.4 / 1.5, abs=1e-13)
