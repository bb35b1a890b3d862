"""Pinning of the wall-emission oracle (oracle.c orc_see_*, orc_constant_emission; CPU only):
yields at hand-worked energies and angles, the process chosen at the thresholds of the
first uniform, the reflection kinematics, the number of true secondaries for yields above
one, their energy budget, and the counts and positions of the constant emission."""
import math

import numpy as np
import pytest

import oracle_ffi as O

EF = 50.0  # energy_factor_eV of the workloads (Te = 2 eV, N_max_vel = 5)


def _see(**m):
    return O.see_from_material(m, EF, 2.0, 5)


def _tables():
    return O.maxwell_tables()


def test_model1_windows():
    o = _see(model_true=1, setD_true=0.5, minE_true_eV=10.0, maxE_true_eV=100.0)
    c = lambda e_eV: O.lib().orc_see_coeff(O.C.byref(o), 2, e_eV / EF, 0.0)
    assert c(10.0) == 0.0 and c(10.0 + 1e-9) == 0.5 and c(100.0) == 0.5 and c(100.0 + 1e-9) == 0.0
    assert o.lowest == 10.0 / EF


def test_classic_yield_maximum_and_angles():
    o = _see(model_true=2, E_see_0_eV=10.0, E_see_max_eV=200.0, maxD_see_classic=2.0, k_smooth=1.0)
    cl = lambda e_eV, th: O.lib().orc_see_coeff(O.C.byref(o), 3, e_eV / EF, th)
    # the maximum sits at E_max with value maxD for normal incidence
    assert cl(200.0, 0.0) == pytest.approx(2.0, rel=1e-14)
    assert cl(10.0, 0.0) == 0.0
    # at angle theta the maximum moves to E_max (1 + k theta^2 / pi) and grows by 1 + k theta^2 / (2 pi)
    th = 0.7
    emax = 200.0 * (1.0 + th * th / math.pi)
    assert cl(emax, th) == pytest.approx(2.0 * (1.0 + th * th / (2.0 * math.pi)), rel=1e-8)
    # twice the distance above threshold: v = 2, exponent 0.25: (2/e)^(1/4)
    assert cl(10.0 + 2 * 190.0, 0.0) == pytest.approx(2.0 * (2.0 / math.e) ** 0.25, rel=1e-14)
    # true-secondary model 2 = classic x (1 - fractions taken by the backscattering models)
    o2 = _see(model_true=2, model_inelastic=2, frac_inelastic=0.1, E_see_0_eV=10.0, E_see_max_eV=200.0,
              maxD_see_classic=2.0, k_smooth=1.0)
    t = O.lib().orc_see_coeff(O.C.byref(o2), 2, 300.0 / EF, 0.2)
    c = O.lib().orc_see_coeff(O.C.byref(o2), 3, 300.0 / EF, 0.2)
    i = O.lib().orc_see_coeff(O.C.byref(o2), 1, 300.0 / EF, 0.2)
    assert t == pytest.approx(0.9 * c, rel=1e-14) and i == pytest.approx(0.1 * c, rel=1e-14)


def test_elastic_model2_continuous_at_its_maximum():
    # the classic-yield parameters always come with the material (its fraction here is 0)
    o = _see(model_elastic=2, E_elast_0_eV=0.5, E_elast_max_eV=5.0, maxD_elast=0.4, dE_elast_eV=8.0,
             E_see_0_eV=10.0, E_see_max_eV=200.0, maxD_see_classic=2.0)
    c = lambda e_eV: O.lib().orc_see_coeff(O.C.byref(o), 0, e_eV / EF, 0.0)
    assert c(5.0) == pytest.approx(0.4, rel=1e-14)
    assert c(5.0 * (1 - 1e-12)) == pytest.approx(0.4, rel=1e-9)
    assert c(0.5) == 0.0
    # decaying side: (E - Emax + dE)/dE exp(-(E - Emax)/dE) at E = Emax + dE: 2/e
    assert c(13.0) == pytest.approx(0.4 * 2.0 / math.e, rel=1e-13)


def _emit(o, dirflag, v, u, x=17.25, y=-0.3):
    maxw, inj = _tables()
    return O.see_emit(o, maxw, inj, dirflag, 32.0, 32.0, x, y, v[0], v[1], v[2], u)


def test_process_chosen_by_the_first_uniform():
    o = _see(model_elastic=1, setD_elastic=0.25, minE_elastic_eV=0.0, maxE_elastic_eV=1e4,
             model_inelastic=1, setD_inelastic=0.25, minE_inelastic_eV=0.0, maxE_inelastic_eV=1e4,
             model_true=1, setD_true=0.25, minE_true_eV=1.0, maxE_true_eV=1e4, T_true_eV=2.0)
    v = (0.3, -1.2, 0.1)
    speed = math.sqrt(sum(c * c for c in v))
    rest = [0.5, 0.5, 0.64]
    rc, e = _emit(o, 4, v, [0.25] + rest)       # elastic, specular
    assert rc == 0 and np.array_equal(e[0], [17.25, 1e-6, 0.3, 1.2, 0.1])
    rc, e = _emit(o, 4, v, [0.25 + 1e-12] + rest)  # inelastic: speed v sqrt(u3), leaves the wall
    assert rc == 0 and e.shape == (1, 5)
    assert np.linalg.norm(e[0, 2:]) == pytest.approx(speed * 0.8, rel=1e-14) and e[0, 3] > 0
    rc, e = _emit(o, 4, v, [0.5] + rest)
    assert e.shape == (1, 5) and np.linalg.norm(e[0, 2:]) == pytest.approx(speed * 0.8, rel=1e-14)
    rc, e = _emit(o, 4, v, [0.5 + 1e-12] + rest)  # true secondary
    assert e.shape == (1, 5) and e[0, 3] > 0
    rc, e = _emit(o, 4, v, [0.75 + 1e-12] + rest)  # attached to the wall
    assert e.shape == (0, 5)
    # the wall above mirrors the normal component and the position
    rc, e = _emit(o, 2, (0.3, 1.2, 0.1), [0.1], y=32.4)
    assert np.array_equal(e[0], [17.25, 32.0 - 1e-6, 0.3, -1.2, 0.1])
    # left wall: the clamp acts on y and the normal is x
    rc, e = _emit(o, 1, (-1.2, 0.3, 0.1), [0.1], x=-0.2, y=40.0)
    assert np.array_equal(e[0], [1e-6, 32.0, 1.2, 0.3, 0.1])


def test_random_angle_reflection():
    o = _see(model_elastic=1, setD_elastic=1.0, minE_elastic_eV=0.0, maxE_elastic_eV=1e4, elastic_reflection=1)
    v = (0.3, -1.2, 0.1)
    speed = math.sqrt(sum(c * c for c in v))
    u1, u2 = 0.6, 0.3
    rc, e = _emit(o, 4, v, [0.5, u1, u2])
    vx, vy, vz = e[0, 2:]
    assert math.sqrt(vx * vx + vy * vy + vz * vz) == pytest.approx(speed, rel=1e-14)
    assert vy / speed == pytest.approx(math.sqrt(1 - u1 * u1), rel=1e-13)  # theta = asin(u1) from the normal
    assert math.atan2(vx, vz) == pytest.approx(2 * math.pi * u2, rel=1e-11)


def test_yield_above_one_and_energy_budget():
    o = _see(model_true=1, setD_true=2.3, minE_true_eV=1.0, maxE_true_eV=1e4, T_true_eV=2.0)
    v = (0.0, -2.0, 0.0)  # 200 eV
    u = [0.9] + [0.5, 0.7, 0.5] * 2 + [0.29] + [0.5, 0.7, 0.5]
    rc, e = _emit(o, 4, v, u)
    assert e.shape[0] == 3  # 2 whole secondaries, then 0.3 > 0.29
    u[7] = 0.31
    rc, e = _emit(o, 4, v, u)
    assert e.shape[0] == 2
    # hot secondaries exhaust the budget E - E_min: the last one is scaled to the rest, then none
    hot = _see(model_true=1, setD_true=3.0, minE_true_eV=150.0, maxE_true_eV=1e4, T_true_eV=400.0)
    rc, e = _emit(hot, 4, v, [0.9] + [0.9, 0.99, 0.9] * 3 + [0.0, 0.9, 0.9, 0.9])
    en = np.sum(e[:, 2:] ** 2, axis=1)
    assert e.shape[0] >= 1 and np.sum(en) == pytest.approx((200.0 - 150.0) / EF, rel=1e-13)
    # below the lowest emission energy nothing happens and no uniform is needed
    rc, e = _emit(hot, 4, (0.0, -1.7, 0.0), [])
    assert rc == 0 and e.shape[0] == 0


def test_backscatter_above_one_is_an_error():
    o = _see(model_elastic=1, setD_elastic=0.7, minE_elastic_eV=0.0, maxE_elastic_eV=1e4,
             model_inelastic=1, setD_inelastic=0.5, minE_inelastic_eV=0.0, maxE_inelastic_eV=1e4)
    rc, e = _emit(o, 4, (0.0, -1.0, 0.0), [0.1])
    assert rc == 5


def test_constant_emission_counts_positions_and_beam():
    maxw, inj = _tables()
    fn, fp, vb = O.emission_factors(1, 1.0, 0.1, 100.0, 2.0, 5)
    assert vb == pytest.approx(math.sqrt(50.0) / 5)
    n_float = float(np.float32(10.3))
    rng = np.random.default_rng(3)
    u = rng.uniform(0, 1, 4 * 12)
    u[1::4] = 0.5  # Maxwell table midpoint: zero spread, the beam speed exactly
    u[40] = n_float - 10 - 1e-9  # fractional electron emitted
    out = O.constant_emission(4, 1, 33, 32, 0, 33, 32, 10.3, 1, fn, fp, vb, maxw, inj, u)
    assert out.shape[0] == 11
    assert np.all(out[:, 1] == 1e-6) and np.all(out[:, 3] == vb)
    assert np.all((out[:, 0] >= 0) & (out[:, 0] <= 32.0))
    assert np.array_equal(out[:10, 0], 0.0 + u[0:40:4] * 32.0)
    u[40] = n_float - 10 + 1e-9
    assert O.constant_emission(4, 1, 33, 32, 0, 33, 32, 10.3, 1, fn, fp, vb, maxw, inj, u).shape[0] == 10
    # thermal emission from the wall above points down, with the half-Maxwellian table
    fn0, fp0, _ = O.emission_factors(0, 2.0, 2.0, 0.0, 2.0, 5)
    out = O.constant_emission(2, 0, 33, 32, 0, 33, 33, 3.0, 0, fn0, fp0, 0.0, maxw, inj, np.full(20, 0.7))
    assert out.shape[0] == 3 and np.all(out[:, 3] < 0) and np.all(out[:, 1] == 32.0 - 1e-6)
    assert out[0, 3] == -O.lib().orc_maxwell_inj_sample(O.dp(inj), 0.7) * fn0


def test_inner_object_emission_leaves_the_hit_face():
    """The inner-object branches of INJECT_* (pic2d_SEEmission.f90:372-549, 696-790): the same
    processes, counts and velocities as at a boundary wall facing the same way, the electrons
    placed 1e-6 outside the face that was hit (right face: x = iright + 1e-6, dirflag 1; bottom
    face: y = jbottom - 1e-6, dirflag 2; left face: x = ileft - 1e-6, dirflag 3; top face:
    y = jtop + 1e-6, dirflag 4) at the crossing's other coordinate."""
    maxw, inj = _tables()
    o = _see(model_elastic=1, setD_elastic=0.2, minE_elastic_eV=0.0, maxE_elastic_eV=1e4, elastic_reflection=1,
             model_inelastic=1, setD_inelastic=0.2, minE_inelastic_eV=0.0, maxE_inelastic_eV=1e4,
             model_true=1, setD_true=1.4, minE_true_eV=1.0, maxE_true_eV=1e4, T_true_eV=2.0)
    box = (10, 14, 10, 13)
    rng = np.random.default_rng(3)
    want = {1: (14.0 + 1e-6, None), 2: (None, 10.0 - 1e-6), 3: (10.0 - 1e-6, None), 4: (None, 13.0 + 1e-6)}
    for dirflag, (wx, wy) in want.items():
        for _ in range(20):
            x, y = (14.0, 11.7) if dirflag in (1, 3) else (12.3, 13.0)
            v = rng.normal(0.0, 1.0, 3)
            u = rng.uniform(0.0, 1.0, 32)
            rc, e_in = O.see_emit_inner(o, maxw, inj, dirflag, box, x, y, *v, u)
            rc2, e_w = O.see_emit(o, maxw, inj, dirflag, 40.0, 30.0, x, y, *v, u)
            assert rc == rc2 == 0 and e_in.shape == e_w.shape
            assert np.array_equal(e_in[:, 2:], e_w[:, 2:])
            if e_in.shape[0]:
                assert np.all(e_in[:, 0] == (wx if wx is not None else x))
                assert np.all(e_in[:, 1] == (wy if wy is not None else y))
