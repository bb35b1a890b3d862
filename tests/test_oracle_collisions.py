"""Pinning of the oracle's collision routines (CPU only): the collision-frequency table
against the constant-cross-section closed form, the process selection against a
brute-force evaluation of the interpolated table, the scattering kinematics against
their conservation laws and the published angle laws, and the charge-exchange test
against its closed-form probability."""
import ctypes as C
import math

import numpy as np
import pytest

E_CL = 1.602189e-19
M_E = 9.109534e-31


def _freq(oracle, E, typ, Et, St, N=1e20, nsub=1, dt=5.6e-11):
    Et = np.asarray(Et, np.float64)
    St = np.asarray(St, np.float64)
    return oracle.lib().orc_en_frequency(E, typ, Et.size, oracle.dp(Et), oracle.dp(St), N, nsub, dt)


def test_en_frequency_constant_sigma(oracle):
    """sigma constant: f = sigma sqrt(2 e E / m) n N_sub dt at the table points, beyond the
    last point and (elastic only) below the first; chords in between."""
    Et = [0.5, 2.0, 10.0, 50.0]
    St = [1e-19] * 4
    scale = 1e20 * 3 * 5.6e-11
    for E in [0.5, 2.0, 10.0, 50.0, 80.0, 200.0]:
        want = 1e-19 * math.sqrt(2 * E * E_CL / M_E) * scale
        assert _freq(oracle, E, 10, Et, St, nsub=3) == pytest.approx(want, rel=1e-14)
    E = 0.1
    assert _freq(oracle, E, 10, Et, St, nsub=3) == pytest.approx(1e-19 * math.sqrt(2 * E * E_CL / M_E) * scale,
                                                                  rel=1e-14)
    assert _freq(oracle, E, 20, Et, St, nsub=3) == 0.0  # inelastic below threshold
    # between points the interpolation is linear in f: the chord lies below sqrt
    for E in [1.0, 5.0, 30.0]:
        f = _freq(oracle, E, 10, Et, St, nsub=3)
        assert f < 1e-19 * math.sqrt(2 * E * E_CL / M_E) * scale
        j = np.searchsorted(Et, E)
        f0 = 1e-19 * math.sqrt(2 * Et[j - 1] * E_CL / M_E) * scale
        f1 = 1e-19 * math.sqrt(2 * Et[j] * E_CL / M_E) * scale
        assert f == pytest.approx(f0 + (f1 - f0) * (E - Et[j - 1]) / (Et[j] - Et[j - 1]), rel=1e-14)


def _table(oracle, procs=None):
    from edgpu import config
    prob = config.workload(1, 8)
    procs = procs or prob.cross_sections()
    sv = [config.ENERGY_SEGMENTS[0]] + [s[0] for s in config.ENERGY_SEGMENTS[1]]
    st = [s[1] for s in config.ENERGY_SEGMENTS[1]]
    T = oracle.MccTable(sv, st, procs, prob.M_i_amu, prob.neutral_N_m3, prob.N_subcycles, prob.dt, prob.V_scale)
    return T, prob, procs


def test_mcc_table_normalisation(oracle):
    """Cumulative frequencies normalised by their maximum; P_max = 1 - exp(-nu_max)."""
    T, prob, procs = _table(oracle)
    P = T.prob2d
    assert np.all(np.diff(P, axis=1) >= 0)  # cumulative over processes
    assert P.max() == 1.0 and P.min() >= 0.0
    nu = np.array([sum(_freq(oracle, e, p["type"], p["E"], p["sigma"], prob.neutral_N_m3, prob.N_subcycles,
                             prob.dt) for p in procs) for e in T.energy])
    assert T.t.max_colliding_fraction == pytest.approx(1 - math.exp(-nu.max()), rel=1e-14)
    assert np.allclose(P[:, -1], np.minimum(nu / nu.max(), 1.0), rtol=1e-13, atol=1e-300)


def test_mcc_select_brute_force(oracle):
    """(E, r) -> process: first k from the top with r > P_k(E), else null, where P_k(E)
    interpolates the table linearly between the bracketing table energies."""
    T, prob, procs = _table(oracle)
    n_proc = len(procs)
    rng = np.random.default_rng(8)
    E = np.concatenate([rng.uniform(0.0, 30.0, 3000), np.exp(rng.uniform(-3, 9.5, 3000)), T.energy[:200]])
    R = rng.uniform(0, 1, E.size)
    idx = C.c_int32()
    a0, a1 = C.c_double(), C.c_double()
    coll = C.c_int32()
    L = oracle.lib()
    near = 0
    for e, r in zip(E, R):
        assert L.orc_mcc_select(C.byref(T.t), e, r, C.byref(idx), C.byref(a0), C.byref(a1), C.byref(coll)) == 0
        if e >= T.energy[-1]:
            Pk = T.prob2d[-1]
        else:
            i = int(np.searchsorted(T.energy, e, side="right") - 1)
            w = (e - T.energy[i]) / (T.energy[i + 1] - T.energy[i])
            Pk = (1 - w) * T.prob2d[i] + w * T.prob2d[i + 1]
        want = n_proc + 1
        for k in range(n_proc, 0, -1):
            if r > Pk[k - 1]:
                want = k + 1
                break
        else:
            want = 1
        if coll.value != want:
            # only a rounding-level tie with a table threshold may differ
            assert np.min(np.abs(Pk - r)) < 1e-12
            near += 1
    assert near <= 2


def _rot_angle(v, w):
    v = np.asarray(v)
    w = np.asarray(w)
    return float(np.dot(v, w) / (np.linalg.norm(v) * np.linalg.norm(w)))


def _scatter(oracle, typ, M, Eth, ivf, E, u, v):
    maxw, _ = oracle.maxwell_tables()
    vx, vy, vz = C.c_double(v[0]), C.c_double(v[1]), C.c_double(v[2])
    prod = C.c_int32()
    ej = np.zeros(3)
    iv = np.zeros(3)
    uu = np.asarray(u, np.float64)
    used = oracle.lib().orc_en_scatter(typ, M, Eth, ivf, oracle.dp(maxw), E, oracle.dp(uu), C.byref(vx), C.byref(vy),
                                       C.byref(vz), C.byref(prod), oracle.dp(ej), oracle.dp(iv))
    return used, np.array([vx.value, vy.value, vz.value]), prod.value, ej, iv


def _okhrimovskyy(E, R):
    # Okhrimovskyy, Bogaerts, Gijbels, Phys. Rev. E 65, 037402 (2002): screened Coulomb
    # angular law with epsilon = E / 27.21 eV
    return 1.0 - 2.0 * R / (1.0 + 8.0 * (E / 27.21) * (1.0 - R))


@pytest.mark.parametrize("typ", [10, 11])
def test_elastic_kinematics(oracle, typ):
    rng = np.random.default_rng(typ)
    M = 40.0
    for _ in range(200):
        v = rng.normal(0, 1, 3)
        E = float(rng.uniform(0.1, 500.0))
        u = rng.uniform(0, 1, 2)
        used, w, prod, _, _ = _scatter(oracle, typ, M, 0.0, 0.0, E, u, v)
        assert used == 2 and prod == 0
        cos_chi = _okhrimovskyy(E, u[0]) if typ == 10 else 1 - 2 * u[0]
        cos_chi = max(min(cos_chi, 0.999999999999), -0.999999999999)
        assert _rot_angle(v, w) == pytest.approx(cos_chi, abs=1e-11)
        # energy loss to the neutral: |v'|^2 = |v|^2 (1 - (2 m_e / M)(1 - cos chi))
        two_m_over_M = 2 * M_E / (M * 1.660565e-27)
        assert np.dot(w, w) == pytest.approx(np.dot(v, v) * (1 - two_m_over_M * (1 - cos_chi)), rel=1e-6)


def test_inelastic_and_ionisation_energy_balance(oracle):
    rng = np.random.default_rng(20)
    Eth = 11.5
    for _ in range(200):
        v = rng.normal(0, 1, 3)
        E = float(rng.uniform(12.0, 800.0))
        u = rng.uniform(0, 1, 2)
        used, w, _, _, _ = _scatter(oracle, 20, 40.0, Eth, 0.0, E, u, v)
        assert used == 2
        assert np.dot(w, w) == pytest.approx(np.dot(v, v) * (E - Eth) / E, rel=1e-12)
        assert _rot_angle(v, w) == pytest.approx(max(min(_okhrimovskyy(E - Eth, u[0]), 1 - 1e-12), -1 + 1e-12),
                                                 abs=1e-11)
    # below threshold nothing happens and nothing is drawn
    used, w, _, _, _ = _scatter(oracle, 20, 40.0, Eth, 0.0, 10.0, [0.3, 0.7], [1.0, 0.0, 0.0])
    assert used == 0 and np.array_equal(w, [1.0, 0.0, 0.0])
    Ei = 15.76
    for _ in range(200):
        v = rng.normal(0, 1, 3)
        E = float(rng.uniform(16.0, 800.0))
        u = rng.uniform(0, 1, 8)
        used, w, prod, ej, iv = _scatter(oracle, 30, 40.0, Ei, 0.01, E, u, v)
        assert used == 8 and prod == 1
        e_sc = E * np.dot(w, w) / np.dot(v, v)
        e_ej = E * np.dot(ej, ej) / np.dot(v, v)
        assert e_sc + e_ej == pytest.approx(E - Ei, rel=1e-12)
        # ejected energy: B tan(R atan((E - I) / 2B)) with B = 10 eV, at most half the excess
        want = min(10.0 * math.tan(u[0] * math.atan(0.5 * (E - Ei) / 10.0)), 0.5 * (E - Ei))
        assert e_ej == pytest.approx(want, rel=1e-12)
        assert np.all(np.abs(iv) <= 3.0 * 0.01 + 1e-15)


def test_rcx_closed_form_rate(oracle):
    """alpha = 0: the exchange probability is sqrt((n dt N_sub |v| sigma)^2 + p_th^2)."""
    r = oracle.Rcx()
    maxw, _ = oracle.maxwell_tables()
    Ms = 40 * 1.660565e-27 / M_E
    V_scale = 4.19e6
    L = oracle.lib()
    L.orc_rcx_setup(C.byref(r), 5e-19, 0.0, 1e20, 300.0, 5.6e-11, 1, Ms, 0.5 * M_E * V_scale ** 2 / E_CL, V_scale,
                    2.0, 5)
    Tg = 300 * 1.38064852e-23 / E_CL
    vmean = math.sqrt(8 / math.pi * Tg * E_CL / (Ms * M_E))
    assert r.probab_thermal == pytest.approx(1e20 * 5e-19 * vmean * 5.6e-11, rel=1e-12)
    n = 400000
    speed = 2e-3  # normalised: |v| V_scale = 8.4e3 m/s
    VX = np.full(n, speed)
    VY = np.zeros(n)
    VZ = np.zeros(n)
    u = np.random.default_rng(9).uniform(0, 1, 4 * n)
    flag = np.zeros(n, np.uint8)
    c = L.orc_rcx_stream(C.byref(r), oracle.dp(maxw), n, oracle.dp(VX), oracle.dp(VY), oracle.dp(VZ),
                         oracle.dp(u), flag.ctypes.data_as(C.POINTER(C.c_uint8)))
    p = math.sqrt((1e20 * 5.6e-11 * speed * V_scale * 5e-19) ** 2 + r.probab_thermal ** 2)
    assert c == flag.sum()
    assert abs(c - p * n) < 5 * math.sqrt(n * p * (1 - p))
    # the exchanged ions are exactly those whose first uniform fell below p
    assert np.array_equal(flag.astype(bool), u[0::4] <= p)
    vf = math.sqrt(300 * 1.38064852e-23 / (2.0 * E_CL * Ms)) / 5
    assert r.vfactor == pytest.approx(vf, rel=1e-14)
    assert np.all(np.abs(VX[flag == 1]) <= 3 * vf * (1 + 1e-14))


# ---- the reference's sampled forms (orc_mcc_en_reference, orc_rcx_reference) -----------------
# pic2d_enCollisionsGeneralProc.f90:416-485 draws, from its one WELL stream: the fractional
# event's test (j = 0 only), random_r, random_n, then random_j = INT(u N) clamped to [1, N],
# redrawn while it names an electron that has already collided; the collision kinematics then
# draw 2 (elastic, inelastic) or 8 (ionisation) numbers.  PERFORM_RESONANT_CHARGE_EXCHANGE
# (pic2d_inCollisionsSpecificProc.f90:95-122) draws one number per ion and three more per
# exchange.

def _elastic_only(oracle):
    """One constant elastic cross-section: the total selection probability is 1 at and above
    the table's last energy (10 keV), so every sampled electron there collides."""
    procs = [dict(type=10, E=[0.0, 1.0, 100.0, 10000.0], sigma=[1e-19] * 4)]
    T, prob, _ = _table(oracle, procs)
    return T, prob


_MAXW = {}


def _mcc_reference(oracle, T, w, V):
    n = V[0].size
    if "v" not in _MAXW:
        _MAXW["v"] = oracle.maxwell_tables()[0]
    maxw = _MAXW["v"]
    V = [np.ascontiguousarray(a) for a in V]
    X = np.zeros(n)
    kind = np.zeros(n, np.int32)
    cnt = np.zeros(8, np.int64)
    nnew = C.c_int64(0)
    enew, inew = np.zeros(5 * n), np.zeros(5 * n)
    rc = oracle.lib().orc_mcc_en_reference(C.byref(T.t), oracle.dp(maxw), C.byref(w), n, oracle.dp(X), oracle.dp(X),
                                           *[oracle.dp(a) for a in V], oracle.ip(kind), oracle.lp(cnt), n,
                                           C.byref(nnew), oracle.dp(enew), oracle.dp(inew))
    assert rc == 0
    return V, kind, cnt


def _fast_electrons(prob, n, rng, energy_eV=20000.0):
    vv = rng.standard_normal((3, n))
    vv /= np.linalg.norm(vv, axis=0)
    speed = math.sqrt(energy_eV / prob.energy_factor_eV)
    return [vv[q] * speed for q in range(3)]


def test_mcc_reference_draw_order_against_second_stream(oracle):
    """Walk a second generator of the same seed through the reference's draw order and
    predict every event: which electron is picked (INT(u N) clamped to [1, N], redrawn while
    already collided) and its speed after the elastic collision from the closed form
    |v'|^2 = |v|^2 (1 - (2 m_e/M)(1 - cos chi)) with the Okhrimovskyy angle of the drawn
    uniform.  The sampler must collide exactly those electrons, leave those speeds, and have
    consumed exactly as many numbers."""
    T, prob = _elastic_only(oracle)
    pmax = T.t.max_colliding_fraction
    rng = np.random.default_rng(71)
    for n, seed in ((600, 11), (1000, 12), (173, 13)):
        V0 = _fast_electrons(prob, n, rng)
        w = oracle.well(seed)
        V, kind, cnt = _mcc_reference(oracle, T, w, [a.copy() for a in V0])
        u = oracle.well_fill(oracle.well(seed), 50 * n)
        at = 0
        R = pmax * n
        Ic = int(R)
        F = R - Ic
        picked, speed2 = [], {}
        two_m_over_M = 0.001097161 / T.t.M_amu
        for jj in range(Ic + 1):
            if jj == 0:
                at += 1
                if u[at - 1] > F:
                    continue
            at += 2  # random_r (every process selected: total probability 1), random_n
            while True:
                rj = min(max(int(u[at] * n), 1), n)
                at += 1
                if not picked or rj not in picked:
                    break
            k = rj - 1
            picked.append(rj)
            E = float((V0[0][k] ** 2 + V0[1][k] ** 2 + V0[2][k] ** 2) * T.t.energy_factor * M_E * 0.5 / E_CL)
            cos_chi = max(min(_okhrimovskyy(E, u[at]), 0.999999999999), -0.999999999999)
            at += 2
            speed2[k] = (V0[0][k] ** 2 + V0[1][k] ** 2 + V0[2][k] ** 2) * (1 - two_m_over_M * (1 - cos_chi))
        assert sorted(np.nonzero(kind)[0] + 1) == sorted(picked)
        assert cnt[0] == len(picked)
        for k, s2 in speed2.items():
            assert V[0][k] ** 2 + V[1][k] ** 2 + V[2][k] ** 2 == pytest.approx(s2, rel=1e-12)
        untouched = kind == 0
        for q in range(3):
            assert np.array_equal(V[q][untouched], V0[q][untouched])
        # the sampler left the stream exactly where the walk ends
        assert oracle.lib().orc_well_next(C.byref(w)) == u[at]


def test_mcc_reference_event_count_and_clamp(oracle):
    """Over many calls: each call makes INT(P_max N) or INT(P_max N) + 1 events, the second
    with probability F = P_max N - INT(P_max N); no electron collides twice in a call; the
    clamp of INT(u N) to [1, N] never picks electron N and picks electron 1 twice as often as
    any other (u in [0, 2/N) maps to it)."""
    T, prob = _elastic_only(oracle)
    pmax = T.t.max_colliding_fraction
    n = 50
    R = pmax * n
    Ic, F = int(R), R - int(R)
    rng = np.random.default_rng(73)
    V0 = _fast_electrons(prob, n, rng)
    w = oracle.well(2026)
    hits = np.zeros(n, np.int64)
    calls, extra = 6000, 0
    for _ in range(calls):
        _, kind, cnt = _mcc_reference(oracle, T, w, [a.copy() for a in V0])
        m = int((kind > 0).sum())
        assert cnt[0] == m and m in (Ic, Ic + 1)  # distinct electrons, one event each
        extra += m == Ic + 1
        hits += kind > 0
    assert abs(extra - F * calls) < 5 * math.sqrt(calls * F * (1 - F))
    assert hits[n - 1] == 0
    mean = hits[1:n - 1].mean()
    assert 1.6 < hits[0] / mean < 2.4
    assert np.all(np.abs(hits[1:n - 1] - mean) < 6 * math.sqrt(mean))


def test_rcx_reference_draws_match_stream_form(oracle):
    """PERFORM_RESONANT_CHARGE_EXCHANGE in the reference's form takes one number per ion and
    three more per exchange from the stream, and decides u <= p with the closed-form
    p = sqrt((n dt N_sub |v| V sigma)^2 + p_th^2): a second generator of the same seed, walked
    with that rule, gives the uniforms on which orc_rcx_stream reproduces the sampler's
    velocities exactly, and the sampler ends where the walk ends."""
    r = oracle.Rcx()
    maxw, _ = oracle.maxwell_tables()
    Ms = 40 * 1.660565e-27 / M_E
    V_scale = 4.19e6
    L = oracle.lib()
    L.orc_rcx_setup(C.byref(r), 5e-19, 0.0, 1e20, 300.0, 5.6e-11, 1, Ms, 0.5 * M_E * V_scale ** 2 / E_CL, V_scale,
                    2.0, 5)
    n = 20000
    rng = np.random.default_rng(79)
    # speeds spread so that p runs from the thermal floor to about 0.6
    p_target = rng.uniform(0.0, 0.6, n)
    speed = p_target / (1e20 * 5.6e-11 * V_scale * 5e-19)
    d = rng.standard_normal((3, n))
    d /= np.linalg.norm(d, axis=0)
    V0 = [np.ascontiguousarray(d[q] * speed) for q in range(3)]
    V = [a.copy() for a in V0]
    w = oracle.well(99)
    c = L.orc_rcx_reference(C.byref(r), oracle.dp(maxw), C.byref(w), n, *[oracle.dp(a) for a in V])
    u = oracle.well_fill(oracle.well(99), 5 * n)
    vabs = np.sqrt(V0[0] ** 2 + V0[1] ** 2 + V0[2] ** 2) * V_scale
    p = np.sqrt((1e20 * 5.6e-11 * vabs * 5e-19) ** 2 + r.probab_thermal ** 2)
    us = np.full((n, 4), 0.5)
    at, walked = 0, 0
    for k in range(n):
        us[k, 0] = u[at]
        at += 1
        assert abs(us[k, 0] - p[k]) > 1e-9 * p[k]  # no decision at the rounding level
        if us[k, 0] <= p[k]:
            us[k, 1:] = u[at:at + 3]
            at += 3
            walked += 1
    assert c == walked and 0.2 * n < c < 0.4 * n
    Vs = [a.copy() for a in V0]
    flag = np.zeros(n, np.uint8)
    c2 = L.orc_rcx_stream(C.byref(r), oracle.dp(maxw), n, *[oracle.dp(a) for a in Vs], oracle.dp(us.ravel()),
                          flag.ctypes.data_as(C.POINTER(C.c_uint8)))
    assert c2 == c
    for q in range(3):
        assert np.array_equal(V[q], Vs[q])
    assert L.orc_well_next(C.byref(w)) == u[at]


def test_mcc_select_index_rounding_at_grid_points(oracle):
    """The segment index is INTEGER + REAL(8) truncated (pic2d_enCollisionsGeneralProc.f90:466-
    467): for an energy one or two ulps below a grid point (100 eV = 60 + 8 x 5 eV) the sum
    rounds up to that point's index, the double check finds the energy below the interval and
    the reference aborts (:469-473).  The oracle returns that error there, and selects normally
    at the point itself and 1e-12 eV below it."""
    T, _, _ = _table(oracle)
    idx, coll = C.c_int32(), C.c_int32()
    a0, a1 = C.c_double(), C.c_double()

    def sel(e):
        return oracle.lib().orc_mcc_select(C.byref(T.t), float(e), 0.5, C.byref(idx), C.byref(a0), C.byref(a1),
                                           C.byref(coll))

    below = np.nextafter(100.0, 0.0)
    assert sel(100.0) == 0 and T.energy[idx.value] == 100.0
    assert sel(below) != 0 and T.energy[idx.value] == 100.0
    assert sel(100.0 - 1e-12) == 0 and T.energy[idx.value + 1] == 100.0
