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
