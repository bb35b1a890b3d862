"""Multi-step runs of the same problem on the oracle (reference form: WELL stream, e-n
collisions by sampling without replacement) and on the GPU (its production path: sorted
tiles, fused collisions, counter-based RNG), reduced to the observables compared over
seed ensembles: density and mean-energy profiles across the gap, the electron energy
distribution, wall fluxes and collision counts."""
import ctypes as C

import numpy as np

from edgpu import config

N_YBINS = 8
EEDF_EDGES = np.array([0.0, 0.5, 1.0, 1.5, 2.0, 3.0, 4.0, 6.0, 8.0, 12.0, 1e9])


def observables(prob, e, i, hits_e, hits_i, coll):
    """e, i: particle arrays (x, y, vx, vy, vz); hits_*: wall hits per object over the run."""
    y_e, y_i = e[1], i[1]
    en = (e[2] ** 2 + e[3] ** 2 + e[4] ** 2) * prob.energy_factor_eV
    edges = np.linspace(0.0, prob.gmj, N_YBINS + 1)
    be = np.clip(np.digitize(y_e, edges) - 1, 0, N_YBINS - 1)
    bi = np.clip(np.digitize(y_i, edges) - 1, 0, N_YBINS - 1)
    n0 = prob.n_particles / N_YBINS
    dens_e = np.bincount(be, minlength=N_YBINS) / n0
    dens_i = np.bincount(bi, minlength=N_YBINS) / n0
    energy = np.bincount(be, weights=en, minlength=N_YBINS) / np.maximum(np.bincount(be, minlength=N_YBINS), 1)
    eedf = np.histogram(en, bins=EEDF_EDGES)[0] / max(en.size, 1)
    return dict(dens_e=dens_e, dens_i=dens_i, energy=energy, eedf=eedf,
                flux_e=np.array([hits_e[1], hits_e[3]], float), flux_i=np.array([hits_i[1], hits_i[3]], float),
                coll=np.asarray(coll, float))


def oracle_run(O, prob, seed, steps):
    w = O.well(seed)
    v, _ = O.maxwell_tables()
    n = prob.n_particles
    e = [np.empty(n) for _ in range(5)]
    O.lib().orc_distribute(C.byref(w), O.dp(v), n, prob.x_limits[0], prob.x_limits[1], prob.y_limits[0],
                           prob.y_limits[1], 0.0, float(prob.gmi), 0.0, float(prob.gmj), prob.factor_convert_e,
                           *[O.dp(a) for a in e])
    i = [np.empty(n) for _ in range(5)]
    O.lib().orc_distribute_ions(C.byref(w), O.dp(v), n, O.dp(e[0]), O.dp(e[1]), prob.factor_convert_i,
                                *[O.dp(a) for a in i])
    g = O.make_geom(prob.geometry())
    tab = O.MccTable([config.ENERGY_SEGMENTS[0]] + [s[0] for s in config.ENERGY_SEGMENTS[1]],
                     [s[1] for s in config.ENERGY_SEGMENTS[1]],
                     [dict(p, ion_velocity_factor=prob.ion_velocity_factor()) for p in prob.cross_sections()],
                     prob.M_i_amu, prob.neutral_N_m3, prob.N_subcycles, prob.dt, prob.V_scale)
    rcx = O.Rcx()
    s1, alpha = prob.rcx_sigma()
    O.lib().orc_rcx_setup(C.byref(rcx), s1, alpha, prob.neutral_N_m3, prob.neutral_T_K, prob.dt, prob.N_subcycles,
                          prob.Ms, prob.energy_factor_eV, prob.V_scale, prob.T_e_eV, prob.N_max_vel)
    a_eq = O.prepare_sys_y(prob.gmi - 1, prob.gmj)
    sp_e = O.make_species(True)
    sp_i = O.make_species(False, Qs=prob.Qs, QM2s=prob.QM2s, N_subcycles=1, sense_B=1)
    B = (prob.Bx, prob.By, prob.Bz)
    nn = prob.nx * prob.ny
    accEX, accEY = np.zeros(nn), np.zeros(nn)
    hits_e = np.zeros(4, np.int64)
    hits_i = np.zeros(4, np.int64)
    coll = np.zeros(8, np.int64)
    n_rcx = 0
    for _ in range(steps):
        rho_e = O.fold_walls(g, O.deposit(g, 1, e[0], e[1]))
        rho_i = O.fold_walls(g, O.deposit(g, 1, i[0], i[1]))
        phi = O.solve_poisson_fftx(g, prob.N_of_particles_cell, rho_e, rho_i, a_eq)
        EX, EY = O.efield(g, phi, prob.factor1_E_from_F, prob.factor2_E_from_F, (accEX, accEY))
        for arrs, sp, fx, fy, hits in ((e, sp_e, EX, EY, hits_e), (i, sp_i, accEX, accEY, hits_i)):
            rc, rem, _, h = O.advance(g, sp, fx, fy, B, prob.Ez, *arrs, 4)
            assert rc == 0
            hits += h
            keep = rem == 0
            for q in range(5):
                arrs[q] = np.ascontiguousarray(arrs[q][keep])
        m = e[0].size
        kind = np.zeros(m, np.int32)
        nnew = C.c_int64(0)
        cap = max(1024, m // 20)
        enew, inew = np.zeros(5 * cap), np.zeros(5 * cap)
        rc = O.lib().orc_mcc_en_reference(C.byref(tab.t), O.dp(v), C.byref(w), m, O.dp(e[0]), O.dp(e[1]),
                                          O.dp(e[2]), O.dp(e[3]), O.dp(e[4]), O.ip(kind), O.lp(coll), cap,
                                          C.byref(nnew), O.dp(enew), O.dp(inew))
        assert rc == 0
        n_rcx += O.lib().orc_rcx_reference(C.byref(rcx), O.dp(v), C.byref(w), i[0].size, O.dp(i[2]), O.dp(i[3]),
                                           O.dp(i[4]))
        k = nnew.value
        if k:
            ne, ni = enew[:5 * k].reshape(k, 5), inew[:5 * k].reshape(k, 5)
            for q in range(5):
                e[q] = np.concatenate([e[q], ne[:, q]])
                i[q] = np.concatenate([i[q], ni[:, q]])
        accEX[:] = 0.0
        accEY[:] = 0.0
    return observables(prob, e, i, hits_e, hits_i, list(coll[:3]) + [n_rcx])


def gpu_run(edgpu, prob, seed, steps, precision, sort_interval=4):
    rt = edgpu.Runtime(prob, precision=precision, seed=seed, sort_interval=sort_interval)
    try:
        rt.initiate_electron_neutral_collisions()
        rt.initiate_ion_neutral_collisions()
        rt.load_initial_particles()
        hits_e = np.zeros(4, np.int64)
        hits_i = np.zeros(4, np.int64)
        coll = np.zeros(4, np.int64)
        for _ in range(steps):
            se, si = rt.step()
            hits_e += np.array(se["hit_count"][:4])
            hits_i += np.array(si["hit_count"][:4])
            coll[:3] += np.array(se["collisions"][:3])
            coll[3] += si["collisions"][0]
        e = rt.download(0)[:5]
        i = rt.download(1)[:5]
        return observables(prob, e, i, hits_e, hits_i, coll)
    finally:
        rt.close()


def ensemble(runs):
    keys = runs[0].keys()
    return {k: (np.mean([r[k] for r in runs], axis=0), np.std([r[k] for r in runs], axis=0, ddof=1))
            for k in keys}


def compare(a, b, n_a, n_b, sigmas=4.0, floors=None):
    """Per-observable check |mean_a - mean_b| <= sigmas * standard error of the difference +
    floor.  Returns a list of (key, index, difference, tolerance) violations."""
    floors = floors or {}
    bad = []
    for k in a:
        ma, sa = a[k]
        mb, sb = b[k]
        se = np.sqrt(sa ** 2 / n_a + sb ** 2 / n_b)
        tol = sigmas * se + floors.get(k, 0.0)
        d = np.abs(np.atleast_1d(ma) - np.atleast_1d(mb))
        for j in np.nonzero(d > np.atleast_1d(tol))[0]:
            bad.append((k, int(j), float(d[j]), float(np.atleast_1d(tol)[j])))
    return bad
