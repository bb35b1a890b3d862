"""Problem set-up for the edgpu runtime: the reference's normalisation and the workloads.

Scales follow INITIATE_PARAMETERS (pic2d_CurProblemValues.f90:719-723, 893-903):
    v_Te = sqrt(2 Te e / m_e), w_pe = sqrt(n e^2 / (eps0 m_e)), lambda_D = v_Te / w_pe,
    dx = lambda_D / N_of_cells_debye, dt = dx / (N_max_vel v_Te),
    V_scale = N_max_vel v_Te, E_scale = m_e V_scale / (e dt), B_scale = m_e / (e dt),
    F_scale = e n dx^2 / eps0.
Ion species constants follow SET_ION_SPECIES (pic2d_CurProblemValues.f90:1503-1507).

Grid convention of the single cluster that covers the domain: nodes 0..gmi in x and
0..gmj in y; an X-periodic domain has FFT length N = gmi-1 (node N+1 is the periodic
image of node 1) and gmj cells between the walls at j=0 and j=gmj.
"""
import math

import numpy as np

E_CL = 1.602189e-19
M_E_KG = 9.109534e-31
EPS0 = 8.854188e-12
AMU_KG = 1.660565e-27
KB_JK = 1.38064852e-23

SIDE_LEFT, SIDE_ABOVE, SIDE_RIGHT, SIDE_BELOW = 0, 1, 2, 3

# energy segments of the probability table (input_data_*/init_neutral_Argon0.dat)
ENERGY_SEGMENTS = (0.0, [(0.04, 0.005), (0.1, 0.01), (0.4, 0.02), (1.0, 0.05), (10.0, 0.2), (20.0, 0.1),
                         (60.0, 0.5), (200.0, 5.0), (1000.0, 10.0), (10000.0, 50.0)])


class Problem:
    """Physical inputs of one run plus every derived scale the hot path needs."""

    def __init__(self, *, T_e_eV, N_plasma_m3, N_of_cells_debye, N_max_vel, N_of_particles_cell, gmi, gmj,
                 periodic_x, init_Te_eV=None, init_Ti_eV=0.026, M_i_amu=40.0, Qs=1, N_subcycles=1, Bz_T=0.0,
                 Bx_T=0.0, By_T=0.0, Ez_Vm=0.0, inner_boxes=(), inner_phi_V=(), wall_phi_V=0.0,
                 neutral_N_m3=1.0e20, neutral_T_K=300.0, collisions=True, ions_sense_magnetic_field=1,
                 wall_emission=None, init_Ne_m3=None, extra_ion_species=(), name="custom"):
        self.name = name
        # {"see": {object: edgpu_see_params fields}, "constant": {object: constant emission}}
        self.wall_emission = wall_emission or {}
        self.T_e_eV = T_e_eV
        self.N_plasma_m3 = N_plasma_m3
        self.N_of_cells_debye = N_of_cells_debye
        self.N_max_vel = N_max_vel
        self.N_of_particles_cell = N_of_particles_cell
        self.N_subcycles = N_subcycles
        self.gmi, self.gmj = gmi, gmj
        self.periodic_x = bool(periodic_x)
        self.init_Te_eV = T_e_eV if init_Te_eV is None else init_Te_eV
        self.init_Ti_eV = init_Ti_eV
        self.M_i_amu = M_i_amu
        self.Qs = Qs
        self.collisions = collisions
        self.neutral_N_m3 = neutral_N_m3
        self.neutral_T_K = neutral_T_K
        self.ions_sense_magnetic_field = ions_sense_magnetic_field
        # scales
        self.v_Te = math.sqrt(2.0 * T_e_eV * E_CL / M_E_KG)
        self.W_plasma = math.sqrt(N_plasma_m3 * E_CL ** 2 / (EPS0 * M_E_KG))
        self.L_debye = self.v_Te / self.W_plasma
        self.dx = self.L_debye / N_of_cells_debye
        self.dt = self.dx / (N_max_vel * self.v_Te)
        self.V_scale = N_max_vel * self.v_Te
        self.E_scale = M_E_KG * self.V_scale / (E_CL * self.dt)
        self.B_scale = M_E_KG / (E_CL * self.dt)
        self.F_scale = E_CL * N_plasma_m3 * self.dx ** 2 / EPS0
        self.energy_factor_eV = 0.5 * M_E_KG * self.V_scale ** 2 / E_CL
        self.factor1_E_from_F = self.F_scale / (self.E_scale * self.dx)
        self.factor2_E_from_F = self.F_scale / (self.E_scale * 2.0 * self.dx)
        self.Ms = (AMU_KG * M_i_amu) / M_E_KG
        self.QM2s = 0.5 * Qs / self.Ms
        # every ion species as (Qs, QM2s, Ms) (SET_ION_SPECIES, pic2d_CurProblemValues.f90:1467-1510):
        # the first is M_i_amu / Qs above, then extra_ion_species = [dict(M_i_amu=.., Qs=..), ...]
        self.ion_species = [(Qs, self.QM2s, self.Ms)]
        for sp in extra_ion_species:
            ms = (AMU_KG * sp["M_i_amu"]) / M_E_KG
            self.ion_species.append((int(sp["Qs"]), 0.5 * sp["Qs"] / ms, ms))
        self.Bx, self.By, self.Bz = Bx_T / self.B_scale, By_T / self.B_scale, Bz_T / self.B_scale
        self.Ez = Ez_Vm / self.E_scale
        # boundary: left(0) / above(1) / right(2) / below(3) objects, then inner boxes
        if self.periodic_x:
            # objects as in the periodic decks: 1 left periodic, 2 top wall, 3 right periodic, 4 bottom wall
            self.segments = [(SIDE_ABOVE, 0, gmi, 1), (SIDE_BELOW, 0, gmi, 3)]
            self.n_boundary_obj = 4
        else:
            self.segments = [(SIDE_LEFT, 0, gmj, 0), (SIDE_ABOVE, 0, gmi, 1), (SIDE_RIGHT, 0, gmj, 2),
                             (SIDE_BELOW, 0, gmi, 3)]
            self.n_boundary_obj = 4
        self.inner_boxes = [tuple(b) for b in inner_boxes]
        wall = wall_phi_V / self.F_scale
        self.object_phi = [wall] * self.n_boundary_obj + [p / self.F_scale for p in inner_phi_V]
        # initial loading extents (DISTRIBUTE_PARTICLES, pic2d_CurProblemValues.f90:3166-3189):
        # x over the N periodic cells (or the full width between walls), y between the walls
        self.x_limits = (0.0, float(gmi - 1)) if self.periodic_x else (0.0, float(gmi))
        self.y_limits = (0.0, float(gmj))
        n_limit_x = (gmi - 1) if self.periodic_x else gmi
        # N_electrons = INT(N_of_particles_cell * (init_Ne_m3 / N_plasma_m3) * n_limit_x * n_limit_y)
        # (pic2d_CurProblemValues.f90:3189)
        ratio = 1.0 if init_Ne_m3 is None else init_Ne_m3 / N_plasma_m3
        self.n_particles = int(N_of_particles_cell * ratio * n_limit_x * gmj)
        self.factor_convert_e = math.sqrt(self.init_Te_eV / T_e_eV) / N_max_vel
        self.factor_convert_i = math.sqrt(init_Ti_eV / T_e_eV) / (N_max_vel * math.sqrt(self.Ms))

    @property
    def nx(self):
        return self.gmi + 1

    @property
    def ny(self):
        return self.gmj + 1

    def geometry(self):
        """Geometry as a plain dict: extents, periodicity, wall segments, object potentials, inner boxes."""
        return dict(gmi=self.gmi, gmj=self.gmj, periodic_x=self.periodic_x, segments=self.segments,
                    object_phi=self.object_phi, n_boundary_obj=self.n_boundary_obj, inner_boxes=self.inner_boxes,
                    vacuum_gaps=list(getattr(self, "vacuum_gaps", [])))

    # ---- collisions -------------------------------------------------------------------
    def cross_sections(self):
        """Synthetic analytic argon cross sections tabulated in the reference's
        init_neutral_*_crsect_*.dat quantisation (f9.3 energy, e10.3 sigma)."""
        return synthetic_argon_cross_sections()

    def ion_velocity_factor(self):
        # pic2d_enCollisionsGeneralProc.f90:221
        return math.sqrt(self.neutral_T_K * KB_JK / (self.T_e_eV * E_CL)) / (self.N_max_vel * math.sqrt(self.Ms))

    def rcx_sigma(self):
        return 5.0e-19, 0.0


def _q_energy(e):
    return float("%.3f" % e)


def _q_sigma(s):
    return float("%.3e" % s)


def synthetic_argon_cross_sections():
    """Elastic 1e-19 m^2 constant; excitation threshold 11.5 eV peaking at 2e-21 m^2 at 30 eV;
    ionisation threshold 15.76 eV with a Lotz-type ln(E/I)/E^a shape peaking at 3e-20 m^2 at
    100 eV.  Returned as the list of processes the table builder takes."""
    E_el = np.concatenate([[0.0], np.geomspace(0.01, 10000.0, 60)])
    el = [(_q_energy(e), _q_sigma(1.0e-19)) for e in E_el]

    def excitation(E, Eth=11.5, Ep=30.0, sp=2.0e-21):
        u = (E - Eth) / (Ep - Eth)
        return sp * u * math.exp(1.0 - u)

    E_ex = np.concatenate([[11.5], 11.5 + np.geomspace(0.05, 10000.0 - 11.5, 50)])
    ex = [(_q_energy(e), _q_sigma(excitation(e))) for e in E_ex]
    I = 15.76
    a = 1.0 / math.log(100.0 / I)

    def lotz(E):
        g = math.log(E / I) / E ** a
        g0 = math.log(100.0 / I) / 100.0 ** a
        return 3.0e-20 * g / g0

    E_io = np.concatenate([[15.76], 15.76 + np.geomspace(0.05, 10000.0 - 15.76, 60)])
    io = [(_q_energy(e), _q_sigma(lotz(e))) for e in E_io]
    return [dict(type=10, E=[p[0] for p in el], sigma=[p[1] for p in el]),
            dict(type=20, E=[p[0] for p in ex], sigma=[p[1] for p in ex]),
            dict(type=30, E=[p[0] for p in io], sigma=[p[1] for p in io], ion_species=0)]


# ---- the workloads ----------------------------------------------------------------------------

def workload(n, scale=1):
    """Problem of configuration n (1..5).  scale>1 divides the grid by `scale` in each
    direction (the same plasma, for tests that must finish quickly)."""
    if n in (1, 2, 3):
        if n == 1:
            N, gmj, ppc, Bz = 256, 256, 50, 0.0
        else:
            N, gmj, ppc, Bz = 2048, 1024, 200, 0.01
        N //= scale
        gmj //= scale
        return Problem(T_e_eV=2.0, N_plasma_m3=1.0e15, N_of_cells_debye=2, N_max_vel=5, N_of_particles_cell=ppc,
                       gmi=N + 1, gmj=gmj, periodic_x=True, init_Ti_eV=0.026, Bz_T=Bz, name="cfg%d" % n)
    if n == 4:
        G = 1024 // scale
        bw, bh = 128 // scale, 64 // scale
        jc = G // 2
        boxes = [(G // 4, G // 4 + bw, jc - bh // 2, jc + bh // 2),
                 (3 * G // 4 - bw, 3 * G // 4, jc - bh // 2, jc + bh // 2)]
        return Problem(T_e_eV=3.0, N_plasma_m3=5.0e15, N_of_cells_debye=2, N_max_vel=5, N_of_particles_cell=100,
                       gmi=G, gmj=G, periodic_x=False, init_Ti_eV=0.026, inner_boxes=boxes,
                       inner_phi_V=(50.0, -50.0), name="cfg4")
    if n == 5:
        N, gmj = 4096 // scale, 2048 // scale
        bw, bh = 256 // scale, 128 // scale
        box = (N // 2 - bw // 2, N // 2 + bw // 2, gmj // 2 - bh // 2, gmj // 2 + bh // 2)
        p = Problem(T_e_eV=2.0, N_plasma_m3=1.0e15, N_of_cells_debye=2, N_max_vel=5, N_of_particles_cell=100,
                    gmi=N + 1, gmj=gmj, periodic_x=True, init_Ti_eV=0.026, inner_boxes=[box],
                    inner_phi_V=(0.0,), name="cfg5")
        # y walls (objects 1 above, 3 below) with a constant true-secondary yield of 0.5 (model 1 over
        # all energies up to 1 keV, secondaries at 2 eV); a 100 eV beam from the wall below at
        # 1e19 m^-2 s^-1: macroparticles per step = flux dt (N dx) / (n dx^2 / ppc)
        see = dict(model_true=1, setD_true=0.5, minE_true_eV=0.0, maxE_true_eV=1000.0, T_true_eV=2.0)
        n_beam = 1.0e19 * p.dt * N * p.N_of_particles_cell / (p.N_plasma_m3 * p.dx)
        p.wall_emission = {"see": {1: see, 3: see},
                           "constant": {3: dict(n_per_step=n_beam, model=1, Te_normal_eV=1.0, Te_parallel_eV=0.1,
                                                We_beam_eV=100.0)}}
        return p
    raise ValueError("no workload %r" % n)
