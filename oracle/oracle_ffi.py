"""ctypes binding of the CPU oracle (oracle.c) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
leg, always as the checker or the timed CPU baseline.  The GPU product path
never imports this module.

Also binds oracle/_ref/libwellref.so (the reference's own WELL19937a RNG path
compiled from its sources) and oracle/_ref/libpushref.so (the reference's own
ADVANCE_ELECTRONS, compiled from its sources and driven by ref_push_harness.f90)
when those libraries have been built.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libwellref.so")
REF_PUSH_PATH = os.path.join(HERE, "_ref", "libpushref.so")

MAX_SEG = 16
MAX_INNER = 8
MAX_PROC = 8
SIDE_LEFT, SIDE_ABOVE, SIDE_RIGHT, SIDE_BELOW = 0, 1, 2, 3

E_CL = 1.602189e-19
M_E_KG = 9.109534e-31
EPS0 = 8.854188e-12
AMU_KG = 1.660565e-27
KB_JK = 1.38064852e-23


class Well(C.Structure):
    _fields_ = [("s", C.c_uint32 * 624), ("i", C.c_int32)]


class Geom(C.Structure):
    _fields_ = [
        ("ixmin", C.c_int32), ("ixmax", C.c_int32), ("iymin", C.c_int32), ("iymax", C.c_int32),
        ("periodic_x", C.c_int32), ("L_period_x", C.c_double),
        ("n_seg", C.c_int32),
        ("seg_side", C.c_int32 * MAX_SEG), ("seg_start", C.c_int32 * MAX_SEG),
        ("seg_end", C.c_int32 * MAX_SEG), ("seg_obj", C.c_int32 * MAX_SEG),
        ("n_obj", C.c_int32), ("n_boundary_obj", C.c_int32),
        ("obj_phi", C.c_double * (MAX_SEG + MAX_INNER)),
        ("n_inner", C.c_int32),
        ("inner_box", (C.c_int32 * 4) * MAX_INNER),
        ("seg_gap", C.c_int32 * MAX_SEG), ("seg_gap_from", C.c_int32 * MAX_SEG), ("seg_gap_to", C.c_int32 * MAX_SEG),
    ]


class Species(C.Structure):
    _fields_ = [
        ("is_electron", C.c_int32), ("Qs", C.c_int32), ("QM2s", C.c_double), ("QM2sNsub", C.c_double),
        ("N_subcycles", C.c_int32), ("sense_B", C.c_int32), ("sense_Ez", C.c_int32),
    ]


class Mcc(C.Structure):
    _fields_ = [
        ("n_proc", C.c_int32), ("type", C.c_int32 * MAX_PROC), ("threshold_eV", C.c_double * MAX_PROC),
        ("ion_velocity_factor", C.c_double * MAX_PROC), ("ion_species", C.c_int32 * MAX_PROC),
        ("M_amu", C.c_double), ("n_seg", C.c_int32), ("seg_value", C.c_double * 33),
        ("seg_step", C.c_double * 33), ("seg_index", C.c_int32 * 33), ("n_values", C.c_int32),
        ("energy_eV", C.POINTER(C.c_double)), ("prob", C.POINTER(C.c_double)),
        ("max_colliding_fraction", C.c_double), ("energy_factor", C.c_double),
    ]


class See(C.Structure):
    _fields_ = [("model", C.c_int32 * 3), ("refl_type", C.c_int32), ("setD", C.c_double * 3),
                ("minE", C.c_double * 3), ("maxE", C.c_double * 3)] + [(n, C.c_double) for n in (
                    "E_elast_0", "E_elast_max", "maxD_elast", "dE_elast", "frac_elast_high", "frac_inelastic",
                    "E_see_0", "E_see_max", "maxD_classic", "k_smooth", "vinj_true", "lowest")]


def see_from_material(m, energy_factor_eV, T_e_eV, N_max_vel):
    """orc_see from init_material_*.dat values (eV, names of the library's edgpu_see_params)
    with the conversions of pic2d_Materials.f90:180-245."""
    o = See()
    models = (m.get("model_elastic", 0), m.get("model_inelastic", 0), m.get("model_true", 0))
    for q in range(3):
        o.model[q] = models[q]
    o.refl_type = m.get("elastic_reflection", 0)
    for q, key in enumerate(("elastic", "inelastic", "true")):
        o.setD[q] = m.get("setD_" + key, 0.0)
        o.minE[q] = m.get("minE_%s_eV" % key, 0.0) / energy_factor_eV
        o.maxE[q] = m.get("maxE_%s_eV" % key, 0.0) / energy_factor_eV
    o.E_elast_0 = m.get("E_elast_0_eV", 0.0) / energy_factor_eV
    o.E_elast_max = m.get("E_elast_max_eV", 0.0) / energy_factor_eV
    o.maxD_elast = m.get("maxD_elast", 0.0)
    o.dE_elast = m.get("dE_elast_eV", 0.0) / energy_factor_eV
    o.frac_elast_high = 0.0 if models[0] == 0 else m.get("frac_elast_highenergy", 0.0)
    o.frac_inelastic = 0.0 if models[1] == 0 else m.get("frac_inelastic", 0.0)
    o.E_see_0 = m.get("E_see_0_eV", 0.0) / energy_factor_eV
    o.E_see_max = m.get("E_see_max_eV", 0.0) / energy_factor_eV
    o.maxD_classic = m.get("maxD_see_classic", 0.0)
    o.k_smooth = m.get("k_smooth", 0.0)
    o.vinj_true = math.sqrt(m.get("T_true_eV", 0.0) / T_e_eV) / N_max_vel
    low = 1.0e6 / energy_factor_eV
    for q, mod in enumerate(models):
        if mod == 1:
            low = min(low, o.minE[q])
        elif mod == 2:
            low = min(low, o.E_elast_0 if q == 0 else o.E_see_0)
    o.lowest = low
    return o


def emission_factors(model, Te_normal_eV, Te_parallel_eV, We_beam_eV, T_e_eV, N_max_vel):
    """(f_normal, f_parallel, v_beam) of PREPARE_SETUP_VALUES, pic2d_Setup.f90:102-126."""
    if model == 0:
        fn = math.sqrt(Te_normal_eV / T_e_eV) / N_max_vel
        vb = 0.0
    else:
        fn = math.sqrt(0.25 * Te_normal_eV ** 2 / (We_beam_eV * T_e_eV)) / N_max_vel
        vb = math.sqrt(We_beam_eV / T_e_eV) / N_max_vel
    return fn, math.sqrt(Te_parallel_eV / T_e_eV) / N_max_vel, vb


def see_emit(o, maxw, maxw_inj, dirflag, xmax, ymax, x, y, vx, vy, vz, u, max_out=64):
    out = np.zeros(5 * max_out)
    n = C.c_int32()
    u = np.ascontiguousarray(u, np.float64)
    rc = lib().orc_see_emit(C.byref(o), dp(maxw), dp(maxw_inj), dirflag, xmax, ymax, x, y, vx, vy, vz, dp(u), u.size,
                            max_out, dp(out), C.byref(n))
    return rc, out[:5 * min(n.value, max_out)].reshape(-1, 5)


def iiee_emit(setD, minE, maxE, vinj, Ms, maxw, maxw_inj, dirflag, xmax, ymax, x, y, vx, vy, vz, u, max_out=64):
    out = np.zeros(5 * max_out)
    n = C.c_int32()
    u = np.ascontiguousarray(u, np.float64)
    rc = lib().orc_iiee_emit(setD, minE, maxE, vinj, Ms, dp(maxw), dp(maxw_inj), dirflag, xmax, ymax, x, y, vx, vy, vz,
                             dp(u), u.size, max_out, dp(out), C.byref(n))
    return rc, out[:5 * min(n.value, max_out)].reshape(-1, 5)


def ion_reflect(dirflag, xmax, ymax, x, y, vx, vy, vz):
    out = np.zeros(5)
    lib().orc_ion_reflect(dirflag, xmax, ymax, x, y, vx, vy, vz, dp(out))
    return out


def see_emit_inner(o, maxw, maxw_inj, dirflag, box, x, y, vx, vy, vz, u, max_out=64):
    out = np.zeros(5 * max_out)
    n = C.c_int32()
    u = np.ascontiguousarray(u, np.float64)
    b = (C.c_int32 * 4)(*box)
    rc = lib().orc_see_emit_inner(C.byref(o), dp(maxw), dp(maxw_inj), dirflag, b, x, y, vx, vy, vz, dp(u), u.size,
                                  max_out, dp(out), C.byref(n))
    return rc, out[:5 * min(n.value, max_out)].reshape(-1, 5)


def inner_crossing(g, xorg, yorg, x, y):
    """(status, object, dirflag, xc, yc) of orc_inner_crossing"""
    o, d = C.c_int32(), C.c_int32()
    xc, yc = C.c_double(), C.c_double()
    rc = lib().orc_inner_crossing(C.byref(g), xorg, yorg, x, y, C.byref(o), C.byref(d), C.byref(xc), C.byref(yc))
    return rc, o.value, d.value, xc.value, yc.value


def constant_emission(dirflag, periodic, gmi, gmj, istart, iend, L, N_emit, model, fn, fp, vb, maxw, maxw_inj, u,
                      max_out=100000):
    out = np.zeros(5 * max_out)
    n = C.c_int64()
    u = np.ascontiguousarray(u, np.float64)
    rc = lib().orc_constant_emission(dirflag, periodic, gmi, gmj, istart, iend, L, N_emit, model, fn, fp, vb,
                                     dp(maxw), dp(maxw_inj), dp(u), u.size, max_out, dp(out), C.byref(n))
    assert rc == 0
    return out[:5 * n.value].reshape(-1, 5)


class Rcx(C.Structure):
    _fields_ = [
        ("sigma_1eV", C.c_double), ("alpha", C.c_double), ("N_m3", C.c_double), ("dt_s", C.c_double),
        ("N_subcycles", C.c_int32), ("Ms", C.c_double), ("energy_factor", C.c_double),
        ("V_scale", C.c_double), ("vfactor", C.c_double), ("probab_thermal", C.c_double),
    ]


def build():
    """Compile liboracle.so (and the reference RNG library when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def ref_lib():
    """The reference's own RNG path, or None when it was not built (e.g. on a machine without the reference)."""
    if not os.path.exists(REF_PATH):
        return None
    r = C.CDLL(REF_PATH)
    r.ref_well_seed.argtypes = [C.c_int32]
    r.ref_well_fill.argtypes = [C.c_int64, C.POINTER(C.c_double)]
    return r


_ref_push = None


def ref_push_lib():
    """The reference's own electron push (ADVANCE_ELECTRONS through ref_push_harness.f90), or
    None when it was not built."""
    global _ref_push
    if _ref_push is None:
        if not os.path.exists(REF_PUSH_PATH):
            return None
        try:
            r = C.CDLL(REF_PUSH_PATH)
        except OSError:  # a host without the MPICH runtime the reference links against
            return None
        Dn = np.ctypeslib.ndpointer(np.float64, flags="C")
        In = np.ctypeslib.ndpointer(np.int32, flags="C")
        Fn = np.ctypeslib.ndpointer(np.float32, flags="C")
        r.refh_setup.argtypes = [C.c_int32] * 4 + [In] * 4 + [C.c_int32, C.c_int32, In] + [C.c_double] * 4
        r.refh_advance_electrons.argtypes = ([C.c_int64] + [Dn] * 7 + [Dn] * 5 +
                                             [In, In, C.c_int64, C.POINTER(C.c_int64), In, In, Fn])
        r.refh_advance_electrons.restype = C.c_int64
        r.refh_setup_ions.argtypes = [C.c_int32, In, Dn, C.c_int32, C.c_int32, C.c_int32]
        r.refh_advance_ions.argtypes = ([C.c_int32, C.c_int64] + [Dn] * 7 + [Dn] * 5 +
                                        [In, In, C.c_int64, C.POINTER(C.c_int64), In, In, Fn])
        r.refh_advance_ions.restype = C.c_int64
        _ref_push = r
    return _ref_push


def _ref_setup(r, cfg, B, Ez):
    segs = np.array(cfg["segments"], np.int32).reshape(-1, 4)
    boxes = cfg.get("inner_boxes", [])
    bx = np.ascontiguousarray(np.array(boxes, np.int32).reshape(-1, 4) if boxes else np.zeros((1, 4), np.int32))
    r.refh_setup(cfg["gmi"], cfg["gmj"], 1 if cfg["periodic_x"] else 0, len(segs),
                 *[np.ascontiguousarray(segs[:, q]) for q in range(4)], cfg["n_boundary_obj"], len(boxes), bx,
                 float(B[0]), float(B[1]), float(B[2]), float(Ez))
    return cfg["n_boundary_obj"] + len(boxes)


def _ref_run(fn, lead, nobj, X, Y, VX, VY, VZ, EX, EY):
    n = len(X)
    out = [np.empty(max(n, 1)) for _ in range(5)]
    tag = np.empty(max(n, 1), np.int32)
    hits = np.zeros(nobj, np.int32)
    nrec = C.c_int64()
    ro, rt, rc = np.empty(max(n, 1), np.int32), np.empty(max(n, 1), np.int32), np.empty(max(n, 1), np.float32)
    m = fn(*lead, n, *[np.ascontiguousarray(a, np.float64) for a in (X, Y, VX, VY, VZ, EX, EY)], *out, tag, hits, n,
           C.byref(nrec), ro, rt, rc)
    k = nrec.value
    return [a[:m] for a in out], tag[:m], hits, (ro[:k], rt[:k], rc[:k])


def ref_advance_ions(cfg, B, Ez, species, s, N_subcycles, sense_B, sense_Ez, X, Y, VX, VY, VZ, accEX, accEY):
    """One ADVANCE_IONS of the reference itself: species = [(Qs, QM2s), ...], ions of species s
    (1-based, tags 1..n) pushed in the accumulated field; returns as ref_advance_electrons
    (records carry token = tag + 100 * segment + 10000 * s)."""
    r = ref_push_lib()
    nobj = _ref_setup(r, cfg, B, Ez)
    qs = np.array([q for q, _ in species], np.int32)
    qm = np.array([m for _, m in species], np.float64)
    r.refh_setup_ions(len(species), qs, qm, N_subcycles, 1 if sense_B else 0, 1 if sense_Ez else 0)
    return _ref_run(r.refh_advance_ions, (s,), nobj, X, Y, VX, VY, VZ, accEX, accEY)


def ref_advance_electrons(cfg, B, Ez, X, Y, VX, VY, VZ, EX, EY):
    """One ADVANCE_ELECTRONS of the reference itself over the electrons given (tags 1..n) in
    the geometry dict cfg (see edgpu.config) with uniform B and Ez.  Returns (electrons left,
    their tags, hit count per object, wall-hit records (object 1-based, token = tag + 100 *
    segment, REAL coordinate along the face))."""
    r = ref_push_lib()
    nobj = _ref_setup(r, cfg, B, Ez)
    return _ref_run(r.refh_advance_electrons, (), nobj, X, Y, VX, VY, VZ, EX, EY)


D = C.POINTER(C.c_double)
I32 = C.POINTER(C.c_int32)
I64 = C.POINTER(C.c_int64)
U8 = C.POINTER(C.c_uint8)


def _declare(L):
    L.orc_well_seed.argtypes = [C.POINTER(Well), C.c_int32]
    L.orc_well_next.argtypes = [C.POINTER(Well)]
    L.orc_well_next.restype = C.c_double
    L.orc_well_fill.argtypes = [C.POINTER(Well), C.c_int64, D]
    L.orc_maxwell_tables.argtypes = [D, D]
    L.orc_maxwell_sample.argtypes = [D, C.c_double]
    L.orc_maxwell_sample.restype = C.c_double
    L.orc_maxwell_inj_sample.argtypes = [D, C.c_double]
    L.orc_maxwell_inj_sample.restype = C.c_double
    L.orc_advance.argtypes = [C.POINTER(Geom), C.POINTER(Species), D, D, C.c_double, C.c_double, C.c_double,
                              C.c_double, C.c_int64, D, D, D, D, D, U8, I32, I64]
    L.orc_advance_ex.argtypes = [C.POINTER(Geom), C.POINTER(Species), D, D, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_int64, D, D, D, D, D, U8, I32, I64, D, D, I32]
    L.orc_inner_crossing.argtypes = [C.POINTER(Geom), C.c_double, C.c_double, C.c_double, C.c_double, I32, I32, D,
                                     D]
    L.orc_see_emit_inner.argtypes = [C.POINTER(See), D, D, C.c_int32, C.POINTER(C.c_int32), C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_double, D, C.c_int32, C.c_int32, D, I32]
    L.orc_deposit.argtypes = [C.POINTER(Geom), C.c_int32, C.c_int64, D, D, U8, D]
    L.orc_fold_walls.argtypes = [C.POINTER(Geom), D]
    L.orc_real8ft.argtypes = [D, C.c_int32, C.c_int32]
    L.orc_prepare_sys_y.argtypes = [C.c_int32, C.c_int32, D]
    L.orc_solve_poisson_fftx.argtypes = [C.POINTER(Geom), C.c_int32, D, D, D, D]
    L.orc_prepare_sys_y_walls.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, D]
    L.orc_solve_poisson_fftx_walls.argtypes = [C.POINTER(Geom), C.c_int32, D, D, D, C.c_uint32, D, D, C.c_int32, D]
    L.orc_surface_charge.argtypes = [C.POINTER(Geom), C.c_uint32, C.c_int64, I32, I32, D, I32, C.c_int32, D, D]
    L.orc_surface_charge_fold.argtypes = [C.POINTER(Geom), D]
    L.orc_ion_reflect.argtypes = [C.c_int32] + [C.c_double] * 7 + [D]
    L.orc_iiee_emit.argtypes = [C.c_double] * 5 + [D, D, C.c_int32] + [C.c_double] * 7 + [D, C.c_int32, C.c_int32, D,
                                                                                          I32]
    L.orc_efield.argtypes = [C.POINTER(Geom), D, C.c_double, C.c_double, D, D, D, D]
    L.orc_krylov_rhs.argtypes = [C.POINTER(Geom), C.c_int32, D, D, D]
    L.orc_krylov_apply.argtypes = [C.POINTER(Geom), D, D]
    L.orc_krylov_cg.argtypes = [C.POINTER(Geom), D, D, C.c_double, C.c_int32, D]
    L.orc_en_frequency.argtypes = [C.c_double, C.c_int32, C.c_int32, D, D, C.c_double, C.c_int32, C.c_double]
    L.orc_en_frequency.restype = C.c_double
    L.orc_mcc_table_size.argtypes = [C.c_int32, D, D]
    L.orc_mcc_build.argtypes = [C.POINTER(Mcc), C.c_int32, I32, I32, D, D, C.c_double, C.c_int32, C.c_double]
    L.orc_mcc_select.argtypes = [C.POINTER(Mcc), C.c_double, C.c_double, I32, D, D, I32]
    L.orc_en_scatter.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_double, D, C.c_double, D, D, D, D, I32, D,
                                 D]
    L.orc_mcc_en_stream.argtypes = [C.POINTER(Mcc), D, C.c_int64, D, D, D, D, D, D, I32, I64, C.c_int64, I64, D, D]
    L.orc_mcc_en_reference.argtypes = [C.POINTER(Mcc), D, C.POINTER(Well), C.c_int64, D, D, D, D, D, I32, I64,
                                       C.c_int64, I64, D, D]
    L.orc_rcx_setup.argtypes = [C.POINTER(Rcx), C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32]
    L.orc_see_coeff.argtypes = [C.POINTER(See), C.c_int32, C.c_double, C.c_double]
    L.orc_see_coeff.restype = C.c_double
    L.orc_see_emit.argtypes = [C.POINTER(See), D, D, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double,
                               C.c_double, C.c_double, C.c_double, D, C.c_int32, C.c_int32, D, I32]
    L.orc_constant_emission.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                        C.c_float, C.c_int32, C.c_double, C.c_double, C.c_double, D, D, D, C.c_int64,
                                        C.c_int64, D, C.POINTER(C.c_int64)]
    L.orc_rcx_stream.argtypes = [C.POINTER(Rcx), D, C.c_int64, D, D, D, D, U8]
    L.orc_rcx_stream.restype = C.c_int64
    L.orc_rcx_reference.argtypes = [C.POINTER(Rcx), D, C.POINTER(Well), C.c_int64, D, D, D]
    L.orc_rcx_reference.restype = C.c_int64
    L.orc_moments_raw.argtypes = [C.POINTER(Geom), C.c_int64, D, D, D, D, D, C.POINTER(C.c_float)]
    L.orc_distribute.argtypes = [C.POINTER(Well), D, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, D, D, D, D, D]
    L.orc_distribute_ions.argtypes = [C.POINTER(Well), D, C.c_int64, D, D, C.c_double, D, D, D, D, D]
    L.orc_remove_inside_inner.argtypes = [C.POINTER(Geom), C.c_int64, D, D, D, D, D]
    L.orc_remove_inside_inner.restype = C.c_int64


def dp(a):
    return a.ctypes.data_as(D)


def ip(a):
    return a.ctypes.data_as(I32)


def lp(a):
    return a.ctypes.data_as(I64)


def up(a):
    return a.ctypes.data_as(U8)


# ---------------------------------------------------------------------------
# thin Python helpers


def well(seed):
    w = Well()
    lib().orc_well_seed(C.byref(w), int(seed))
    return w


def well_fill(w, n):
    out = np.empty(n, np.float64)
    lib().orc_well_fill(C.byref(w), n, dp(out))
    return out


def maxwell_tables():
    v = np.zeros(9001, np.float64)
    vi = np.zeros(4501, np.float64)
    rc = lib().orc_maxwell_tables(dp(v), dp(vi))
    if rc != 0:
        raise RuntimeError("PrepareMaxwellDistribIntegral failed")
    return v, vi


def make_geom(cfg):
    """Build a Geom from a geometry dict (see edgpu.config for the keys)."""
    g = Geom()
    g.ixmin, g.ixmax, g.iymin, g.iymax = 0, cfg["gmi"], 0, cfg["gmj"]
    g.periodic_x = 1 if cfg["periodic_x"] else 0
    g.L_period_x = float(cfg["gmi"] - 1)
    segs = cfg["segments"]
    g.n_seg = len(segs)
    for k, (side, s, e, obj) in enumerate(segs):
        g.seg_side[k], g.seg_start[k], g.seg_end[k], g.seg_obj[k] = side, s, e, obj
    phis = cfg["object_phi"]
    g.n_obj = len(phis)
    g.n_boundary_obj = cfg["n_boundary_obj"]
    for k, p in enumerate(phis):
        g.obj_phi[k] = p
    boxes = cfg.get("inner_boxes", [])
    g.n_inner = len(boxes)
    for k, b in enumerate(boxes):
        for q in range(4):
            g.inner_box[k][q] = b[q]
    for n in cfg.get("vacuum_gaps", []):
        k = [q for q, sg in enumerate(segs) if sg[3] == n]
        if len(k) != 1:
            raise ValueError("a vacuum gap has one segment")
        k = k[0]
        g.seg_gap[k] = 1
        g.seg_gap_from[k] = _gap_neighbour(cfg, k, 1)
        g.seg_gap_to[k] = _gap_neighbour(cfg, k, 2)
    return g


def _seg_points(cfg, seg):
    """(i, j) of a segment's start and end node"""
    side, s, e, _ = seg
    if side == SIDE_BELOW:
        return (s, 0), (e, 0)
    if side == SIDE_ABOVE:
        return (s, cfg["gmj"]), (e, cfg["gmj"])
    if side == SIDE_LEFT:
        return (0, s), (0, e)
    return (cfg["gmi"], s), (cfg["gmi"], e)


def _gap_neighbour(cfg, k, which):
    """n_connected_to_start (which = 1) / _end (2) of the vacuum gap of segment k: the first
    other boundary object, in object order, one of whose segments (in their order) starts or
    ends at that end of the gap (pic2d_CurProblemValues.f90:350-380)"""
    segs = cfg["segments"]
    point = _seg_points(cfg, segs[k])[which - 1]
    for nn in range(cfg["n_boundary_obj"]):
        if nn == segs[k][3]:
            continue
        for sg in segs:
            if sg[3] == nn and point in _seg_points(cfg, sg):
                return nn
    raise ValueError("an end of vacuum gap %d is not connected to any other boundary object" % segs[k][3])


def make_species(is_electron, Qs=1, QM2s=0.0, N_subcycles=1, sense_B=1, sense_Ez=1):
    s = Species()
    s.is_electron = 1 if is_electron else 0
    s.Qs = Qs
    s.QM2s = QM2s
    s.QM2sNsub = QM2s * N_subcycles
    s.N_subcycles = N_subcycles
    s.sense_B = sense_B
    s.sense_Ez = sense_Ez
    return s


def advance(g, sp, EX, EY, B, Ez, X, Y, VX, VY, VZ, n_obj):
    n = X.size
    removed = np.zeros(n, np.uint8)
    hit_obj = np.zeros(n, np.int32)
    hits = np.zeros(max(n_obj, 1), np.int64)
    rc = lib().orc_advance(C.byref(g), C.byref(sp), dp(EX), dp(EY), B[0], B[1], B[2], Ez, n, dp(X), dp(Y), dp(VX),
                           dp(VY), dp(VZ), up(removed), ip(hit_obj), lp(hits))
    return rc, removed, hit_obj, hits


def advance_records(g, sp, EX, EY, B, Ez, X, Y, VX, VY, VZ, n_obj):
    """advance() plus, per removed particle, the recorded position and the dirflag of the wall
    or object face crossed (orc_advance_ex)"""
    n = X.size
    removed = np.zeros(n, np.uint8)
    hit_obj = np.zeros(n, np.int32)
    hits = np.zeros(max(n_obj, 1), np.int64)
    rx, ry = np.zeros(n), np.zeros(n)
    rd = np.zeros(n, np.int32)
    rc = lib().orc_advance_ex(C.byref(g), C.byref(sp), dp(EX), dp(EY), B[0], B[1], B[2], Ez, n, dp(X), dp(Y), dp(VX),
                              dp(VY), dp(VZ), up(removed), ip(hit_obj), lp(hits), dp(rx), dp(ry), ip(rd))
    return rc, removed, hit_obj, hits, rx, ry, rd


def remove_inside_inner(g, X, Y, VX, VY, VZ):
    """orc_remove_inside_inner on copies: the electrons left, in the reference's order"""
    a = [np.array(v, np.float64) for v in (X, Y, VX, VY, VZ)]
    n = lib().orc_remove_inside_inner(C.byref(g), a[0].size, *[dp(v) for v in a])
    return [v[:n].copy() for v in a]


def deposit(g, Qs, X, Y, skip=None):
    nx, ny = g.ixmax - g.ixmin + 1, g.iymax - g.iymin + 1
    rho = np.zeros(nx * ny, np.float64)
    rc = lib().orc_deposit(C.byref(g), Qs, X.size, dp(X), dp(Y), up(skip) if skip is not None else None, dp(rho))
    if rc != 0:
        raise RuntimeError("deposit failed %d" % rc)
    return rho


def fold_walls(g, rho):
    lib().orc_fold_walls(C.byref(g), dp(rho))
    return rho


def prepare_sys_y(N, gmj):
    a = np.zeros((gmj + 1) * (N // 2 + 1), np.float64)
    lib().orc_prepare_sys_y(N, gmj, dp(a))
    return a


def solve_poisson_fftx(g, N_ppc, rho_e, rho_i, a_eq):
    phi = np.zeros_like(rho_e)
    rc = lib().orc_solve_poisson_fftx(C.byref(g), N_ppc, dp(rho_e), dp(rho_i), dp(a_eq), dp(phi))
    if rc != 0:
        raise RuntimeError("poisson failed %d" % rc)
    return phi


def prepare_sys_y_walls(N, gmj, diel_below, diel_above, eps_below=1.0, eps_above=1.0):
    a = np.zeros((gmj + 1) * (N // 2 + 1), np.float64)
    lib().orc_prepare_sys_y_walls(N, gmj, int(diel_below), int(diel_above), eps_below, eps_above, dp(a))
    return a


def solve_poisson_fftx_walls(g, N_ppc, rho_e, rho_i, a_eq, diel_mask, sigma_below, sigma_above, two_dielectric):
    phi = np.zeros_like(rho_e)
    sb = np.ascontiguousarray(sigma_below, np.float64)
    sa = np.ascontiguousarray(sigma_above, np.float64)
    rc = lib().orc_solve_poisson_fftx_walls(C.byref(g), N_ppc, dp(rho_e), dp(rho_i), dp(a_eq), int(diel_mask), dp(sb),
                                            dp(sa), int(two_dielectric), dp(phi))
    if rc != 0:
        raise RuntimeError("poisson failed %d" % rc)
    return phi


def surface_charge(g, diel_mask, obj, dirflag, x, q, sigma_below, sigma_above, count=None):
    """adds the records' charge to sigma_below / sigma_above in place (no periodic fold)"""
    obj = np.ascontiguousarray(obj, np.int32)
    dirflag = np.ascontiguousarray(dirflag, np.int32)
    x = np.ascontiguousarray(x, np.float64)
    cnt = None if count is None else np.ascontiguousarray(count, np.int32)
    lib().orc_surface_charge(C.byref(g), int(diel_mask), obj.size, ip(obj), ip(dirflag), dp(x),
                             None if cnt is None else ip(cnt), int(q), dp(sigma_below), dp(sigma_above))


def surface_charge_fold(g, sigma):
    lib().orc_surface_charge_fold(C.byref(g), dp(sigma))


def efield(g, phi, f1, f2, acc=None):
    EX = np.zeros_like(phi)
    EY = np.zeros_like(phi)
    if acc is None:
        lib().orc_efield(C.byref(g), dp(phi), f1, f2, dp(EX), dp(EY), None, None)
    else:
        lib().orc_efield(C.byref(g), dp(phi), f1, f2, dp(EX), dp(EY), dp(acc[0]), dp(acc[1]))
    return EX, EY


class MccTable:
    """Holds an orc_mcc plus the numpy buffers it points into."""

    def __init__(self, seg_values, seg_steps, procs, M_amu, N_m3, N_subcycles, dt_s, V_scale):
        self.t = Mcc()
        t = self.t
        t.n_seg = len(seg_steps)
        t.seg_value[0] = seg_values[0]
        for i in range(1, t.n_seg + 1):
            t.seg_value[i] = seg_values[i]
            t.seg_step[i] = seg_steps[i - 1]
        sv = np.array([t.seg_value[i] for i in range(t.n_seg + 1)])
        ss = np.array([0.0] + [t.seg_step[i] for i in range(1, t.n_seg + 1)])
        nv = lib().orc_mcc_table_size(t.n_seg, dp(sv), dp(ss))
        self.energy = np.zeros(nv + 1, np.float64)
        self.prob = np.zeros((nv + 1) * len(procs), np.float64)
        t.energy_eV = dp(self.energy)
        t.prob = dp(self.prob)
        t.M_amu = M_amu
        t.energy_factor = V_scale * V_scale
        types = np.array([p["type"] for p in procs], np.int32)
        npts = np.array([len(p["E"]) for p in procs], np.int32)
        E = np.concatenate([np.asarray(p["E"], np.float64) for p in procs])
        S = np.concatenate([np.asarray(p["sigma"], np.float64) for p in procs])
        for k, p in enumerate(procs):
            t.ion_velocity_factor[k] = p.get("ion_velocity_factor", 0.0)
            t.ion_species[k] = p.get("ion_species", 0)
        rc = lib().orc_mcc_build(C.byref(t), len(procs), ip(types), ip(npts), dp(E), dp(S), N_m3, N_subcycles, dt_s)
        if rc != 0:
            raise RuntimeError("mcc build failed")
        self.n_values = t.n_values
        self.prob2d = self.prob.reshape(nv + 1, len(procs))
