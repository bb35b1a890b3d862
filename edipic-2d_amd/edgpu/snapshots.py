"""Snapshots in the reference's binary format, from the device state.

SAVE_GLOBAL_2D_ARRAY (pic2d_Snapshots.f90:1207-1357) writes one FP32 matrix of (gmj + 2)
rows by (gmi + 2) columns, row-major, native byte order: the first row is [gmi + 1,
x_0 .. x_gmi] (node x-coordinates in metres), every following row j = 0 .. gmj is
[y_j, a(0, j) .. a(gmi, j)] -- the layout gnuplot's "binary matrix" and the reference's
Postprocessing/plot_snapshots.py read.

The moments are the device's raw accumulators (edgpu_collect_moments), normalised as
COLLECT_ELECTRON_MOMENTS / COLLECT_ION_MOMENTS do (pic2d_ElectronMoments.f90:350-419,
pic2d_IonMoments.f90:280-345) in single precision, with the wall correction
ADJUST_DENSITY_AT_WALL_BOUNDARIES (pic2d_ElectronMoments.f90:1829-1933).  A single
X-periodic cluster has no horizontal neighbours in SYNCHRONIZE_MOMENTS_IN_OVERLAP_NODES,
so its overlap columns are written as accumulated.  Unit conversions and file names follow
CREATE_SNAPSHOT (pic2d_Snapshots.f90:380-1000).
"""
import os

import numpy as np

from .config import E_CL, M_E_KG

f32 = np.float32
MOMENT_NAMES = ("N", "VX", "VY", "VZ", "WX", "WY", "WZ", "VXVY", "VXVZ", "VYVZ", "QX", "QY", "QZ")


def save_global_2d_array(arr, path, gmi, gmj, delta_x_m):
    """arr: (gmj+1, gmi+1) values a(i, j) at [j, i], written as FP32."""
    a = np.asarray(arr, np.float32).reshape(gmj + 1, gmi + 1)
    out = np.empty((gmj + 2, gmi + 2), np.float32)
    out[0, 0] = f32(gmi + 1)
    out[0, 1:] = (np.arange(gmi + 1, dtype=np.float64) * delta_x_m).astype(np.float32)
    out[1:, 0] = (np.arange(gmj + 1, dtype=np.float64) * delta_x_m).astype(np.float32)
    out[1:, 1:] = a
    out.tofile(path)


def read_global_2d_array(path, gmi, gmj):
    """(x, y, a[j, i]) of a file written by save_global_2d_array."""
    m = np.fromfile(path, np.float32).reshape(gmj + 2, gmi + 2)
    return m[0, 1:], m[1:, 0], m[1:, 1:]


def normalise_moments(raw, periodic_x):
    """Single-precision normalisation of the 13 raw accumulators (shape (13, ny, nx)):
    averages, twice the mixed products, heat flows; then the wall correction of N."""
    N, VX, VY, VZ, WX, WY, WZ, VXVY, VXVZ, VYVZ, QX, QY, QZ = [np.array(raw[q], np.float32) for q in range(13)]
    ok = N > f32(1.0e-9)
    inv = np.where(ok, f32(1.0) / np.where(ok, N, f32(1.0)), f32(0.0)).astype(np.float32)
    VX, VY, VZ = VX * inv, VY * inv, VZ * inv
    WX, WY, WZ = WX * inv, WY * inv, WZ * inv
    VXVY = f32(2.0) * VXVY * inv
    VXVZ = f32(2.0) * VXVZ * inv
    VYVZ = f32(2.0) * VYVZ * inv
    rtemp = VX * VX + VY * VY + VZ * VZ
    rtemp = rtemp + rtemp - WX - WY - WZ
    QX = QX * inv + (rtemp - WX - WX) * VX - VXVY * VY - VXVZ * VZ
    QY = QY * inv - VXVY * VX + (rtemp - WY - WY) * VY - VYVZ * VZ
    QZ = QZ * inv - VXVZ * VX - VYVZ * VY + (rtemp - WZ - WZ) * VZ
    m = dict(N=N, VX=VX, VY=VY, VZ=VZ, WX=WX, WY=WY, WZ=WZ, VXVY=VXVY, VXVZ=VXVZ, VYVZ=VYVZ, QX=QX, QY=QY,
             QZ=QZ)
    for k in m:
        m[k] = np.where(ok, m[k], f32(0.0)).astype(np.float32)
    # ADJUST_DENSITY_AT_WALL_BOUNDARIES: walls double the density of their nodes, corners of a
    # walled box (SURROUNDED_BY_WALL in the single cluster) take 4
    n = m["N"]
    if periodic_x:
        n[0, :] *= f32(2.0)
        n[-1, :] *= f32(2.0)
    else:
        n[1:-1, 0] *= f32(2.0)
        n[1:-1, -1] *= f32(2.0)
        n[-1, 1:-1] *= f32(2.0)
        n[0, 1:-1] *= f32(2.0)
        for (j, i) in ((0, 0), (-1, 0), (0, -1), (-1, -1)):
            n[j, i] *= f32(4.0)
    return m


def scales(prob):
    """The snapshot conversion factors of INITIATE_PARAMETERS (pic2d_CurProblemValues.f90:893-903)."""
    V = prob.V_scale
    n_scale = prob.N_plasma_m3 / prob.N_of_particles_cell
    return dict(N_scale_part_m3=n_scale, current_factor_Am2=E_CL * V * n_scale, V_scale_ms=V,
                energy_factor_eV=0.5 * M_E_KG * V ** 2 / E_CL, temperature_factor_eV=M_E_KG * V ** 2 / E_CL,
                heat_flow_factor_Wm2=0.5 * M_E_KG * V ** 2 * n_scale, F_scale_V=prob.F_scale, E_scale_Vm=prob.E_scale)


def _temperature(N, W, V, factor):
    t = np.maximum(f32(0.0), W - V * V) * f32(factor)
    return np.where(N > f32(0.0), t, f32(0.0)).astype(np.float32)


def species_files(m, snap, sc, species, Qs=1, Ms=1.0):
    """{file name: FP32 grid} of one species' moment snapshot (CREATE_SNAPSHOT :575-1000)."""
    out = {}
    N = m["N"]
    if species == 0:
        tag = lambda q, unit: "_%04d_%se_%s_2D.bin" % (snap, q, unit)
        jf, sign, ef, tf, hf = sc["current_factor_Am2"], -1, sc["energy_factor_eV"], sc["temperature_factor_eV"], \
            sc["heat_flow_factor_Wm2"]
        out[tag("N", "m3")] = N * f32(sc["N_scale_part_m3"])
    else:
        tag = lambda q, unit: "_%04d_%si_%d_%s_2D.bin" % (snap, q, species, unit)
        jf, sign = Qs * sc["current_factor_Am2"], 1
        ef, tf, hf = Ms * sc["energy_factor_eV"], Ms * sc["temperature_factor_eV"], Ms * sc["heat_flow_factor_Wm2"]
        out[tag("N", "m3")] = N * f32(sc["N_scale_part_m3"])
    for c in "XYZ":
        V = m["V" + c]
        out[tag("J" + c, "Am2")] = ((-N if sign < 0 else N) * V * f32(jf)).astype(np.float32)
        out[tag("V" + c, "ms")] = V * f32(sc["V_scale_ms"])
        out[tag("W" + c, "eV")] = m["W" + c] * f32(ef)
        out[tag("T" + c, "eV")] = _temperature(N, m["W" + c], V, tf)
        out[tag("Q" + c, "Wm2")] = N * m["Q" + c] * f32(hf)
    return out


def write_snapshot(rt, directory, snap, species=(0, 1)):
    """All snapshot files of step `snap`: F_V, EX_Vm, EY_Vm and the moments of `species`
    (the electron velocity half-advanced, as COLLECT_ELECTRON_MOMENTS does).  Returns the
    file names written."""
    prob = rt.prob
    sc = scales(prob)
    gmi, gmj, dx = prob.gmi, prob.gmj, prob.dx
    files = {"_%04d_F_V_2D.bin" % snap: (rt.get_grid("phi") * sc["F_scale_V"]).astype(np.float32),
             "_%04d_EX_Vm_2D.bin" % snap: (rt.get_grid("EX") * sc["E_scale_Vm"]).astype(np.float32),
             "_%04d_EY_Vm_2D.bin" % snap: (rt.get_grid("EY") * sc["E_scale_Vm"]).astype(np.float32)}
    for s in species:
        raw = rt.collect_moments(s).reshape(13, gmj + 1, gmi + 1)
        m = normalise_moments(raw, prob.periodic_x)
        files.update({k: v.ravel() for k, v in species_files(m, snap, sc, s, prob.Qs, prob.Ms).items()})
    for name, a in files.items():
        save_global_2d_array(a, os.path.join(directory, name), gmi, gmj, dx)
    return sorted(files)
