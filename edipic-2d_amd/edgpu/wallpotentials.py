"""Time-dependent electrode potentials: UPDATE_WALL_POTENTIALS (pic2d_WallPotentials.f90:1-140)
for the host, which hands the result to edgpu_update_wall_potentials before each step's solve
(the reference calls it after the charge gathers, pic2d_MainProgram.f90:164).

A metal object's potential at step T_cntr is

    phi_const + (harmonic(T_cntr) + waveform(T_cntr)) * amplitude_profile(T_cntr)

with the harmonic phi_var sin(omega T_cntr + phase) (its phase re-aligned at the start of each
amplitude-profile period so that pulses start in the same phase), the waveform interpolated
linearly in a periodic table of (time step, potential) read from init_bo_NN_waveform.dat, and
the amplitude factor interpolated in a periodic table read from init_bo_NN_amplitude_profile.dat
(pic2d_Setup.f90:150-400).  The tables keep the reference's single-precision reads and
interpolation weight."""
import math
import os

import numpy as np


class Electrode:
    def __init__(self, phi_const=0.0, phi_var=0.0, omega=0.0, phase=0.0):
        self.phi_const, self.phi_var, self.omega, self.phase = phi_const, phi_var, omega, phase
        self.wf_T = self.wf_phi = None
        self.ap_T = self.ap_factor = None
        self.phase_adjusted = phase

    @property
    def varies(self):
        return self.phi_var != 0.0 or self.wf_T is not None or self.ap_T is not None

    def harmonic(self, T):
        if self.ap_T is not None:
            period = self.ap_T[-1]
            if T == (T // period) * period:
                n_periods = int((self.omega * T) / (2.0 * math.pi))
                self.phase_adjusted = n_periods * 2.0 * math.pi + self.phase - self.omega * T
            return self.phi_var * math.sin(self.omega * T + self.phase_adjusted)
        return self.phi_var * math.sin(self.omega * T + self.phase)

    @staticmethod
    def _interp(Ts, vals, T):
        period = Ts[-1]
        t = T - (T // period) * period
        for i in range(len(Ts) - 1):
            if Ts[i] <= t <= Ts[i + 1]:
                ai = np.float32((Ts[i + 1] - t) / (Ts[i + 1] - Ts[i]))
                return float(vals[i] * float(ai) + vals[i + 1] * (1.0 - float(ai)))
        return None

    def waveform(self, T):
        if self.wf_T is None:
            return 0.0
        v = self._interp(self.wf_T, self.wf_phi, T)
        return 0.0 if v is None else v

    def amplitude(self, T):
        if self.ap_T is None:
            return 1.0
        v = self._interp(self.ap_T, self.ap_factor, T)
        return 1.0 if v is None else v

    def phi(self, T):
        return self.phi_const + (self.harmonic(T) + self.waveform(T)) * self.amplitude(T)


def _pairs(path, skip):
    """(column 1, column 2) rows of a table file after `skip` lines until the first line that
    does not hold two numbers, as a list-directed READ loop does; column 1 goes through a REAL
    (float32) variable, column 2 into a REAL(8) array"""
    rows = []
    with open(path) as f:
        lines = f.read().splitlines()[skip:]
    for line in lines:
        parts = line.replace(",", " ").split()
        try:
            a, b = float(parts[0]), float(parts[1])
        except (IndexError, ValueError):
            break
        rows.append((np.float32(a), b))
    return rows


def read_waveform(el, path, dt_s, F_scale_V):
    """init_bo_NN_waveform.dat (pic2d_Setup.f90:185-270): amplitude [V], frequency [Hz], two
    comment lines, then (fraction of the period, relative potential) rows."""
    with open(path) as f:
        head = f.read().splitlines()[:2]
    rows = _pairs(path, 4)
    if len(rows) <= 1:
        return
    wf_phi_V = float(head[0].split()[0])
    wf_nu_Hz = float(head[1].split()[0])
    if wf_phi_V == 0.0:
        return
    period = int(1.0 / (wf_nu_Hz * dt_s))
    if period < len(rows) - 1:
        return
    T = [int(np.float32(r[0]) * np.float32(period)) for r in rows]
    phi = [(wf_phi_V / F_scale_V) * float(r[1]) for r in rows]
    T[0], T[-1] = 0, period
    for i in range(1, len(T) - 1):
        T[i] = max(T[i], T[i - 1] + 1)
    if T[-1] > T[-2]:
        el.wf_T, el.wf_phi = T, phi


def read_amplitude_profile(el, path, dt_s):
    """init_bo_NN_amplitude_profile.dat (pic2d_Setup.f90:300-400): a comment line, then (time
    [ns], factor) rows; with a waveform the ends of the non-zero intervals move to whole
    waveform periods."""
    rows = _pairs(path, 1)
    if len(rows) <= 1:
        return
    T = [int(float(r[0]) * 1.0e-9 / dt_s) for r in rows]
    fac = [float(r[1]) for r in rows]
    T[0] = 0
    for i in range(1, len(T) - 1):
        T[i] = max(T[i], T[i - 1] + 1)
    if el.wf_T is not None:
        wp = el.wf_T[-1]
        for i in range(len(T) - 1):
            if fac[i] == 0.0 and fac[i + 1] != 0.0:
                d = T[i + 1] - T[i]
                T[i] = wp * (T[i] // wp)
                T[i + 1] = T[i] + d
        for i in range(1, len(T)):
            if fac[i - 1] != 0.0 and fac[i] == 0.0:
                d = T[i] - T[i - 1]
                T[i] = wp * (T[i] // wp)
                T[i - 1] = T[i] - d
        T[-1] = wp * (T[-1] // wp)
    if all(T[i + 1] > T[i] for i in range(len(T) - 1)):
        el.ap_T, el.ap_factor = T, fac


class WallPotentials:
    """The electrodes of a deck: phi(T_cntr) gives the normalised potential of every object
    (metal objects from their schedule, the others as the problem holds them)."""

    def __init__(self, directory, bo, metal, base_phi, dt_s, F_scale_V):
        self.base = list(base_phi)
        self.el = {}
        for n, b in enumerate(bo):
            if n not in metal:
                continue
            el = Electrode(b["phi_V"] / F_scale_V, b.get("phi_var_V", 0.0) / F_scale_V,
                           b.get("omega_MHz", 0.0) * 2.0e6 * math.pi * dt_s, b.get("phase_deg", 0.0) * math.pi / 180.0)
            wf = os.path.join(directory, "init_bo_%02d_waveform.dat" % (n + 1))
            if os.path.exists(wf):
                read_waveform(el, wf, dt_s, F_scale_V)
            ap = os.path.join(directory, "init_bo_%02d_amplitude_profile.dat" % (n + 1))
            if os.path.exists(ap):
                read_amplitude_profile(el, ap, dt_s)
            if el.varies:
                self.el[n] = el

    @property
    def varies(self):
        return bool(self.el)

    def phi(self, T_cntr):
        out = list(self.base)
        for n, el in self.el.items():
            out[n] = el.phi(T_cntr)
        return out
