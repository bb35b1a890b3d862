"""Snapshot files written from the device state against files built from the oracle's
moments of the same particles (half-advanced electron velocities, as the reference)."""
import ctypes as C
import os

import numpy as np
import pytest

from edgpu import config
from edgpu import snapshots as S
from helpers import oracle_particles, synthetic_field

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [2, 4])
def test_snapshot_matches_oracle_moments(edgpu, oracle, tmp_path, n):
    O = oracle
    prob = config.workload(n, 32)
    rt = edgpu.Runtime(prob, fuse_collisions=0)
    try:
        e, i = oracle_particles(O, prob)
        if prob.inner_boxes:
            inside = np.zeros(e[0].size, bool)
            for (i0, i1, j0, j1) in prob.inner_boxes:
                inside |= (e[0] > i0) & (e[0] < i1) & (e[1] > j0) & (e[1] < j1)
            e = [a[~inside] for a in e]
            i = [a[~inside] for a in i]
        rt.upload(0, *e)
        rt.upload(1, *i)
        EX, EY = synthetic_field(prob, amp=0.05)
        rt.set_grid("EX", EX)
        rt.set_grid("EY", EY)
        phi = np.linspace(-1.0, 1.0, prob.nx * prob.ny)
        rt.set_grid("phi", phi)
        names = S.write_snapshot(rt, str(tmp_path), 3)
        # oracle: raw moments of the same particles, electrons half-advanced
        g = O.make_geom(prob.geometry())
        pushed = [a.copy() for a in e]
        O.advance(g, O.make_species(True), EX, EY, (prob.Bx, prob.By, prob.Bz), prob.Ez, *pushed,
                  prob.n_boundary_obj + len(prob.inner_boxes))
        half = [0.5 * (e[q] + pushed[q]) for q in (2, 3, 4)]
        sc = S.scales(prob)
        want = {}
        for s, (x, y, vx, vy, vz) in ((0, (e[0], e[1], *half)), (1, tuple(i))):
            acc = np.zeros(13 * prob.nx * prob.ny, np.float32)
            O.lib().orc_moments_raw(C.byref(g), x.size, O.dp(x), O.dp(y), O.dp(np.ascontiguousarray(vx)),
                                    O.dp(np.ascontiguousarray(vy)), O.dp(np.ascontiguousarray(vz)),
                                    acc.ctypes.data_as(C.POINTER(C.c_float)))
            m = S.normalise_moments(acc.reshape(13, prob.ny, prob.nx), prob.periodic_x)
            want.update(S.species_files(m, 3, sc, s, prob.Qs, prob.Ms))
        want["_0003_F_V_2D.bin"] = (phi * sc["F_scale_V"]).astype(np.float32).reshape(prob.ny, prob.nx)
        want["_0003_EX_Vm_2D.bin"] = (EX * sc["E_scale_Vm"]).astype(np.float32).reshape(prob.ny, prob.nx)
        assert set(want) <= set(names)
        worst = 0.0
        for name, ref in want.items():
            path = os.path.join(str(tmp_path), name)
            assert os.path.getsize(path) == 4 * (prob.gmi + 2) * (prob.gmj + 2)
            x, y, got = S.read_global_2d_array(path, prob.gmi, prob.gmj)
            assert x[1] == np.float32(prob.dx) and y[1] == np.float32(prob.dx)
            scale = max(float(np.max(np.abs(ref))), 1e-30)
            err = float(np.max(np.abs(got - ref))) / scale
            worst = max(worst, err)
            assert err < 2e-4, (name, err)
        print("snapshot files %d, worst relative difference %.2e" % (len(want), worst), flush=True)
    finally:
        rt.close()
