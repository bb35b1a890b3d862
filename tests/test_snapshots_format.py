"""Snapshot writer (CPU): the SAVE_GLOBAL_2D_ARRAY matrix layout worked by hand, the
single-precision normalisation of the moments against closed forms (a cold beam has no
temperature and no heat flow, two opposite beams have temperature v^2), the wall density
correction, and the file names of CREATE_SNAPSHOT."""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi as O
from edgpu import config
from edgpu import snapshots as S


def test_matrix_layout_by_hand(tmp_path):
    p = tmp_path / "a.bin"
    a = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]], np.float32)  # gmi = 2, gmj = 1: a[j, i]
    S.save_global_2d_array(a, p, 2, 1, 0.5)
    raw = np.fromfile(p, np.float32)
    assert raw.tolist() == [3.0, 0.0, 0.5, 1.0,
                            0.0, 1.0, 2.0, 3.0,
                            0.5, 4.0, 5.0, 6.0]
    x, y, b = S.read_global_2d_array(p, 2, 1)
    assert np.array_equal(b, a) and x.tolist() == [0.0, 0.5, 1.0] and y.tolist() == [0.0, 0.5]


def _raw(prob, x, y, vx, vy, vz):
    g = O.make_geom(prob.geometry())
    acc = np.zeros(13 * prob.nx * prob.ny, np.float32)
    O.lib().orc_moments_raw(C.byref(g), x.size, O.dp(x), O.dp(y), O.dp(vx), O.dp(vy), O.dp(vz),
                            acc.ctypes.data_as(C.POINTER(C.c_float)))
    return acc.reshape(13, prob.ny, prob.nx)


def test_cold_beam_and_two_streams():
    prob = config.workload(1, 32)  # 8 x 8 cells, X-periodic
    rng = np.random.default_rng(1)
    n = 20000
    x = rng.uniform(1.0, prob.gmi - 2.0, n)
    y = rng.uniform(1.0, prob.gmj - 1.0, n)
    v = (0.3, -0.2, 0.1)
    m = S.normalise_moments(_raw(prob, x, y, *[np.full(n, c) for c in v]), True)
    inner = m["N"] > 10
    for q, c in zip(("VX", "VY", "VZ"), v):
        assert np.allclose(m[q][inner], c, rtol=1e-5)
    for q in ("QX", "QY", "QZ"):
        assert np.max(np.abs(m[q][inner])) < 1e-5  # a cold beam carries no heat
    t = S._temperature(m["N"], m["WX"], m["VX"], 1.0)
    assert np.max(t[inner]) < 1e-6
    # two opposite beams along y: mean 0, temperature v^2
    s = np.where(rng.uniform(0, 1, n) < 0.5, -0.4, 0.4)
    m = S.normalise_moments(_raw(prob, x, y, np.zeros(n), s, np.zeros(n)), True)
    ty = S._temperature(m["N"], m["WY"], m["VY"], 1.0)
    full = m["N"] > 200
    assert np.allclose(ty[full] + m["VY"][full] ** 2, 0.16, rtol=1e-5)


def test_wall_density_correction():
    raw = np.zeros((13, 4, 5), np.float32)
    raw[0] = 1.0
    m = S.normalise_moments(raw, False)
    n = m["N"]
    assert n[0, 0] == n[-1, -1] == n[0, -1] == n[-1, 0] == 4.0
    assert np.all(n[0, 1:-1] == 2.0) and np.all(n[1:-1, 0] == 2.0) and np.all(n[1:-1, 1:-1] == 1.0)
    n = S.normalise_moments(raw, True)["N"]
    assert np.all(n[0] == 2.0) and np.all(n[-1] == 2.0) and np.all(n[1:-1] == 1.0)


def test_file_names():
    prob = config.workload(1, 32)
    m = {k: np.zeros((2, 2), np.float32) for k in S.MOMENT_NAMES}
    names = sorted(S.species_files(m, 7, S.scales(prob), 0))
    assert "_0007_Ne_m3_2D.bin" in names and "_0007_JXe_Am2_2D.bin" in names
    assert "_0007_TZe_eV_2D.bin" in names and "_0007_QYe_Wm2_2D.bin" in names and len(names) == 16
    names = sorted(S.species_files(m, 7, S.scales(prob), 1, 1, 7.3e4))
    assert "_0007_Ni_1_m3_2D.bin" in names and "_0007_VZi_1_ms_2D.bin" in names


@pytest.mark.parametrize("x,y,weights", [
    # on a node: everything to that node
    (3.0, 2.0, {(3, 2): 1.0}),
    # a cell centre: a quarter to each corner
    (3.5, 2.5, {(3, 2): 0.25, (4, 2): 0.25, (3, 3): 0.25, (4, 3): 0.25}),
    # a quarter of the way along x on a row of nodes
    (3.25, 2.0, {(3, 2): 0.75, (4, 2): 0.25}),
    # the upper edges (x = gmi = 9, y = gmj = 8) belong to the last cell, whose far corner
    # node takes the whole weight
    (9.0, 8.0, {(9, 8): 1.0}),
    (9.0, 5.5, {(9, 5): 0.5, (9, 6): 0.5}),
])
def test_raw_moments_single_particle(x, y, weights):
    """The 13 accumulators of COLLECT_*_MOMENTS (pic2d_ElectronMoments.f90:140-229) for one
    particle with v = (1, 2, 3), worked by hand: N, then vx, vy, vz, then vx^2, vy^2, vz^2,
    then vx vy, vx vz, vy vz, then |v|^2 (vx, vy, vz) with |v|^2 = 14, each times the
    bilinear weight of the node."""
    prob = config.workload(1, 32)  # 8 x 8 cells
    acc = _raw(prob, np.array([x]), np.array([y]), np.array([1.0]), np.array([2.0]), np.array([3.0]))
    per_moment = [1, 1, 2, 3, 1, 4, 9, 2, 3, 6, 14, 28, 42]
    want = np.zeros_like(acc)
    for (i, j), w in weights.items():
        want[:, j, i] = [w * m for m in per_moment]
    assert np.array_equal(acc, want)
