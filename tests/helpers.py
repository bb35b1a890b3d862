"""Shared helpers of the parity tests: oracle-side state for a workload."""
import numpy as np


def oracle_particles(O, prob, seed=123456789, warmup=1000):
    """DISTRIBUTE_PARTICLES through the oracle's WELL stream (seed, warm-up draws, then
    electrons x, y, vx, vy, vz, less those placed inside inner objects, and ions vx, vy as the
    reference draws them)."""
    w = O.well(seed)
    if warmup:
        O.well_fill(w, warmup)
    v, _ = O.maxwell_tables()
    n = prob.n_particles
    e = [np.empty(n) for _ in range(5)]
    O.lib().orc_distribute(O.C.byref(w), O.dp(v), n, prob.x_limits[0], prob.x_limits[1], prob.y_limits[0],
                           prob.y_limits[1], 0.0, float(prob.gmi), 0.0, float(prob.gmj), prob.factor_convert_e,
                           *[O.dp(a) for a in e])
    if prob.inner_boxes:
        e = O.remove_inside_inner(O.make_geom(prob.geometry()), *e)
        n = e[0].size
    i = [np.empty(n) for _ in range(5)]
    O.lib().orc_distribute_ions(O.C.byref(w), O.dp(v), n, O.dp(e[0]), O.dp(e[1]), prob.factor_convert_i,
                                *[O.dp(a) for a in i])
    return e, i


def rel_err(a, b):
    scale = max(np.max(np.abs(a)), np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def sorted_rows(*cols):
    A = np.stack(cols, axis=1)
    order = np.lexsort(A.T[::-1])
    return A[order]


def synthetic_field(prob, amp=0.05, seed=7):
    """A smooth field on the nodes with both signs, strong enough to move particles across cells."""
    rng = np.random.default_rng(seed)
    x = np.arange(prob.nx)[None, :]
    y = np.arange(prob.ny)[:, None]
    kx = 2 * np.pi / max(prob.gmi - 1, 1)
    ky = np.pi / prob.gmj
    a, b, c = rng.uniform(0.5, 1.5, 3)
    EX = amp * (a * np.sin(3 * kx * x + 0.3) * np.cos(2 * ky * y) + 0.2 * b)
    EY = amp * (c * np.cos(2 * kx * x) * np.sin(3 * ky * y + 0.1) - 0.1)
    return np.ascontiguousarray(EX.ravel()), np.ascontiguousarray(EY.ravel())
