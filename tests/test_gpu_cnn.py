"""GPU checks of the learned warm start: the MFMA lane layout, the network's forward pass
and parameter gradients against the numpy oracle, and what the trained network gains over
the warm start from the previous potential in the Krylov solve."""
import numpy as np
import pytest

import cnn_ref as R
from edgpu import config
from helpers import oracle_particles

pytestmark = pytest.mark.gpu


def test_mfma_lane_layout(edgpu):
    """The lane layout of v_mfma_f32_16x16x32_bf16 the convolutions assume: a product whose
    every entry depends on all 32 terms of its row and column, so a misplaced operand lane
    shows."""
    rng = np.random.default_rng(1)
    A = rng.normal(0, 1, (16, 32)).astype(np.float32)
    B = rng.normal(0, 1, (32, 16)).astype(np.float32)
    D = edgpu.mfma_probe(A, B)
    ref = R.bf16(A).astype(np.float64) @ R.bf16(B).astype(np.float64)
    assert np.max(np.abs(D - ref)) < 1e-5 * np.max(np.abs(ref))


@pytest.mark.parametrize("H,W", [(24, 40), (17, 33)])
def test_network_forward_and_gradients_match_oracle(edgpu, H, W):
    """Prediction and every layer's parameter gradient within BF16 tolerance of the oracle,
    which rounds to bfloat16 at the same points (features, weights, activations, gradients)."""
    prob = config.workload(4, 16)
    rt = edgpu.Runtime(prob, fuse_collisions=0)
    try:
        rt.cnn_init(seed=5, max_pairs=1)
        p = R.random_params(7).astype(np.float32)
        rt.cnn_set_params(p)
        rng = np.random.default_rng(H * W)
        feat = rng.normal(0, 1, (H, W, 4)).astype(np.float32)
        feat[:, :, 3] = (rng.uniform(0, 1, (H, W)) > 0.15).astype(np.float32)
        target = rng.normal(0, 1, (H, W)).astype(np.float32)
        pred, grad, loss = rt.cnn_eval(feat, target)
        rpred, rloss, rgrad = R.forward_backward(p, feat, target, bf16_mode=True)
        print("network %dx%d: pred rel err %.2e, loss %.6f vs %.6f" % (
            H, W, np.max(np.abs(pred - rpred)) / np.max(np.abs(rpred)), loss, rloss), flush=True)
        assert np.max(np.abs(pred - rpred)) < 2e-2 * np.max(np.abs(rpred))
        assert abs(loss - rloss) < 2e-2 * rloss
        off = 0
        for l, (ci, co) in enumerate(R.shapes()):
            n = 9 * ci * co + co
            g, r = grad[off:off + n], rgrad[off:off + n]
            err = np.linalg.norm(g - r) / np.linalg.norm(r)
            print("  layer %d gradient rel err %.2e" % (l, err), flush=True)
            assert err < 3e-2, (l, err)
            off += n
        # frozen parameters get no gradient
        assert np.all(grad[R.frozen_mask() == 0] == 0)
    finally:
        rt.close()


def _krylov_step(rt, warm):
    rt.gather_ion_charge_density(host=False)
    rt.gather_electron_charge_density(host=False)
    _, it, rr = rt.solve_potential_krylov(rtol=1e-8, warm_start=warm, host=False)
    assert rr <= 1e-8
    rt.clear_accumulated_fields()
    rt.calculate_electric_field(host=False)
    rt.advance_electrons()
    rt.advance_ions()
    return it


def test_learned_warm_start_gain(edgpu, oracle):
    """Two-box geometry at 128x128 cells: record (rho, phi) pairs from the run's own solves,
    train on the device, then compare Krylov iterations to 1e-8 of the learned guess with
    the previous-potential guess and the zero guess over the following steps."""
    prob = config.workload(4, 8)
    rt = edgpu.Runtime(prob, fuse_collisions=0)
    try:
        # Jacobi preconditioning: hundreds of iterations, so the effect of the starting guess
        # is resolved by the iteration count (multigrid needs about ten either way)
        rt.set_krylov_preconditioner(0)
        e, i = oracle_particles(oracle, prob)
        inside = np.zeros(e[0].size, bool)
        for (i0, i1, j0, j1) in prob.inner_boxes:
            inside |= (e[0] > i0) & (e[0] < i1) & (e[1] > j0) & (e[1] < j1)
        rt.upload(0, *[a[~inside] for a in e])
        rt.upload(1, *[a[~inside] for a in i])
        rt.cnn_init(seed=3, max_pairs=32)
        _krylov_step(rt, 1)  # the first solve has no previous potential
        for _ in range(24):
            _krylov_step(rt, 1)
            rt.cnn_record_pair()
        print("scales", rt.cnn_scales(), flush=True)
        loss0 = rt.cnn_train(1, lr=1e-3)
        loss = rt.cnn_train(400, lr=1e-3)
        print("training loss %.4f -> %.4f" % (loss0, loss), flush=True)
        assert loss < 0.8 * loss0
        its = {}
        for warm in (0, 1, 2, 1, 2):
            its.setdefault(warm, []).extend(_krylov_step(rt, warm) for _ in range(4))
        m = {k: float(np.mean(v)) for k, v in its.items()}
        print("Krylov iterations to 1e-8: zero guess %.1f, previous potential %.1f, learned %.1f" % (
            m[0], m[1], m[2]), flush=True)
        # the learned guess must beat the zero guess and must not be worse than the previous
        # potential beyond the step-to-step spread of the iteration count
        assert m[2] < m[0]
        assert m[2] <= 1.02 * m[1]
    finally:
        rt.close()
