"""Pinning of the network oracle (oracle/cnn_ref.py, CPU only): its backward pass against
central finite differences of its own forward pass in FP64, and its bfloat16 rounding
against hand-worked bit patterns."""
import numpy as np

import cnn_ref as R


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], np.float32)
    got = R.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins; just above rounds up;
    # 1 + 3*2^-8 ties between 1+2^-7 and 1+2^-6: even is 1 + 2^-6
    want = np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, np.float32(3.0e-3)], np.float32)
    assert np.array_equal(got[:5], want[:5])
    assert abs(got[5] - 3.0e-3) / 3.0e-3 < 2 ** -8


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(11)
    H, W = 7, 9
    feat = rng.normal(0, 1, (H, W, 4))
    feat[:, :, 3] = (rng.uniform(0, 1, (H, W)) > 0.2).astype(float)
    target = rng.normal(0, 1, (H, W))
    p = R.random_params(3)
    _, loss, grad = R.forward_backward(p, feat, target, bf16_mode=False)
    trained = np.nonzero(R.frozen_mask())[0]
    pick = rng.choice(trained, 40, replace=False)
    # include parameters of every layer
    offs = np.cumsum([0] + [9 * ci * co + co for ci, co in R.shapes()])
    pick = np.concatenate([pick, [offs[l] + 5 for l in range(R.LAYERS)], [offs[-1] - 16]])
    h = 1e-6
    for q in pick:
        e = np.zeros_like(p)
        e[q] = h
        lp = R.forward_backward(p + e, feat, target, bf16_mode=False)[1]
        lm = R.forward_backward(p - e, feat, target, bf16_mode=False)[1]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, grad[q])


def test_zero_last_layer_predicts_bias():
    rng = np.random.default_rng(2)
    H, W = 5, 6
    feat = rng.normal(0, 1, (H, W, 4))
    feat[:, :, 3] = 1.0
    p = R.random_params(4)
    ws, bs = R.unpack(p)
    ws[-1][:] = 0.0
    bs[-1][0] = 0.25
    pred, _, _ = R.forward_backward(p, feat, np.zeros((H, W)), bf16_mode=False)
    assert np.allclose(pred, 0.25)


def test_conv_matches_direct_sum():
    """The oracle's 3x3 convolution against a per-pixel loop with explicit zero padding
    (tap = 3*(dy+1) + (dx+1), row offset dy, column offset dx)."""
    rng = np.random.default_rng(5)
    H, W, ci, co = 4, 5, 3, 2
    x = rng.normal(0, 1, (H, W, ci))
    w = rng.normal(0, 1, (9, co, ci))
    got = R.conv(x, w)
    want = np.zeros((H, W, co))
    for i in range(H):
        for j in range(W):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    if 0 <= i + dy < H and 0 <= j + dx < W:
                        want[i, j] += w[3 * (dy + 1) + (dx + 1)] @ x[i + dy, j + dx]
    assert np.allclose(got, want, rtol=1e-13, atol=1e-13)
    # an asymmetric kernel distinguishes the row and column offsets
    w2 = np.zeros((9, 1, 1))
    w2[3 * (1 + 1) + (0 + 1), 0, 0] = 1.0  # dy = +1, dx = 0: out[i, j] = x[i + 1, j]
    x2 = np.arange(H * W, dtype=float).reshape(H, W, 1)
    out2 = R.conv(x2, w2)[:, :, 0]
    assert np.array_equal(out2[:-1], x2[1:, :, 0]) and np.all(out2[-1] == 0)
