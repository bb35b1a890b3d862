"""CPU oracle of the learned warm-start network -- TEST INFRASTRUCTURE ONLY.

numpy restatement of the network the GPU library trains (6 layers of 3x3 convolutions,
32 channels, ReLU between layers, zero padding, MSE over the free nodes), used by tests/
to check the GPU forward pass and parameter gradients.  The reference has no counterpart:
it cold-starts every Krylov solve.

Parameter vector layout (as the library's): per layer l the weights [tap][cout][cin]
(tap = 3*(dy+1) + (dx+1)) then the bias [cout]; layer 0 has cin = 16 (4 features, 12 zero
channels), layers 1..4 32 -> 32, layer 5 32 -> 16 (only output channel 0 is used).

With bf16=True the values are rounded to bfloat16 where the GPU stores them: features,
weights, the activations after each ReLU, the loss gradient and the data gradients; all
sums are FP32.  With bf16=False everything is FP64 (for finite-difference checks of this
file's own backward pass).
"""
import numpy as np

LAYERS = 6


def shapes():
    out = []
    for l in range(LAYERS):
        cin = 16 if l == 0 else 32
        cout = 16 if l == LAYERS - 1 else 32
        out.append((cin, cout))
    return out


def n_params():
    return sum(9 * ci * co + co for ci, co in shapes())


def unpack(p):
    ws, bs = [], []
    off = 0
    for ci, co in shapes():
        ws.append(p[off:off + 9 * co * ci].reshape(9, co, ci))
        off += 9 * co * ci
        bs.append(p[off:off + co])
        off += co
    return ws, bs


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def _shift(x, dy, dx):
    """y[i, j] = x[i + dy, j + dx] with zeros outside (x: H, W, C)."""
    H, W = x.shape[:2]
    y = np.zeros_like(x)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    if ys < ye and xs < xe:
        y[ys:ye, xs:xe] = x[ys + dy:ye + dy, xs + dx:xe + dx]
    return y


def conv(x, w):
    """out[p, co] = sum_tap sum_ci w[tap, co, ci] x[p + off(tap), ci]."""
    out = None
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        t = np.tensordot(_shift(x, dy, dx), w[tap], axes=([2], [1]))
        out = t if out is None else out + t
    return out


def forward_backward(params, feat, target, bf16_mode=True):
    """feat (H, W, 4), target (H, W).  Returns prediction (H, W), loss, gradient vector."""
    dt = np.float32 if bf16_mode else np.float64
    rnd = bf16 if bf16_mode else (lambda a: a)
    H, W = target.shape
    p = np.asarray(params, dt)
    ws, bs = unpack(p)
    wq = [rnd(w).astype(dt) for w in ws]
    x0 = np.zeros((H, W, 16), dt)
    x0[:, :, :4] = rnd(np.asarray(feat, dt))
    acts = [x0]
    for l in range(LAYERS - 1):
        a = np.maximum(conv(acts[-1], wq[l]) + bs[l].astype(dt), 0).astype(dt)
        acts.append(rnd(a).astype(dt))
    out = conv(acts[-1], wq[-1]) + bs[-1].astype(dt)
    pred = out[:, :, 0]
    free = np.asarray(feat)[:, :, 3] > 0.5
    nfree = int(free.sum())
    e = np.where(free, pred - np.asarray(target, dt), 0).astype(dt)
    loss = float((e.astype(np.float64) ** 2).sum() / nfree)
    g = np.zeros((H, W, 16), dt)
    g[:, :, 0] = rnd((2.0 / nfree) * e).astype(dt)
    grads_w, grads_b = [None] * LAYERS, [None] * LAYERS
    for l in range(LAYERS - 1, -1, -1):
        xin = acts[l]
        gw = np.zeros_like(ws[l], dtype=dt)
        for tap in range(9):
            dy, dx = tap // 3 - 1, tap % 3 - 1
            gw[tap] = np.tensordot(g, _shift(xin, dy, dx), axes=([0, 1], [0, 1]))
        grads_w[l] = gw
        grads_b[l] = g.sum(axis=(0, 1))
        if l == 0:
            break
        # data gradient: correlation with the flipped, transposed weights, then the ReLU gate
        wt = np.stack([wq[l][8 - tap].T for tap in range(9)])  # [tap][cin][cout]
        gin = conv(g, wt)
        gin = np.where(acts[l] > 0, gin, 0).astype(dt)
        g = rnd(gin).astype(dt)
    flat = []
    for l, (ci, co) in enumerate(shapes()):
        flat.append(grads_w[l].reshape(-1))
        flat.append(grads_b[l].reshape(-1))
    grad = np.concatenate(flat)
    # frozen parameters (padded channels) carry no gradient
    return pred, loss, grad * frozen_mask(dt)


def frozen_mask(dt=np.float64):
    """1 for trained parameters, 0 for the padded channels kept at zero."""
    m = []
    for l, (ci, co) in enumerate(shapes()):
        w = np.ones((9, co, ci), dt)
        b = np.ones(co, dt)
        if l == 0:
            w[:, :, 4:] = 0
        if l == LAYERS - 1:
            w[:, 1:, :] = 0
            b[1:] = 0
        m.append(w.reshape(-1))
        m.append(b)
    return np.concatenate(m)


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    out = []
    for l, (ci, co) in enumerate(shapes()):
        cin_real = 4 if l == 0 else ci
        out.append(rng.normal(0, scale * np.sqrt(2.0 / (9 * cin_real)), 9 * co * ci))
        out.append(rng.normal(0, 0.05, co))
    return np.concatenate(out) * frozen_mask()
