"""TEST INFRASTRUCTURE: numpy restatement of the learned con2prim pressure
guess (3 -> 64 -> 64 -> 1 ReLU network) for checking the device forward and
backward passes.  No reference counterpart: the network is this library's
addition.  Checker only.

Parameter vector (float32, 4481): W1[64][3], b1[64], W2[64][64] ([out][in]),
b2[64], W3[64], b3.  With `emulate=True` the values are rounded to bf16
where the device rounds them (W1, W2, W3 working copies; the layer-1
activations and the layer-2 pre-activation gradient before the matrix-core
products); everything else is evaluated in float64.  With `emulate=False`
the network is plain float64, which the finite-difference test pins.
"""
import numpy as np

NPARAM = 4481
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 192, 256, 4352, 4416, 4480


def bf16(x):
    """round float32 values to bfloat16 (round to nearest even), as float32"""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32)
    r = ((u + np.uint32(0x7FFF) + ((u >> np.uint32(16)) & np.uint32(1))) & np.uint32(0xFFFF0000)).astype(np.uint32)
    out = r.view(np.float32).copy()
    out[np.isnan(a)] = np.nan
    return out


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(64, 3), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(64, 64),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3, [b3]])


def _r(x, emulate):
    return bf16(x).astype(np.float64) if emulate else np.asarray(x, dtype=np.float64)


def inputs(norm, feat, target=None):
    """normalised inputs / targets exactly as the device forms them (float32)"""
    norm = np.asarray(norm, dtype=np.float32)
    f = np.asarray(feat, dtype=np.float32).reshape(-1, 3)
    x = (f - norm[0:3]) * norm[3:6]
    t = None if target is None else (np.asarray(target, dtype=np.float32) - norm[6]) / norm[7]
    return x.astype(np.float64), (None if t is None else t.astype(np.float64))


def forward(p, x, emulate=True):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    W1q, W2q, W3q = _r(W1, emulate), _r(W2, emulate), _r(W3, emulate)
    z1 = x @ W1q.T + b1
    h1 = np.maximum(z1, 0.0)
    h1q = _r(h1, emulate)
    z2 = h1q @ W2q.T + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ W3q + b3
    return y, dict(x=x, h1q=h1q, h2=h2, W2q=W2q, W3q=W3q)


def loss_and_grad(p, x, t, emulate=True):
    """mean squared error over all points and its gradient (float64)"""
    B = x.shape[0]
    y, c = forward(p, x, emulate)
    r = y - t
    loss = np.mean(r * r)
    dy = 2.0 * r / B
    gW3 = c["h2"].T @ dy
    gb3 = dy.sum()
    dz2 = (c["h2"] > 0) * c["W3q"][None, :] * dy[:, None]
    gb2 = dz2.sum(axis=0)
    dz2q = _r(dz2, emulate)
    gW2 = dz2q.T @ c["h1q"]
    dh1 = dz2q @ c["W2q"]
    dz1 = dh1 * (c["h1q"] > 0)
    gW1 = dz1.T @ c["x"]
    gb1 = dz1.sum(axis=0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)
