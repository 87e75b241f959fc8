"""CPU reference of the learned warm-start MLP (test infrastructure only).

Same math as riftphys-rust_amd/csrc/rp_mlp.h / rp_mlp.hip: h1 = relu(W1 x + b1),
h2 = relu(W2 h1 + b2), y = W3 h2 + b3 on normalised inputs, loss = mean over rows of
sum_o (y_o - t_o / y_scale_o)^2.  `bf16=True` rounds to bfloat16 (nearest even) where the device
feeds the matrix cores (inputs, weights, activations, back-propagated deltas) and accumulates in
float64, so device and reference differ only by accumulation order; `bf16=False` is the exact
float64 model whose gradient the finite-difference test pins.
"""
import numpy as np

NPARAM = 64 * 12 + 64 + 64 * 64 + 64 + 3 * 64 + 3 + 12 + 12 + 3
NW = NPARAM - 27


def to_bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    out = {}
    for name, shape in (("W1", (64, 12)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (3, 64)), ("b3", (3,)),
                        ("xm", (12,)), ("xs", (12,)), ("ys", (3,))):
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def init(seed=0, xm=None, xs=None, ys=None):
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    o = 0
    for (fo, fi) in ((64, 12), (64, 64), (3, 64)):
        a = np.sqrt(6.0 / (fi + fo))
        p[o:o + fo * fi] = rng.uniform(-a, a, fo * fi)
        o += fo * fi
        p[o:o + fo] = rng.uniform(-0.1, 0.1, fo)
        o += fo
    p[o:o + 12] = 0.0 if xm is None else xm
    p[o + 12:o + 24] = 1.0 if xs is None else xs
    p[o + 24:o + 27] = 1.0 if ys is None else ys
    return p


def _q(a, bf16):
    return to_bf16(np.asarray(a, np.float32)).astype(np.float64) if bf16 else np.asarray(a, np.float64)


def forward(p, feat, bf16=True):
    """Returns (y in normalised units, y in output units, cache)."""
    P = unpack(p)
    x = (np.asarray(feat, np.float32) - P["xm"].astype(np.float32)) / P["xs"].astype(np.float32)
    xb = _q(x, bf16)
    W1, W2, W3 = _q(P["W1"], bf16), _q(P["W2"], bf16), _q(P["W3"], bf16)
    a1 = xb @ W1.T + P["b1"]
    h1 = _q(np.maximum(a1, 0.0), bf16)
    a2 = h1 @ W2.T + P["b2"]
    h2 = _q(np.maximum(a2, 0.0), bf16)
    y = h2 @ W3.T + P["b3"]
    return y, y * P["ys"], dict(x=xb, h1=h1, h2=h2, W1=W1, W2=W2, W3=W3, P=P)


def loss_grad(p, feat, targ, bf16=True):
    """(loss, gradient of the first NW parameters) of one step over all given rows."""
    y, _, c = forward(p, feat, bf16)
    P = c["P"]
    B = len(feat)
    t = np.asarray(targ, np.float64) / P["ys"]
    diff = y - t
    loss = float(np.sum(diff * diff) / B)
    dy = 2.0 * diff / B
    dyq = _q(dy, bf16)
    g2 = (dyq @ c["W3"]) * (c["h2"] > 0)
    g2q = _q(g2, bf16)
    g1 = (g2q @ c["W2"]) * (c["h1"] > 0)
    g1q = _q(g1, bf16)
    grad = np.concatenate([(g1q.T @ c["x"]).ravel(), g1.sum(0), (g2q.T @ c["h1"]).ravel(), g2.sum(0),
                           (dyq.T @ c["h2"]).ravel(), dy.sum(0)])
    assert grad.size == NW
    return loss, grad
