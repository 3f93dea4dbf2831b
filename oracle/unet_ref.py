"""unet_ref.py — CPU oracle of the learned pressure warm start.  TEST INFRASTRUCTURE ONLY.

A numpy restatement of the 3-D U-Net that coffee_sim_amd/csrc/cs_unet.hip runs
on the GPU, forward and backward, used only by tests/ to check the GPU
kernels' activations and gradients.  The network has no counterpart in the
reference (the reference resets the pressure to zero every substep,
shader.rs:1187); its shape is the one the benchmark configuration states:
3 levels with 16/32/64 channels, 3x3x3 kernels, bf16 weights and activations,
fp32 accumulation.

Layout: activations are channels-last arrays [n, n, n, C] indexed [z, y, x, c];
a 3x3x3 kernel is W[tap, cin, cout] with tap = (dz+1)*9 + (dy+1)*3 + (dx+1),
zero padding at the grid edge.

Rounding points mirror the GPU: conv inputs and weights are bf16 values, sums
are exact here (float64) where the GPU accumulates in fp32, and every stored
activation or activation gradient is rounded to bf16 (round to nearest even).
The backward pass is pinned independently by central finite differences in
float64 with rounding switched off (tests/test_unet_oracle.py).
"""
from __future__ import annotations

import numpy as np

TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
IN_CH = 16      # 4 features padded to 16
LEVELS = (16, 32, 64)


def bf16(x: np.ndarray, on: bool = True) -> np.ndarray:
    """Round float values to bfloat16 (nearest, ties to even), returned as float64."""
    if not on:
        return np.asarray(x, np.float64)
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, 0x7FC00000, r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def _shift(x: np.ndarray, dz: int, dy: int, dx: int) -> np.ndarray:
    """y[z, y, x] = x[z + dz, y + dy, x + dx] with zeros outside."""
    n = x.shape[0]
    out = np.zeros_like(x)
    zs, ys, xs = [slice(max(0, -d), n - max(0, d)) for d in (dz, dy, dx)]
    zd, yd, xd = [slice(max(0, d), n - max(0, -d)) for d in (dz, dy, dx)]
    out[zs, ys, xs] = x[zd, yd, xd]
    return out


def conv3(x: np.ndarray, w: np.ndarray, b: np.ndarray | None) -> np.ndarray:
    """Pre-activation of a 3x3x3 'same' convolution: sum over taps and input channels."""
    out = np.zeros(x.shape[:3] + (w.shape[2],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        out += _shift(x, dz, dy, dx) @ w[t]
    if b is not None:
        out += b
    return out


def conv3_dgrad(g: np.ndarray, w: np.ndarray) -> np.ndarray:
    """d/dx of conv3 given the gradient g of its pre-activation."""
    out = np.zeros(g.shape[:3] + (w.shape[1],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        out += _shift(g, -dz, -dy, -dx) @ w[t].T
    return out


def conv3_wgrad(x: np.ndarray, g: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """d/dW and d/db of conv3 given the gradient g of its pre-activation."""
    cin, cout = x.shape[3], g.shape[3]
    dw = np.zeros((27, cin, cout))
    gm = g.reshape(-1, cout)
    for t, (dz, dy, dx) in enumerate(TAPS):
        dw[t] = _shift(x, dz, dy, dx).reshape(-1, cin).T @ gm
    return dw, gm.sum(axis=0)


def pool2(x):
    n = x.shape[0] // 2
    return x.reshape(n, 2, n, 2, n, 2, -1).mean(axis=(1, 3, 5))


def pool2_back(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def up2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up2_back(g):
    n = g.shape[0] // 2
    return g.reshape(n, 2, n, 2, n, 2, -1).sum(axis=(1, 3, 5))


CONVS = [  # (name, cin, cout)
    ("e1a", IN_CH, 16), ("e1b", 16, 16),
    ("e2a", 16, 32), ("e2b", 32, 32),
    ("ba", 32, 64), ("bb", 64, 64),
    ("d2a", 96, 32), ("d2b", 32, 32),
    ("d1a", 48, 16), ("d1b", 16, 16),
]


def init_params(seed: int) -> dict:
    """He-normal weights, zero biases (the GPU net is initialised the same way
    from the same seed by tests that compare the two)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, cin, cout in CONVS:
        p[name + ".w"] = rng.normal(0.0, np.sqrt(2.0 / (27 * cin)), size=(27, cin, cout))
        p[name + ".b"] = np.zeros(cout)
    p["head.w"] = rng.normal(0.0, np.sqrt(1.0 / 16), size=(16,))
    p["head.b"] = np.zeros(1)
    return p


def forward(p: dict, x: np.ndarray, rounding: bool = True) -> tuple[np.ndarray, dict]:
    """Prediction [n, n, n] and the saved activations."""
    r = lambda a: bf16(a, rounding)  # noqa: E731
    W = {k: (r(v) if k.endswith(".w") and not k.startswith("head") else np.asarray(v, np.float64))
         for k, v in p.items()}
    s = {"x": r(x)}

    def cr(name, inp):
        return r(np.maximum(conv3(inp, W[name + ".w"], W[name + ".b"]), 0.0))

    s["e1a"] = cr("e1a", s["x"])
    s["e1b"] = cr("e1b", s["e1a"])
    s["p1"] = r(pool2(s["e1b"]))
    s["e2a"] = cr("e2a", s["p1"])
    s["e2b"] = cr("e2b", s["e2a"])
    s["p2"] = r(pool2(s["e2b"]))
    s["ba"] = cr("ba", s["p2"])
    s["bb"] = cr("bb", s["ba"])
    s["c2"] = np.concatenate([up2(s["bb"]), s["e2b"]], axis=3)
    s["d2a"] = cr("d2a", s["c2"])
    s["d2b"] = cr("d2b", s["d2a"])
    s["c1"] = np.concatenate([up2(s["d2b"]), s["e1b"]], axis=3)
    s["d1a"] = cr("d1a", s["c1"])
    s["d1b"] = cr("d1b", s["d1a"])
    pred = s["d1b"] @ W["head.w"] + W["head.b"][0]
    s["W"] = W
    return pred, s


def loss_and_grad(pred: np.ndarray, target: np.ndarray, mask: np.ndarray) -> tuple[float, np.ndarray]:
    """Masked mean squared error and its gradient with respect to pred."""
    m = np.asarray(mask, np.float64)
    cnt = max(m.sum(), 1.0)
    d = (pred - target) * m
    return float((d * d).sum() / cnt), 2.0 * d / cnt


def backward(p: dict, s: dict, gpred: np.ndarray, rounding: bool = True) -> dict:
    """Gradients of every parameter given d loss / d pred."""
    r = lambda a: bf16(a, rounding)  # noqa: E731
    W = s["W"]
    g = {}
    g["head.w"] = (s["d1b"] * gpred[..., None]).reshape(-1, 16).sum(axis=0)
    g["head.b"] = np.array([gpred.sum()])
    d = r(gpred[..., None] * W["head.w"])  # d loss / d d1b

    def conv_back(name, inp, out, dout, want_input=True):
        gz = dout * (out > 0)  # through the ReLU
        dw, db = conv3_wgrad(inp, gz)
        g[name + ".w"], g[name + ".b"] = dw, db
        return r(conv3_dgrad(gz, W[name + ".w"])) if want_input else None

    d = conv_back("d1b", s["d1a"], s["d1b"], d)
    dc1 = conv_back("d1a", s["c1"], s["d1a"], d)
    d_up1, d_e1b_skip = dc1[..., :32], dc1[..., 32:]
    d = r(up2_back(d_up1))  # d loss / d d2b
    d = conv_back("d2b", s["d2a"], s["d2b"], d)
    dc2 = conv_back("d2a", s["c2"], s["d2a"], d)
    d_up2, d_e2b_skip = dc2[..., :64], dc2[..., 64:]
    d = r(up2_back(d_up2))  # d loss / d bb
    d = conv_back("bb", s["ba"], s["bb"], d)
    d = conv_back("ba", s["p2"], s["ba"], d)
    d_e2b = r(d_e2b_skip + pool2_back(d))
    d = conv_back("e2b", s["e2a"], s["e2b"], d_e2b)
    d = conv_back("e2a", s["p1"], s["e2a"], d)
    d_e1b = r(d_e1b_skip + pool2_back(d))
    d = conv_back("e1b", s["e1a"], s["e1b"], d_e1b)
    conv_back("e1a", s["x"], s["e1a"], d, want_input=False)
    return g


def param_vector(p: dict) -> np.ndarray:
    """Parameters flattened in the GPU's order (cs_unet.hip, `param_layout`)."""
    parts = []
    for name, _, _ in CONVS:
        parts += [np.asarray(p[name + ".w"]).reshape(-1), np.asarray(p[name + ".b"]).reshape(-1)]
    parts += [np.asarray(p["head.w"]).reshape(-1), np.asarray(p["head.b"]).reshape(-1)]
    return np.concatenate(parts)


def param_dict(v: np.ndarray) -> dict:
    p, o = {}, 0
    for name, cin, cout in CONVS:
        n = 27 * cin * cout
        p[name + ".w"] = v[o:o + n].reshape(27, cin, cout)
        o += n
        p[name + ".b"] = v[o:o + cout]
        o += cout
    p["head.w"] = v[o:o + 16]
    o += 16
    p["head.b"] = v[o:o + 1]
    return p


def adam_step(p: np.ndarray, g: np.ndarray, m: np.ndarray, v: np.ndarray, t: int, lr: float,
              b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8, folded: bool = False):
    """One Adam step (Kingma & Ba 2015) in float64, step t >= 1.  Returns the
    new (parameters, first moment, second moment).  `folded` gives the form of
    the paper's Section 2 closing remark that the device uses: the bias
    corrections folded into the step, lr_t = lr sqrt(1 - b2^t) / (1 - b1^t),
    and p -= lr_t m / (sqrt(v) + eps), which is Algorithm 1 with eps divided
    by sqrt(1 - b2^t)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    if folded:
        lr_t = lr * np.sqrt(1.0 - b2 ** t) / (1.0 - b1 ** t)
        return p - lr_t * m / (np.sqrt(v) + eps), m, v
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
