"""Pins the CPU oracle of the learned warm start (oracle/unet_ref.py) before it
is used to check the GPU kernels: the convolution against a direct loop, its
input gradient by the adjoint identity, and every parameter gradient of the
whole network by central finite differences in float64."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import unet_ref as U  # noqa: E402


def test_conv3_matches_a_direct_loop_with_zero_padding():
    rng = np.random.default_rng(0)
    n, cin, cout = 4, 3, 2
    x = rng.normal(size=(n, n, n, cin))
    w = rng.normal(size=(27, cin, cout))
    b = rng.normal(size=cout)
    got = U.conv3(x, w, b)
    want = np.zeros((n, n, n, cout))
    for z in range(n):
        for y in range(n):
            for xx in range(n):
                acc = b.copy()
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            zz, yy, x2 = z + dz, y + dy, xx + dx
                            if 0 <= zz < n and 0 <= yy < n and 0 <= x2 < n:
                                t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
                                acc += x[zz, yy, x2] @ w[t]
                want[z, y, xx] = acc
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_conv3_input_gradient_is_the_adjoint():
    rng = np.random.default_rng(1)
    n, cin, cout = 6, 4, 5
    x = rng.normal(size=(n, n, n, cin))
    y = rng.normal(size=(n, n, n, cout))
    w = rng.normal(size=(27, cin, cout))
    lhs = np.sum(U.conv3(x, w, None) * y)
    rhs = np.sum(x * U.conv3_dgrad(y, w))
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)


def test_pool_and_upsample_gradients_are_adjoints():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(8, 8, 8, 3))
    c = rng.normal(size=(4, 4, 4, 3))
    assert abs(np.sum(U.pool2(a) * c) - np.sum(a * U.pool2_back(c))) < 1e-10
    assert abs(np.sum(U.up2(c) * a) - np.sum(c * U.up2_back(a))) < 1e-10


def test_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 3.0e38], np.float64)
    got = U.bf16(x)
    # 1 + 2^-8 is the tie between 1 and 1 + 2^-7: even mantissa wins (1.0);
    # 1 + 3*2^-8 ties between 1 + 2^-7 and 1 + 2^-6: the even one is 1 + 2^-6
    np.testing.assert_array_equal(got[:5], [1.0, 1.0, 1.0 + 2 ** -6, -2.5, 0.0])


def test_network_gradients_match_finite_differences():
    n = 8
    rng = np.random.default_rng(3)
    p = U.init_params(7)
    # biases away from zero so every ReLU is exercised in both states
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.normal(0.0, 0.1, size=p[k].shape)
    x = rng.normal(size=(n, n, n, U.IN_CH))
    target = rng.normal(size=(n, n, n))
    mask = (rng.random((n, n, n)) < 0.7).astype(np.float64)

    def loss_of(v):
        pred, _ = U.forward(U.param_dict(v), x, rounding=False)
        return U.loss_and_grad(pred, target, mask)[0]

    v = U.param_vector(p)
    pred, s = U.forward(p, x, rounding=False)
    _, gp = U.loss_and_grad(pred, target, mask)
    gv = U.param_vector(U.backward(p, s, gp, rounding=False))
    picks = rng.choice(v.size, 40, replace=False)
    # include the head and a bias of every level
    picks = np.concatenate([picks, [v.size - 1, v.size - 2, v.size - 17]])
    eps = 1e-6
    for i in picks:
        vp, vm = v.copy(), v.copy()
        vp[i] += eps
        vm[i] -= eps
        fd = (loss_of(vp) - loss_of(vm)) / (2 * eps)
        assert abs(fd - gv[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, gv[i])


def test_adam_closed_forms():
    """Adam's bias correction makes a constant gradient move every parameter by
    lr * g / (|g| + eps) at every step (the first step included); a gradient
    that flips sign at step 2 gives m_hat = -g (0.01 / 0.19),
    v_hat = g^2 (0.001999 / 0.001999) (Kingma & Ba 2015, Sec. 2-3)."""
    rng = np.random.default_rng(0)
    g = rng.normal(size=64) * np.logspace(-3, 2, 64)
    p0 = rng.normal(size=64)
    lr = 1e-3
    p, m, v = p0.copy(), np.zeros(64), np.zeros(64)
    for t in range(1, 6):
        p1, m, v = U.adam_step(p, g, m, v, t, lr)
        assert np.allclose(p1 - p, -lr * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-18), t
        p = p1
    p, m, v = U.adam_step(p0, g, np.zeros(64), np.zeros(64), 1, lr)
    p2, _, _ = U.adam_step(p, -g, m, v, 2, lr)
    mh = -g * 0.01 / (1.0 - 0.81)
    vh = g * g * (0.999 * 0.001 + 0.001) / (1.0 - 0.999 ** 2)
    assert np.allclose(p2 - p, -lr * mh / (np.sqrt(vh) + 1e-8), rtol=1e-12)
    # the step is bounded by lr (here about 0.0526 lr / 1.0)
    assert np.all(np.abs(p2 - p) <= lr)
    # the folded form is Algorithm 1 with eps divided by sqrt(1 - b2^t)
    m, v = np.zeros(64), np.zeros(64)
    ma, va = np.zeros(64), np.zeros(64)
    pa = pb = p0
    for t in range(1, 4):
        gt = g * (1.0 + 0.3 * t)
        pa, ma, va = U.adam_step(pa, gt, ma, va, t, lr, folded=True)
        pb, m, v = U.adam_step(pb, gt, m, v, t, lr, eps=1e-8 / np.sqrt(1.0 - 0.999 ** t))
        assert np.allclose(pa, pb, rtol=1e-13, atol=1e-16), t
