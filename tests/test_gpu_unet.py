"""The learned warm start on the device against its CPU oracle
(oracle/unet_ref.py), and its use inside the solve.

Activations and activation gradients are bf16 on both sides (rounded at the
same points); the device accumulates in fp32 where the oracle sums exactly,
so values that sit on a bf16 rounding boundary can land one bf16 step apart
and the comparisons allow a few per cent of relative L2 error.  The solve
that starts from the network's prediction is compared with the oracle bit for
bit (the prediction enters both as the same float field)."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

import coffee_sim_amd as cs
from helpers import assert_bitwise, compare_state

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import unet_ref as U  # noqa: E402

pytestmark = pytest.mark.gpu

N = 32


def _inputs(seed: int):
    rng = np.random.default_rng(seed)
    f4 = rng.normal(0.0, 1.0, size=(N, N, N, 4)).astype(np.float32)
    f4[..., 1] = np.clip(f4[..., 1], 0.0, 1.0)
    target = rng.normal(0.0, 1.0, size=(N, N, N)).astype(np.float32)
    mask = (rng.random((N, N, N)) < 0.6).astype(np.uint8)
    return f4, target, mask


def _oracle_params(seed: int) -> dict:
    rng = np.random.default_rng(seed + 1)
    p = U.init_params(seed)
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.normal(0.0, 0.05, size=p[k].shape)
    # the device holds fp32 parameters
    return {k: np.asarray(v, np.float32).astype(np.float64) for k, v in p.items()}


def _x16(f4):
    x = np.zeros(f4.shape[:3] + (U.IN_CH,))
    x[..., :4] = f4
    return x


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def net_and_oracle(ctx):
    p = _oracle_params(5)
    net = cs.UNet(ctx, N, seed=0)
    assert net.param_count == U.param_vector(p).size
    net.set_params(U.param_vector(p).astype(np.float32))
    yield net, p
    net.close()


def test_parameters_round_trip(net_and_oracle):
    net, p = net_and_oracle
    np.testing.assert_array_equal(net.get_params(), U.param_vector(p).astype(np.float32))


def test_forward_matches_oracle(net_and_oracle):
    net, p = net_and_oracle
    f4, _, _ = _inputs(1)
    got = net.forward(f4)
    want, _ = U.forward(p, _x16(f4))
    assert rel_l2(got, want) <= 2e-2, rel_l2(got, want)


def test_gradients_match_oracle(net_and_oracle):
    """Every parameter group's gradient (weights and bias of each convolution,
    the head) from the device's forward and backward kernels against the
    oracle's backward pass."""
    net, p = net_and_oracle
    f4, target, mask = _inputs(2)
    loss, grad = net.gradient(f4, target, mask)
    pred, s = U.forward(p, _x16(f4))
    want_loss, gp = U.loss_and_grad(pred, target.astype(np.float64), mask)
    assert abs(loss - want_loss) <= 2e-2 * want_loss
    want = U.param_dict(U.param_vector(U.backward(p, s, gp)))
    got = U.param_dict(grad.astype(np.float64))
    for k in want:
        err = rel_l2(got[k], want[k])
        assert err <= 5e-2, (k, err)


def test_training_lowers_the_loss(ctx):
    net = cs.UNet(ctx, N, seed=3)
    f4, target, mask = _inputs(4)
    target = (0.5 * f4[..., 0] + 0.2).astype(np.float32)  # a learnable map
    first = net.train_step(f4, target, mask, 1e-3)
    for _ in range(30):
        last = net.train_step(f4, target, mask, 1e-3)
    assert last < 0.5 * first, (first, last)
    net.close()


def test_device_adam_matches_the_oracle_adam(ctx):
    """Two train steps on the device against oracle/unet_ref.adam_step (its
    folded form, which the device implements: the bias corrections in the
    step size) applied to the device's own gradient at each step; the first
    step moves no parameter by more than lr."""
    net = cs.UNet(ctx, N, seed=7)
    f4, target, mask = _inputs(9)
    lr = 1e-3
    p = net.get_params().astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for t in (1, 2):
        _, g = net.gradient(f4, target, mask)
        g = g.astype(np.float64)
        want, m, v = U.adam_step(p, g, m, v, t, lr, folded=True)
        net.train_step(f4, target, mask, lr)
        got = net.get_params().astype(np.float64)
        big = np.abs(g) > 1e-6
        assert big.sum() > 1000
        d_got, d_want = got - p, want - p
        # float32 parameters and moments: a few ulps of p plus 1e-4 of the step
        assert np.all(np.abs(d_got[big] - d_want[big]) <= 1e-4 * lr + 4e-7 * np.abs(p[big])), t
        if t == 1:  # lr g / (|g| + eps) in the first step
            assert np.all(np.abs(d_got) <= lr * (1 + 1e-3) + 1e-7 * np.abs(p))
        p = got
    net.close()


def _learned_settings(w, fallback: float = 10.0):
    s = w.settings.copy()
    s.warm_start = cs.abi.CS_WARM_LEARNED
    s.warm_fallback_residual = fallback
    return s


@pytest.mark.parametrize("fallback", [1.0, 1e30], ids=["falls-back", "accepted"])
def test_learned_start_inside_the_solve_matches_the_host_path_and_the_oracle(ctx, orc, fallback):
    """v60_working at 1/4 size (32^3): the solve seeded on the device by the
    network equals (1) the network run on the captured features through the
    host API and (2) the oracle seeded with that field, bit for bit; once with
    the untrained prediction rejected (residual above 1) and once accepted."""
    w = cs.SeededWorkload("v60_working", seed=6, scale=4)
    net = cs.UNet(ctx, N, seed=1)
    net.set_scales(0.05, 500.0)
    sim = cs.MpmSim3D(ctx, _learned_settings(w, fallback), w.scene)
    sim.set_warm_net(net)
    ext = w.settings.copy()
    ext.warm_start = cs.abi.CS_WARM_EXTERNAL
    ext.warm_fallback_residual = fallback
    ref = orc.OracleSim(ext, w.scene)
    for k in range(2):
        sim.step_frame(w.frame_dt)
        field = sim.read_warm_start()
        net.capture(sim)
        f4, _, _ = net.read_sample(net.dataset_size() - 1)
        host = (net.forward(f4).reshape(-1) * np.float32(500.0)).astype(np.float32)
        assert_bitwise(f"warm field substep {k}", field, host)
        ref.set_warm_start(field)
        ref.step_frame(w.frame_dt)
        sg, so = sim.solve_stats(), ref.solve_stats()
        assert sg.iterations == so.iterations and sg.fell_back == so.fell_back, (sg.as_dict(), so.as_dict())
        assert sg.fell_back == (1 if fallback == 1.0 else 0)
        compare_state(sim, ref, f"learned start substep {k}")
    sim.close()
    net.close()


def test_trained_network_cuts_rbgs_iterations(ctx):
    """Pairs dumped by the plain solver from six seeded 1/4-size pours (96
    solves, each run to 1e-7 or 40,000 iterations so the target carries little
    of the solver's own smooth error) train the network on the device for 40
    epochs; on a held-out seed its prediction reaches the 1e-4 residual in at
    least 25% fewer RBGS iterations than the zero start (1.9x fewer when this
    was written).  A prediction is accepted while its initial residual is
    below 10: its excess over the zero start's residual of 1 is high-frequency
    error that the first sweeps remove."""
    net = cs.UNet(ctx, N, seed=2)
    train_seeds, held_out = (11, 12, 13, 14, 16, 17), 15
    # scales from the data: rms of the right-hand side and of the pressure
    def pair_settings(w):
        s = w.settings.copy()
        s.rbgs_tolerance = 1e-7
        s.rbgs_max_iters = 40000
        return s

    probe = cs.SeededWorkload("v60_working", seed=train_seeds[0], scale=4)
    sp = cs.MpmSim3D(ctx, pair_settings(probe), probe.scene)
    sp.step_frame(probe.frame_dt)
    net.capture(sp)
    f4, t, m = net.read_sample(0)
    sel = m.astype(bool)
    net.set_scales(1.0 / max(float(np.sqrt(np.mean(f4[sel, 0] ** 2))), 1e-6),
                   max(float(np.sqrt(np.mean(t[sel] ** 2))), 1e-6))
    net.clear_dataset()
    sp.close()
    for sd in train_seeds:
        wk = cs.SeededWorkload("v60_working", seed=sd, scale=4)
        sim = cs.MpmSim3D(ctx, pair_settings(wk), wk.scene)
        for _ in range(16):
            sim.step_frame(wk.frame_dt)
            net.capture(sim)
        sim.close()
    losses = [net.train_epoch(2e-3 * 0.2 ** (e / 39), seed=e) for e in range(40)]
    assert losses[-1] < 0.05 * losses[0], losses
    wk = cs.SeededWorkload("v60_working", seed=held_out, scale=4)
    plain = cs.MpmSim3D(ctx, wk.settings, wk.scene)
    learned = cs.MpmSim3D(ctx, _learned_settings(wk), wk.scene)
    learned.set_warm_net(net)
    zero_it, learned_it = [], []
    for _ in range(6):
        plain.step_frame(wk.frame_dt)
        learned.step_frame(wk.frame_dt)
        zero_it.append(plain.solve_stats().iterations)
        learned_it.append(learned.solve_stats().iterations)
        assert learned.solve_stats().converged
    print(f"zero-start iterations {zero_it}, learned-start iterations {learned_it}")
    assert sum(learned_it) <= 0.75 * sum(zero_it), (zero_it, learned_it)
