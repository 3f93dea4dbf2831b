"""The oracle's scalar helpers against what the mathematics fixes."""
from __future__ import annotations

import math

import numpy as np


def ulp_err(got: float, want: float) -> float:
    g = np.float32(got)
    return abs(float(g) - want) / float(np.spacing(np.float32(abs(want))))


def test_expf_within_two_ulp_of_exact(orc):
    L = orc.lib()
    xs = np.concatenate([
        np.linspace(-0.35, 0.35, 70001, dtype=np.float32),
        np.linspace(-87.0, 88.0, 50001, dtype=np.float32),
    ])
    worst = max(ulp_err(L.orc_expf(float(x)), math.exp(float(x))) for x in xs)
    assert worst <= 2.0, worst


def test_expf_special_values(orc):
    L = orc.lib()
    assert L.orc_expf(0.0) == 1.0
    assert math.isnan(L.orc_expf(float("nan")))
    assert L.orc_expf(-1000.0) > 0.0  # clamped, no underflow to 0 in the J update range
    assert abs(L.orc_expf(1.0) - math.e) < 2e-7


def test_smoothstep_closed_form(orc):
    L = orc.lib()
    assert L.orc_smoothstep(0.35, 0.85, 0.0) == 0.0
    assert L.orc_smoothstep(0.35, 0.85, 1.0) == 1.0
    assert abs(L.orc_smoothstep(0.35, 0.85, 0.6) - 0.5) < 1e-6
    t = (0.5 - 0.35) / 0.5
    assert abs(L.orc_smoothstep(0.35, 0.85, 0.5) - t * t * (3 - 2 * t)) < 1e-6


def test_float_to_int_truncates_saturates_and_maps_nan_to_zero(orc):
    L = orc.lib()
    assert L.orc_f2i(2.9) == 2 and L.orc_f2i(-2.9) == -2
    assert L.orc_f2i(3.0e9) == 2**31 - 1
    assert L.orc_f2i(-3.0e9) == -(2**31)
    assert L.orc_f2i(float("nan")) == 0
    assert L.orc_f2i(float("inf")) == 2**31 - 1


def _sincos(L, xs):
    import ctypes as C

    s, c = C.c_float(), C.c_float()
    out = np.empty((xs.size, 2), np.float32)
    for i, x in enumerate(xs):
        L.orc_sincosf(float(x), C.byref(s), C.byref(c))
        out[i] = (s.value, c.value)
    return out


def test_emitter_sincos_is_correctly_rounded_on_golden_angles(orc):
    """sin/cos of theta = sample * 2.399963 (the emitter's golden angle, up to
    2^28 samples) and of spread arguments: each result is the float nearest the
    exact value (math.sin/cos in double, whose own error is far below half a
    float ulp), except within 1e-12 of a rounding midpoint, and never off by
    more than one ulp.  A wrong quadrant, constant or coefficient breaks this."""
    L = orc.lib()
    rng = np.random.default_rng(3)
    samples = np.concatenate([np.arange(0, 4096), rng.integers(0, 1 << 28, 20000)]).astype(np.uint32)
    xs = np.concatenate([
        samples.astype(np.float32) * np.float32(2.399963),
        rng.uniform(-10.0, 10.0, 5000).astype(np.float32),
        np.array([0.0, -0.0, 1e-30, np.pi / 4, np.pi / 2, np.pi, 3 * np.pi / 2, 1e9], np.float32),
    ])
    got = _sincos(L, xs)
    exact = np.array([[math.sin(float(x)), math.cos(float(x))] for x in xs])
    want = exact.astype(np.float32)
    diff = got != want
    for i, j in zip(*np.nonzero(diff)):
        e = exact[i, j]
        lo, hi = sorted((float(got[i, j]), float(want[i, j])))
        assert np.nextafter(np.float32(lo), np.float32(np.inf)) == np.float32(hi), (xs[i], j)
        mid = (lo + hi) / 2.0
        assert abs(e - mid) <= 1e-12 * max(abs(e), 1e-30), (xs[i], j, e, got[i, j], want[i, j])
    assert diff.sum() <= 2, int(diff.sum())
