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
