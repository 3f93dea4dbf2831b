"""Known-answer tests ported from the reference's own unit tests, run against
the CPU oracle's restatement and against the library's host code (presets,
SDF and class generation).  Each test names the reference test it ports."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import coffee_sim_amd as cs
from coffee_sim_amd import abi


def _f32(x):
    return float(np.float32(x))


@pytest.fixture(scope="module")
def supm():
    s = cs.preset("default_v60")
    # gravity = -9.80665 * SIM_UNITS_PER_METER (units.rs:23-24)
    return float(np.float32(-s.gravity) / np.float32(9.80665))


# ---- units.rs:44-67 ---------------------------------------------------------------
def test_v60_height_sets_scene_length_scale(supm):
    modeled_height_m = np.float32(2.75 - (-3.02)) * np.float32(1.0 / supm)
    assert abs(modeled_height_m - 0.10) < 1e-6


def test_volume_scale_matches_length_scale():
    s = cs.preset("default_v60")
    assert abs(s.spout.volume_to_ml - 5.20) < 0.05  # ML_PER_SIM_UNIT_CUBED


def test_gentle_pour_speed_matches_kettle_scale(supm):
    s = cs.preset("default_v60")
    assert abs(s.spout.gentle_exit_speed / supm - 0.12) < 1e-6
    assert s.spout.max_exit_speed > s.spout.gentle_exit_speed


def test_current_filter_opening_is_larger_than_reference_paper(supm):
    assert 2.0 * 4.10 / supm > 0.10


# ---- brew_config.rs:170-196 ------------------------------------------------------
def test_kozeny_carman_permeability_tracks_grind_size_squared(orc):
    L = orc.lib()
    fine = L.orc_kozeny_carman_permeability(400.0, 0.40)
    coarse = L.orc_kozeny_carman_permeability(800.0, 0.40)
    assert fine > 0.0 and coarse > fine
    assert abs(coarse / fine - 4.0) < 0.01


def test_default_brew_permeability_is_above_minimum(orc):
    k = orc.lib().orc_kozeny_carman_permeability(450.0, 0.40)
    assert k > 1.0e-12
    assert 1.0e-6 / k > 0.0  # darcy_resistance_rate_s


# ---- inflow.rs:299-398 -------------------------------------------------------------
def _spout():
    return cs.preset("default_v60").spout


def test_effective_head_is_zero_below_activation_and_monotonic(orc):
    L, sp = orc.lib(), _spout()
    h = lambda a: L.orc_effective_head_from_angle(a, C.byref(sp))  # noqa: E731
    assert h(0.0) == 0.0
    assert h(sp.activation_angle_deg) == 0.0
    low, mid, high = h(16.0), h(30.0), h(48.0)
    assert 0.0 < low < mid < high
    assert abs(h(sp.full_flow_angle_deg) - sp.head_at_full_angle) < 1e-6


def test_exit_speed_from_head_is_monotonic_and_capped(orc):
    L = orc.lib()
    low, mid, high = (L.orc_exit_speed_from_head(h, 22.0) for h in (0.25, 2.0, 40.0))
    assert 0.0 < low < mid <= high <= 22.0
    assert L.orc_exit_speed_from_head(0.0, 22.0) == 0.0


def test_exit_speed_cap_tracks_pour_angle(orc):
    L, sp = orc.lib(), _spout()
    cap = lambda a: L.orc_exit_speed_cap_from_angle(a, C.byref(sp))  # noqa: E731
    barely, medium, full = cap(9.0), cap(36.0), cap(sp.full_flow_angle_deg)
    assert sp.gentle_exit_speed <= barely <= sp.gentle_exit_speed * 1.01
    assert medium > barely and full > medium
    assert abs(full - sp.max_exit_speed) < 1e-6


def test_flow_rate_from_speed_is_monotonic_and_capped(orc):
    L = orc.lib()
    r = 0.18
    low, mid, high = (L.orc_flow_rate_from_speed(v, r, 0.92, 5.4, 4.0) for v in (4.0, 10.0, 40.0))
    assert 0.0 < low < mid <= high <= 4.0
    assert L.orc_flow_rate_from_speed(0.0, r, 0.92, 5.4, 4.0) == 0.0


def test_translating_spout_preserves_jet_direction(orc):
    L, sp = orc.lib(), _spout()
    L.orc_spout_aim_at(C.byref(sp), 0.0, 0.4, 0.0)
    o, t, d = list(sp.origin), list(sp.target), list(sp.direction)
    L.orc_spout_translate_origin_to(C.byref(sp), o[0] - 0.3, o[1], o[2])
    assert abs(sp.target[0] - t[0] + 0.3) < 1e-5
    assert abs(sp.target[1] - t[1]) < 1e-5 and abs(sp.target[2] - t[2]) < 1e-5
    for c in range(3):
        assert abs(sp.direction[c] - d[c]) < 1e-5


# ---- mod.rs:695-797 -------------------------------------------------------------
def test_default_v60_shader_constants_in_sync():
    s = cs.preset("default_v60")
    cone = s.obstacles[0]
    assert cone.kind == abi.CS_OBSTACLE_TRUNCATED_CONE
    slope = (4.10 - 0.0) / (2.75 - -3.02)
    assert abs(cone.top_radius - (0.42 + slope * (cone.top_y - cone.bot_y))) < 0.01
    assert abs(cone.bot_radius - 0.42) < 1e-6
    assert (cone.top_y, cone.bot_y) == (3.0, -3.0)
    cup = s.obstacles[1]
    assert (cup.radius, cup.top_y, cup.bot_y) == (3.0, -3.5, -8.0)
    # the analytic WGSL fallback (shader.rs:101, 751-756) matches the cone
    assert abs(s.geometry.cone_top_radius - _f32(_f32(0.42) + _f32(4.2634315))) < 1e-6
    assert abs((s.geometry.cone_top_radius - s.geometry.cone_bot_radius) / 6.0 - slope) < 1e-5


def test_default_v60_grid_uses_uniform_dx():
    s = cs.preset("default_v60")
    dx = s.bounds_size[0] / s.grid_dims[0]
    dz = s.bounds_size[2] / s.grid_dims[2]
    assert abs(dx - dz) < 1e-5
    assert s.grid_dims[1] * dx >= s.bounds_size[1] - dx
    assert list(s.grid_dims) == [80, 115, 80]


def test_default_v60_uses_slow_gooseneck_angle(orc, supm):
    s = cs.preset("default_v60")
    assert s.initial_kettle_angle_deg <= 10.0
    assert s.spout.max_flow_rate_ml_s <= 12.0
    L = orc.lib()
    head = L.orc_effective_head_from_angle(s.initial_kettle_angle_deg, C.byref(s.spout))
    cap = L.orc_exit_speed_cap_from_angle(s.initial_kettle_angle_deg, C.byref(s.spout))
    assert L.orc_exit_speed_from_head(head, cap) / supm <= 0.13
    assert s.spout.max_exit_speed / supm <= 0.50


def test_occupancy_threshold_accepts_lone_particle():
    nominal = 80.0 / 320.0
    assert nominal * 0.1 < nominal * 0.75 ** 3


def test_grid_lanes_fit_storage_binding_limit():
    # expanded_grid_atomic_lanes_fit_storage_binding_limit (mod.rs:766-784)
    s = cs.preset("default_v60")
    assert 6 * s.num_cells * 4 <= 128 * 1024 * 1024  # wgpu default max_storage_buffer_binding_size


# ---- presets and scene generators: library host code == oracle restatement ----------
@pytest.mark.parametrize("name", ["default_v60", "benchmark_free_stream", "benchmark_center_pour"])
def test_presets_match_oracle_bit_for_bit(orc, name):
    assert bytes(cs.preset(name)) == bytes(orc.preset(name))


def test_center_pour_and_free_stream_aim(orc):
    fs = cs.preset("benchmark_free_stream")
    assert list(fs.spout.origin) == [0.0, _f32(6.8), 0.0] and list(fs.spout.direction) == [0.0, -1.0, 0.0]
    cp = cs.preset("benchmark_center_pour")
    assert list(cp.spout.origin) == [0.0, _f32(7.1), 0.0]
    assert list(cp.spout.target) == [0.0, _f32(0.4), 0.0]


def test_sdf_and_class_generation_match_oracle(orc):
    s = cs.preset("default_v60")
    a, b = cs.generate_sdf(s), orc.generate_sdf(s)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    ca, cb = cs.generate_sdf_class(s, a), orc.generate_sdf_class(s, b)
    assert np.array_equal(ca, cb)


def _sdf_closed_form(s, px, py, pz):
    """The obstacle SDFs of state.rs:206-276 in float64.  Union rule of
    state.rs:264-270: a texel still at the "no constraint" value (>= 998) takes
    the obstacle's value, otherwise the max; an obstacle that does not
    constrain a point contributes 999 - wall/2."""
    half = 0.5 * s.wall_thickness
    out = np.full(px.shape, 999.0)
    for o in list(s.obstacles)[: s.num_obstacles]:
        r = np.hypot(px - o.center[0], pz - o.center[2])
        ly = py - o.center[1]
        if o.kind == abi.CS_OBSTACLE_TRUNCATED_CONE:
            t = (ly - o.bot_y) / (o.top_y - o.bot_y)
            d = o.bot_radius + (o.top_radius - o.bot_radius) * t - r
            d = np.where((ly > o.top_y) | (ly < o.bot_y), 999.0, d)
        else:
            d = np.minimum(o.radius - r, ly - o.bot_y)
            d = np.where(ly > o.top_y, 999.0, d)
        d = d - half
        out = np.where(out >= 998.0, d, np.maximum(out, d))
    return out


def test_sdf_matches_closed_form_obstacles(orc):
    s = cs.preset("default_v60")
    n = s.sdf_res
    sdf = orc.generate_sdf(s).reshape(n, n, n).astype(np.float64)
    i = (np.arange(n) + 0.5) / n
    z, y, x = np.meshgrid(i, i, i, indexing="ij")
    px = -s.bounds_size[0] / 2 + x * s.bounds_size[0]
    py = -s.bounds_size[1] / 2 + y * s.bounds_size[1]
    pz = -s.bounds_size[2] / 2 + z * s.bounds_size[2]
    want = _sdf_closed_form(s, px, py, pz)
    assert np.max(np.abs(sdf - want)) < 2e-5


def test_sdf_class_matches_live_cell_center_sampling(orc):
    """sdf_class_data_matches_live_cell_center_sampling (state.rs:441-468), with
    the live sample taken in float64 trilinear interpolation."""
    s = cs.preset("default_v60")
    n = s.sdf_res
    sdf = orc.generate_sdf(s)
    cls = orc.generate_sdf_class(s, sdf).reshape(s.grid_dims[2], s.grid_dims[1], s.grid_dims[0])
    d = sdf.reshape(n, n, n).astype(np.float64)
    gx, gy, gz = s.grid_dims
    dx = s.bounds_size[0] / gx
    iz, iy, ix = np.meshgrid(np.arange(gz), np.arange(gy), np.arange(gx), indexing="ij")
    pos = [(-s.bounds_size[c] / 2 + (g + 0.5) * dx) for c, g in enumerate((ix, iy, iz))]
    uv = [(pos[c] + s.bounds_size[c] / 2) / s.bounds_size[c] * n - 0.5 for c in range(3)]
    b = [np.floor(u).astype(int) for u in uv]
    f = [u - np.floor(u) for u in uv]
    val = np.zeros(ix.shape)
    for corner in range(8):
        o = [(corner >> c) & 1 for c in range(3)]
        idx = [np.clip(b[c] + o[c], 0, n - 1) for c in range(3)]
        w = np.ones(ix.shape)
        for c in range(3):
            w = w * (f[c] if o[c] else 1.0 - f[c])
        val += w * d[idx[2], idx[1], idx[0]]
    live = val < 0.0
    mismatch = (live != (cls != 0)) & (np.abs(val) > 1e-5)
    assert not mismatch.any()
    assert 0.05 < cls.mean() < 0.95
