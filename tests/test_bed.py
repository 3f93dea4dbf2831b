"""The bed of particles ("reference mode" of the bed field).

CPU: the reference's own tests of init_bed_particles and BedConfig
(bed.rs:326-523) run against cs_bed_init and cs_bed_config_seated_in_filter;
the oracle's per-substep stamp (bed_lookup_clear / bed_lookup_scatter,
shader.rs:2291-2335) is checked against a brute-force numpy stamp, and a
reference-mode oracle step against a field-mode oracle step handed the same
lookup and the records that the BedExtracts and particles give.

GPU: the device stamp, bed table and the whole substep with a seated bed and
water soaking into it -- including bed_coupling, extraction_advect and the
grains' own motion (bed_dynamics) -- equal the oracle bit for bit.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import coffee_sim_amd as cs
from helpers import assert_bitwise, compare_state

BOUNDS = (14.0, 20.0, 14.0)


def small_config():
    b = cs.bed_config_default()
    b.num_particles = 200
    return b


def inner_radius_at_y(f, y):
    height = max(f.top_y - f.bot_y, 1e-6)
    t = min(max((y - f.bot_y) / height, 0.0), 1.0)
    return max(f.bot_radius + (f.top_radius - f.bot_radius) * t - f.thickness, f.hole_radius)


# ---- bed.rs tests -----------------------------------------------------------------------

def test_empty_bed_when_height_collapses():
    cfg = small_config()
    cfg.top_y = 0.0
    cfg.bot_y = 0.0
    b = cs.bed_init(cfg, (16, 16, 16), (8.0, 8.0, 8.0))
    assert len(b["particles"]) == len(b["affine"]) == len(b["extracts"]) == 0
    assert b["cell_lookup"].shape == (16 * 16 * 16,)
    assert (b["cell_lookup"] == -1).all()


def test_produces_at_most_target_particles():
    cfg = small_config()
    b = cs.bed_init(cfg, (32, 32, 32), BOUNDS)
    assert 0 < len(b["particles"]) <= cfg.num_particles
    assert len(b["particles"]) == len(b["affine"]) == len(b["extracts"])


def test_downsampled_bed_spans_configured_height():
    cfg = cs.bed_config_default()
    b = cs.bed_init(cfg, (32, 32, 32), BOUNDS)
    y = b["particles"][:, 1]
    assert y.min() <= cfg.center[1] + cfg.bot_y + 0.2
    assert y.max() >= cfg.center[1] + cfg.top_y - 0.2


def test_seated_bed_bottom_is_rounded_not_planar():
    cfg = cs.bed_config_seated_in_filter()
    p = cs.bed_init(cfg, (64, 64, 64), BOUNDS)["particles"]
    bottom = cfg.center[1] + cfg.bot_y
    low = p[p[:, 1] <= bottom + 0.45]
    r = np.sqrt(low[:, 0] ** 2 + low[:, 2] ** 2)
    centre = low[r < cfg.bot_radius * 0.45, 1]
    edge = low[r > cfg.bot_radius * 0.80, 1]
    assert len(centre) > 0 and len(edge) > 0
    assert edge.mean() > centre.mean() + 0.03


def test_bed_particles_carry_bed_phase_marker():
    b = cs.bed_init(small_config(), (32, 32, 32), BOUNDS)
    assert (b["affine"][:, 3] == 1.0).all()
    # the rest position rides in col1 (shader.rs:1999) and col2.x
    assert_bitwise("rest position", b["affine"][:, 4:7], b["particles"][:, 0:3])
    assert_bitwise("rest y", b["affine"][:, 8], b["particles"][:, 1])
    assert (b["particles"][:, 3] == 1.0).all() and (b["particles"][:, 7] == 1.0).all()
    assert (b["particles"][:, 4:7] == 0.0).all()


def test_bed_extracts_initialised_within_bounds():
    cfg = small_config()
    e = cs.bed_init(cfg, (32, 32, 32), BOUNDS)["extracts"]
    assert (e[:, 0] == 0.0).all() and (e[:, 3] == 0.0).all() and (e[:, 5] == 0.0).all() and (e[:, 7] == 0.0).all()
    assert np.abs(e[:, 1] - cfg.initial_porosity).max() < 1e-6
    assert np.abs(e[:, 2] - cfg.initial_permeability).max() < 1e-6
    assert np.abs(e[:, 4] - cfg.extractable_mass).max() < 1e-6


def test_cell_lookup_only_indexes_existing_particles():
    cfg = small_config()
    b = cs.bed_init(cfg, (32, 32, 32), BOUNDS)
    lk, p = b["cell_lookup"], b["particles"]
    assert lk.max() < len(p) and lk.min() >= -1
    assert (lk >= 0).any()
    # each indexed cell's particle is a near lattice neighbour of the cell centre
    height = cfg.top_y - cfg.bot_y
    ravg = (cfg.top_radius + cfg.bot_radius) * 0.5
    vol = math.pi * ravg * ravg * height / 3.0 * (1 + cfg.bot_radius / ravg + (cfg.bot_radius / ravg) ** 2)
    spacing = (vol / cfg.num_particles) ** (1.0 / 3.0)
    dx = BOUNDS[0] / 32
    idx = np.nonzero(lk >= 0)[0]
    iz, iy, ix = idx // (32 * 32), (idx // 32) % 32, idx % 32
    centre = np.stack([-BOUNDS[0] / 2 + (ix + 0.5) * dx, -BOUNDS[1] / 2 + (iy + 0.5) * dx,
                       -BOUNDS[2] / 2 + (iz + 0.5) * dx], 1)
    d = np.linalg.norm(p[lk[idx], 0:3] - centre, axis=1)
    assert d.max() <= 2.5 * spacing


def test_default_bed_sits_inside_filter_interior():
    f = cs.filter_config_default()
    bed = cs.bed_config_seated_in_filter(f)
    fbot, ftop = f.center[1] + f.bot_y, f.center[1] + f.top_y
    top_abs, bot_abs = bed.center[1] + bed.top_y, bed.center[1] + bed.bot_y
    assert bed.top_radius < inner_radius_at_y(f, top_abs - f.center[1])
    assert bed.bot_radius < inner_radius_at_y(f, bot_abs - f.center[1])
    assert bot_abs > fbot
    assert bot_abs - fbot >= bed.bot_radius * 0.65
    assert top_abs < ftop


def test_seated_in_filter_does_not_panic_on_narrow_filter():
    f = cs.filter_config_default()
    f.top_y, f.bot_y, f.top_radius, f.bot_radius, f.thickness, f.hole_radius = 0.1, -0.1, 0.3, 0.2, 0.02, 0.05
    bed = cs.bed_config_seated_in_filter(f)
    for v in (bed.top_y, bed.bot_y, bed.top_radius, bed.bot_radius):
        assert math.isfinite(v)


def test_seated_in_filter_uses_filter_center_offset():
    f = cs.filter_config_default()
    f.center[1] = 2.0
    bed = cs.bed_config_seated_in_filter(f)
    assert bed.center[1] + bed.top_y <= f.center[1] + f.top_y - 0.3
    assert bed.center[1] + bed.bot_y >= f.center[1] + f.bot_y + 0.3


def test_default_permeability_is_kozeny_carman():
    """kozeny_carman_permeability_m2(450 um, 0.40) (brew_config.rs:92-96): d^2 phi^3 / (180 (1 - phi)^2)."""
    b = cs.bed_config_default()
    d, phi = 450e-6, 0.40
    want = d * d * phi ** 3 / (180.0 * (1.0 - phi) ** 2)
    assert abs(b.initial_permeability - want) <= 1e-6 * want
    assert b.num_particles == 12000 and abs(b.initial_porosity - 0.40) < 1e-7


# ---- the scene: the default V60 with its seated bed and water soaking into it ------------

def bed_scene(seed: int = 0, n_water: int = 20000, substeps: int = 2):
    s = cs.preset("default_v60")
    s.substeps = substeps
    bed = cs.bed_init(cs.bed_config_seated_in_filter(), list(s.grid_dims), list(s.bounds_size))
    cfg = cs.bed_config_seated_in_filter()
    rng = np.random.default_rng(seed)
    top = cfg.center[1] + cfg.top_y
    r = cfg.top_radius * 0.85 * np.sqrt(rng.random(n_water))
    th = rng.random(n_water) * 2 * np.pi
    part = np.zeros((n_water, 8), np.float32)
    part[:, 0] = r * np.cos(th)
    part[:, 2] = r * np.sin(th)
    part[:, 1] = top - 1.2 + 1.8 * rng.random(n_water)
    part[:, 3] = 1.0
    part[:, 5] = -0.5
    part[:, 7] = 0.25
    aff = np.zeros((n_water, 12), np.float32)
    sc = cs.Scene()
    keep = [part, aff]
    sc.particles = part.ctypes.data_as(C.POINTER(C.c_float))
    sc.affine = aff.ctypes.data_as(C.POINTER(C.c_float))
    sc.num_particles = n_water
    sc._keep = keep
    cs.attach_bed(sc, bed)
    return s, sc, bed


def brute_force_stamp(s, bed):
    """bed_lookup_scatter by brute force: the highest bed index whose
    world_to_cell lies within one cell of each cell."""
    gx, gy, gz = (int(d) for d in s.grid_dims)
    dx = np.float32(np.float32(s.bounds_size[0]) / np.float32(gx))
    inv_dx = np.float32(np.float32(1.0) / dx)
    origin = (-np.asarray(list(s.bounds_size), np.float32) * np.float32(0.5)).astype(np.float32)
    p = bed["particles"][:, 0:3].astype(np.float32)
    cell = np.floor(((p - origin).astype(np.float32) * inv_dx).astype(np.float32)).astype(np.int64)
    look = np.full((gz, gy, gx), -1, np.int64)
    keep = bed["affine"][:, 3] >= 0.5
    for bid in np.nonzero(keep)[0]:
        cx, cy, cz = cell[bid]
        x0, x1 = max(cx - 1, 0), min(cx + 1, gx - 1)
        y0, y1 = max(cy - 1, 0), min(cy + 1, gy - 1)
        z0, z1 = max(cz - 1, 0), min(cz + 1, gz - 1)
        if x0 > x1 or y0 > y1 or z0 > z1:
            continue
        blk = look[z0:z1 + 1, y0:y1 + 1, x0:x1 + 1]
        np.maximum(blk, bid, out=blk)
    return look.reshape(-1).astype(np.int32)


def test_oracle_render_packs_the_bed_first(orc):
    """prepare_render over the reference's particle order (sh:2242-2276): the
    bed particles first, at radius 0.62 dx with colour -1 - saturation, then
    the water."""
    s, sc, bed = bed_scene(substeps=1, n_water=500)
    bed["extracts"][:, 7] = np.linspace(0.0, 1.0, len(bed["extracts"]), dtype=np.float32)
    cs.attach_bed(sc, bed)
    ref = orc.OracleSim(s, sc)
    out = ref.pack_render()
    nb = len(bed["particles"])
    assert out.shape == (nb + 500, 8)
    dx = np.float32(np.float32(s.bounds_size[0]) / np.float32(s.grid_dims[0]))
    assert_bitwise("bed positions", out[:nb, 0:3], bed["particles"][:, 0:3])
    assert_bitwise("bed colour", out[:nb, 3], (np.float32(-1.0) - bed["extracts"][:, 7]).astype(np.float32))
    assert (out[:nb, 4] == np.float32(dx * np.float32(0.62))).all()
    assert (out[nb:, 3] >= 0.0).all() and (out[nb:, 3] <= 2.0).all()


def test_oracle_bed_coupling_conserves_water(orc):
    """bed_coupling moves water mass from the particles into pore water and
    extraction_advect credits it (shader.rs:1885-1987): over one substep of a
    pour onto the seated bed, the mass the particles lose after G2P equals the
    pore water gained, short only of the fixed-point truncation of each credit
    (at most 7 credits of 2^-18 per absorbing particle); extractable plus
    dissolved mass stays 0.15 per bed particle; saturation is pore water over
    the per-particle capacity 42 ml * 80 / n_bed, clamped to 1."""
    s, sc, bed = bed_scene(substeps=1)
    s.emitter_mode = cs.abi.CS_EMIT_FIXED_COUNT  # no emission: the stopped copy emits nothing
    s.emit_per_substep = 0
    sub_dt = 1.0 / 600.0
    a, b = orc.OracleSim(s, sc), orc.OracleSim(s, sc)
    for k in range(3):  # both advance together; a stops before the coupling in the last substep
        if k < 2:
            a.step_frame(sub_dt)
        b.step_frame(sub_dt)
    e0 = a.read_bed_extracts()
    a.debug_step_until(sub_dt, "g2p")
    m_after_g2p = a.read_particles()[:, 7].astype(np.float64)
    assert_bitwise("state before coupling", a.read_bed_extracts(), e0)
    e1 = b.read_bed_extracts()
    m_after = b.read_particles()[:, 7].astype(np.float64)
    lost = m_after_g2p - m_after
    absorbing = int((lost > 0).sum())
    gained = float((e1[:, 0].astype(np.float64) - e0[:, 0]).sum())
    assert absorbing > 50 and (lost >= 0).all()
    slack = absorbing * 7 * 2.0 ** -18
    assert 0.0 <= lost.sum() - gained <= slack + 1e-6 * lost.sum(), (lost.sum(), gained, slack)
    n = len(e1)
    cap = np.float32(np.float32(42.0 * 80.0) / np.float32(n))
    wet = e1[:, 0] > 0
    want_sat = np.clip(e1[wet, 0] / cap, 0.0, 1.0).astype(np.float32)
    assert np.abs(e1[wet, 7] - want_sat).max() <= 1e-6
    assert np.abs(e1[:, 4].astype(np.float64) + e1[:, 5] - 0.15).max() <= 1e-6
    assert (e1[wet, 5] > 0).any()  # extraction started where the bed is wet


def test_oracle_stamp_matches_brute_force(orc):
    s, sc, bed = bed_scene(substeps=1)
    ref = orc.OracleSim(s, sc)
    ref.step_frame(1.0 / 600.0)
    got = ref.read_bed_lookup()
    want = brute_force_stamp(s, bed)
    assert (want >= 0).sum() > 1000
    assert_bitwise("bed lookup", got, want)


def test_oracle_reference_mode_equals_field_mode_with_the_same_records(orc):
    """Up to G2P, a bed of particles is a field bed whose lookup is the stamp
    and whose records are the BedExtracts' porosity, permeability, compaction
    and saturation with the particles' height and velocity (bed_coupling,
    which only particles have, follows G2P)."""
    s, sc, bed = bed_scene(substeps=1)
    ref = orc.OracleSim(s, sc)
    ref.debug_step_until(1.0 / 600.0, "g2p")
    e, p = bed["extracts"], bed["particles"]
    rec = np.zeros((len(p), 8), np.float32)
    rec[:, 0], rec[:, 1], rec[:, 2], rec[:, 3] = e[:, 1], e[:, 2], e[:, 3], e[:, 7]
    rec[:, 4] = p[:, 1]
    rec[:, 5:8] = p[:, 4:7]
    look = brute_force_stamp(s, bed)
    s2 = s.copy()
    s2.has_bed_field = 1
    sc2 = cs.Scene()
    sc2.particles, sc2.affine, sc2.num_particles = sc.particles, sc.affine, sc.num_particles
    sc2.bed_lookup = look.ctypes.data_as(C.POINTER(C.c_int32))
    sc2.bed_records = rec.ctypes.data_as(C.POINTER(cs.BedRecord))
    sc2.num_bed_records = len(rec)
    field = orc.OracleSim(s2, sc2)
    field.debug_step_until(1.0 / 600.0, "g2p")
    for name in ("read_particles", "read_affine", "read_grid_vel", "read_p2g_lanes"):
        assert_bitwise(name, getattr(ref, name)(), getattr(field, name)())
    pr, kr = ref.read_pressure()
    pf, kf = field.read_pressure()
    assert_bitwise("pressure", pr, pf)
    assert_bitwise("kind", kr, kf)
    assert (kr == 3).sum() > 100  # bed-coupled cells took part


# ---- GPU ---------------------------------------------------------------------------------

@pytest.mark.gpu
def test_bed_particles_match_oracle(ctx, orc):
    """The device stamp and every substep of a pour soaking into the seated
    bed equal the oracle bit for bit (particles, lanes, velocities, pressure,
    kinds, metrics), with the emitter's capacity counting the bed."""
    s, sc, bed = bed_scene(substeps=2)
    sim, ref = cs.MpmSim3D(ctx, s, sc), orc.OracleSim(s, sc)
    assert sim.counts() == ref.counts() == (sc.num_particles, len(bed["particles"]), s.max_particles)
    for k in range(2):
        sim.step_frame(1.0 / 300.0)
        ref.step_frame(1.0 / 300.0)
        assert_bitwise(f"bed lookup frame {k}", sim.read_bed_lookup(), ref.read_bed_lookup())
        compare_state(sim, ref, f"seated bed frame {k}")
        assert_bitwise(f"bed extracts frame {k}", sim.read_bed_extracts(), ref.read_bed_extracts())
        for name, g, o in zip(("bed grains", "bed affine"), sim.read_bed_particles(), ref.read_bed_particles()):
            assert_bitwise(f"{name} frame {k}", g, o)
    assert (sim.read_bed_extracts()[:, 0] > 0).sum() > 100  # the bed took up water
    moved = np.abs(sim.read_bed_particles()[0][:, 0:3] - bed["particles"][:, 0:3]).max(axis=1)
    assert (moved > 0).sum() > 100  # bed_dynamics moved the grains
    # prepare_render: bed particles first, radius 0.62 dx, colour -1 - saturation
    rg, ro = sim.pack_render(), ref.pack_render()
    assert rg.shape == (sim.counts()[0] + len(bed["particles"]), 8) == (sim.particle_count(), 8)
    assert_bitwise("render", rg, ro)
    nb = len(bed["particles"])
    sat = sim.read_bed_extracts()[:, 7]
    assert_bitwise("bed colour", rg[:nb, 3], (np.float32(-1.0) - sat).astype(np.float32))
    assert (sat > 0).any() and (rg[nb:, 3] >= 0.0).all()
    _, kind = sim.read_pressure()
    assert (kind == 3).sum() > 100


@pytest.mark.gpu
def test_bed_grains_follow_the_oracle_over_a_pour(ctx, orc):
    """Six frames of a pour onto the seated bed: the grains' positions,
    velocities, phases and rest positions, the extracts and the water equal the
    oracle bit for bit at every frame, and a reset restores the scene's bed."""
    s, sc, bed = bed_scene(seed=3, n_water=30000, substeps=2)
    sim, ref = cs.MpmSim3D(ctx, s, sc), orc.OracleSim(s, sc)
    for k in range(6):
        sim.step_frame(1.0 / 200.0)
        ref.step_frame(1.0 / 200.0)
        for name, g, o in zip(("bed grains", "bed affine"), sim.read_bed_particles(), ref.read_bed_particles()):
            assert_bitwise(f"{name} frame {k}", g, o)
        assert_bitwise(f"bed extracts frame {k}", sim.read_bed_extracts(), ref.read_bed_extracts())
        compare_state(sim, ref, f"pour frame {k}")
    part, aff = sim.read_bed_particles()
    assert np.abs(part[:, 0:3] - bed["particles"][:, 0:3]).max() > 1e-4
    sim.reset()
    part, aff = sim.read_bed_particles()
    assert_bitwise("reset grains", part, bed["particles"])
    assert_bitwise("reset affine", aff, bed["affine"])
    assert_bitwise("reset extracts", sim.read_bed_extracts(), bed["extracts"])


@pytest.mark.gpu
def test_bed_particles_are_validated(ctx):
    s, sc, bed = bed_scene(substeps=1, n_water=100)
    s2 = s.copy()
    s2.has_bed_field = 1
    with pytest.raises(cs.CoffeeSimError):
        cs.MpmSim3D(ctx, s2, sc)
    s3 = s.copy()
    s3.max_particles = len(bed["particles"]) + 50
    with pytest.raises(cs.CoffeeSimError):
        cs.MpmSim3D(ctx, s3, sc)
    sim = cs.MpmSim3D(ctx, s, sc)
    with pytest.raises(cs.CoffeeSimError):
        cs.Batch(ctx, [sim])
    plain = cs.MpmSim3D(ctx, s, None)
    plain.step_frame(1.0 / 600.0)
    assert (plain.read_bed_lookup() == -1).all()


@pytest.mark.gpu
def test_seated_bed_checkpoint_resumes_bitwise(ctx):
    """A checkpoint carries the bed particles, their affine state and their
    BedExtracts: saved after a frame and loaded into a fresh simulation, the
    next frame equals the uninterrupted run bit for bit."""
    s, sc, bed = bed_scene(substeps=2)
    a = cs.MpmSim3D(ctx, s, sc)
    a.step_frame(1.0 / 300.0)
    blob = a.save()
    b = cs.MpmSim3D(ctx, s, sc)
    b.step_frame(1.0 / 600.0)  # a different bed state, overwritten by the load
    b.load(blob)
    a.step_frame(1.0 / 300.0)
    b.step_frame(1.0 / 300.0)
    compare_state(a, b, "resumed seated bed")
    assert_bitwise("bed extracts", a.read_bed_extracts(), b.read_bed_extracts())
    for name, g, o in zip(("bed grains", "bed affine"), a.read_bed_particles(), b.read_bed_particles()):
        assert_bitwise(name, g, o)
    # a checkpoint of another bed (none) is rejected
    plain = cs.MpmSim3D(ctx, s, None)
    with pytest.raises(cs.CoffeeSimError):
        plain.load(blob)


@pytest.mark.gpu
def test_water_count_leaves_room_for_the_bed(ctx):
    """The water may use max_particles less the bed (mod.rs:239-249); a larger
    count is a range error, returned before any device work."""
    s, sc, bed = bed_scene(substeps=1, n_water=100)
    sim = cs.MpmSim3D(ctx, s, sc)
    nb = len(bed["particles"])
    with pytest.raises(cs.CoffeeSimError):
        sim.set_particle_count(s.max_particles)
    with pytest.raises(cs.CoffeeSimError):
        sim.set_particle_count(s.max_particles - nb + 1)
    sim.set_particle_count(s.max_particles - nb)
