"""HIP path vs the CPU oracle on identical inputs, through the C ABI.

Integer and index work (P2G lanes, kinds, pressure, metrics) must match bit for
bit.  The floating-point passes use the same binary32 operation order on both
sides (no FMA contraction, correctly rounded / and sqrt, the same exp), so
particle and grid floats are required to match bit for bit as well.  The
stated tolerances are asserted on top: grid mass and momentum within 1e-5
relative L2 after one substep (`test_v60_small_one_substep_...`), positions
within 1e-4*dx after the 200 substeps of the small V60 workload, and the
divergence left after projection (metric slots 5 and 6) at or below the
oracle's (`test_v60_small_200_substeps_...`).
"""
from __future__ import annotations

import numpy as np
import pytest

import coffee_sim_amd as cs
from helpers import assert_bitwise, compare_state, rel_l2

pytestmark = pytest.mark.gpu


def _pair(ctx, orc, settings, scene=None):
    return cs.MpmSim3D(ctx, settings, scene), orc.OracleSim(settings, scene)


def test_device_exp_matches_oracle_exp(ctx, orc):
    x = np.concatenate([
        np.linspace(-0.35, 0.35, 20001, dtype=np.float32),
        np.linspace(-87.0, 88.0, 20001, dtype=np.float32),
        np.array([0.0, -0.0, 1e-30, -1e-30, 0.34999999], np.float32),
    ])
    got = ctx.device_expf(x)
    want = np.array([orc.lib().orc_expf(float(v)) for v in x], np.float32)
    assert_bitwise("expf", got, want)


def test_device_sincos_matches_oracle_sincos(ctx, orc):
    """The emitter's sin/cos (double-precision reduction and series on both
    sides) on golden angles of up to 2^28 samples and spread arguments."""
    import ctypes as C

    rng = np.random.default_rng(11)
    samples = np.concatenate([np.arange(0, 4096), rng.integers(0, 1 << 28, 20000)]).astype(np.uint32)
    x = np.concatenate([samples.astype(np.float32) * np.float32(2.399963),
                        rng.uniform(-10.0, 10.0, 4000).astype(np.float32)])
    sn, cs_ = ctx.device_sincosf(x)
    L = orc.lib()
    a, b = C.c_float(), C.c_float()
    want = np.empty((x.size, 2), np.float32)
    for i, v in enumerate(x):
        L.orc_sincosf(float(v), C.byref(a), C.byref(b))
        want[i] = (a.value, b.value)
    assert_bitwise("sin", sn, want[:, 0])
    assert_bitwise("cos", cs_, want[:, 1])


@pytest.mark.parametrize("name", ["default_v60", "benchmark_free_stream", "benchmark_center_pour", "pour96"])
def test_device_scene_generation_matches_oracle(ctx, orc, name):
    """A simulation created without scene arrays generates its SDF and class
    mask on the device (generate_sdf_data / generate_sdf_class_data,
    state.rs:206-344): equal bit for bit to the oracle's and to the library's
    host generator, for the presets and a 96^3 pour grid with a 96^3 field."""
    if name == "pour96":
        w = cs.SeededWorkload.pour(96, 20000, 200, 40.0, seed=5)
        s = w.settings.copy()
        s.sdf_res = 96
    else:
        s = cs.preset(name)
    sim = cs.MpmSim3D(ctx, s)
    sdf, cls = sim.read_sdf()
    sim.close()
    want_sdf = orc.generate_sdf(s)
    assert_bitwise("sdf", sdf, want_sdf)
    assert_bitwise("sdf host", sdf, cs.generate_sdf(s))
    assert_bitwise("class", cls, orc.generate_sdf_class(s, want_sdf))
    assert 0 < int(cls.sum()) < cls.size


def test_free_stream_reference_scene_bitwise(ctx, orc):
    """benchmark_free_stream (mod.rs:145-152): the whole reference substep with
    its own SDF, class mask, emitter and 40 fixed RBGS pairs."""
    s = cs.preset("benchmark_free_stream")
    sim, ref = _pair(ctx, orc, s)
    for a in (sim, ref):
        a.set_kettle_angle(36.0)
    for frame in range(6):
        sim.step_frame(1.0 / 60.0)
        ref.step_frame(1.0 / 60.0)
        compare_state(sim, ref, f"frame {frame}")
    assert sim.times() == ref.times()
    assert sim.inflow() == ref.inflow()
    st_g, st_o = sim.solve_stats(), ref.solve_stats()
    assert st_g.iterations == st_o.iterations == 40
    assert st_g.final_residual == pytest.approx(st_o.final_residual, rel=1e-9)


def test_v60_small_one_substep_grid_tolerance_and_bitwise(ctx, orc):
    w = cs.SeededWorkload("v60_small", seed=0)
    sim, ref = _pair(ctx, orc, w.settings, w.scene)
    sim.step_frame(w.frame_dt)
    ref.step_frame(w.frame_dt)
    lg, lo = sim.read_p2g_lanes(), ref.read_p2g_lanes()
    # grid mass and momentum after one substep within 1e-5 rel L2
    assert rel_l2(lg[0], lo[0]) <= 1e-5
    assert rel_l2(lg[1:4], lo[1:4]) <= 1e-5
    compare_state(sim, ref, "v60_small substep 1")


def test_v60_small_200_substeps_positions_and_post_projection_divergence(ctx, orc):
    """The small V60 workload in full: 64^3, dx 2 mm, 32,768 particles, 40
    fixed red+black pairs, 200 substeps of 1e-4 s.  Every 20 substeps: the
    post-projection max|div| and max|div - target| (slots 5, 6) are at or
    below the oracle's, and the whole state is bitwise equal; at the end the
    positions are within 1e-4*dx of the oracle's."""
    w = cs.SeededWorkload("v60_small", seed=0)
    assert w.substeps_total == 200 and w.settings.substeps == 1
    sim, ref = _pair(ctx, orc, w.settings, w.scene)
    dx = w.settings.bounds_size[0] / w.settings.grid_dims[0]
    for k in range(1, 201):
        sim.step_frame(w.frame_dt)
        ref.step_frame(w.frame_dt)
        if k % 20 == 0:
            mg, mo = sim.read_metrics(), ref.read_metrics()
            assert mg[5] <= mo[5] and mg[6] <= mo[6], (k, mg, mo)
            assert mg[1] > 0  # fluid cells
            compare_state(sim, ref, f"v60_small substep {k}")
    n = sim.counts()[0]
    pg = sim.read_particles(0, n)[:, 0:3].astype(np.float64)
    po = ref.read_particles(0, n)[:, 0:3].astype(np.float64)
    assert np.abs(pg - po).max() <= 1e-4 * dx
    assert np.all(np.isfinite(pg))


@pytest.mark.parametrize("variant", [
    {"CS_SOLVER": "dense"},
    {"CS_RBGS_HALO": "1"},
    {"CS_RBGS_HALO": "2", "CS_RBGS_THREADS": "256"},
    {"CS_RBGS_HALO": "2", "CS_RBGS_THREADS": "512"},
    {"CS_RBGS_HALO": "4", "CS_RBGS_MULTI": "0", "CS_RBGS_BATCH": "1"},
    {"CS_RBGS_HALO": "4", "CS_RBGS_MULTI": "1"},
    {"CS_RBGS_HALO": "4", "CS_RBGS_MULTI": "1", "CS_RBGS_GROUPS": "8"},
    {"CS_RBGS_HALO": "4", "CS_RBGS_MULTI": "1", "CS_RBGS_GROUPS": "3"},
    {"CS_RBGS_HALO": "4", "CS_RBGS_MULTI": "1", "CS_RBGS_FUSE": "0"},
    {"CS_RBGS_HALO": "4", "CS_RBGS_MULTI": "1", "CS_GRAPHS": "0"},
    {"CS_RBGS_HALO": "4", "CS_RBGS_MULTI": "1", "CS_RBGS_BATCH": "2"},
], ids=["dense", "brick-h1", "brick-h2-t256", "brick-h2-t512", "brick-h4-single-batch1", "brick-h4-multi",
        "brick-h4-multi-8groups", "brick-h4-multi-3groups", "brick-h4-multi-unfused", "brick-h4-multi-nographs",
        "brick-h4-multi-batch2"])
def test_solver_variants_match_oracle(ctx, orc, monkeypatch, variant):
    """Every pressure-solver layout (one thread per cell over the whole grid;
    brick-sparse with 1, 2 or 4 half sweeps per launch; the multi-brick launch
    with few workgroups, so each walks several bricks and prefetches; with and
    without the residual test folded into the sweeps and the steady-state
    batch replayed as a graph) gives the oracle's pressure, iteration count
    and particles bit for bit."""
    for k, v in variant.items():
        monkeypatch.setenv(k, v)
    w = cs.SeededWorkload("v60_working", seed=3, scale=4)
    sim, ref = _pair(ctx, orc, w.settings, w.scene)
    for k in range(2):
        sim.step_frame(w.frame_dt)
        ref.step_frame(w.frame_dt)
        assert sim.solve_stats().iterations == ref.solve_stats().iterations
        compare_state(sim, ref, f"{variant} substep {k}")
    # fixed-pairs mode through the same layout
    s = w.settings.copy()
    s.rbgs_mode = 0
    s.pressure_rbgs_pairs = 13
    sim2, ref2 = _pair(ctx, orc, s, w.scene)
    sim2.step_frame(w.frame_dt)
    ref2.step_frame(w.frame_dt)
    compare_state(sim2, ref2, f"{variant} fixed 13 pairs")


def test_v60_working_scaled_tolerance_solve(ctx, orc):
    """Residual-driven RBGS (tolerance 1e-4, checked every 8 iterations) with the
    value-noise bed and the fixed-count pour, at 1/4 linear size."""
    w = cs.SeededWorkload("v60_working", seed=1, scale=4)
    sim, ref = _pair(ctx, orc, w.settings, w.scene)
    for k in range(3):
        sim.step_frame(w.frame_dt)
        ref.step_frame(w.frame_dt)
        sg, so = sim.solve_stats(), ref.solve_stats()
        assert sg.iterations == so.iterations, (k, sg.as_dict(), so.as_dict())
        assert sg.converged == so.converged
        compare_state(sim, ref, f"v60_working/4 substep {k}")


def test_previous_step_warm_start_matches_oracle(ctx, orc):
    """Each solve seeded with the previous substep's converged pressure (fluid
    cells only) reaches the tolerance in the oracle's iteration count."""
    w = cs.SeededWorkload("v60_working", seed=2, scale=4)
    s = w.settings.copy()
    s.warm_start = cs.abi.CS_WARM_PREVIOUS
    sim, ref = _pair(ctx, orc, s, w.scene)
    iters = []
    for k in range(4):
        sim.step_frame(w.frame_dt)
        ref.step_frame(w.frame_dt)
        sg, so = sim.solve_stats(), ref.solve_stats()
        assert sg.iterations == so.iterations, (k, sg.as_dict(), so.as_dict())
        iters.append(sg.iterations)
        compare_state(sim, ref, f"previous warm start substep {k}")
    # the carried pressure pays off after the first substep
    assert min(iters[1:]) < iters[0]


def test_external_warm_start_and_fallback_match_oracle(ctx, orc):
    """A supplied pressure field seeds the solve; one worse than zero (residual
    above 1) is rejected and the solve restarts from zero, on both sides."""
    w = cs.SeededWorkload("v60_working", seed=4, scale=4)
    s = w.settings.copy()
    s.warm_start = cs.abi.CS_WARM_EXTERNAL
    nc = int(np.prod(list(s.grid_dims)))
    rng = np.random.default_rng(5)
    for label, field, fell_back in (("good", None, 0), ("bad", rng.normal(0, 50.0, nc).astype(np.float32), 1)):
        sim, ref = _pair(ctx, orc, s, w.scene)
        if field is None:
            # the zero-start solution of this very substep, as a perfect predictor would give it
            z = w.settings.copy()
            probe = cs.MpmSim3D(ctx, z, w.scene)
            probe.step_frame(w.frame_dt)
            p_fixed, _ = probe.read_pressure()
            field = (p_fixed.astype(np.float64) / cs.abi.FP_SCALE).astype(np.float32)
            probe.close()
        sim.set_warm_start(field)
        ref.set_warm_start(field)
        sim.step_frame(w.frame_dt)
        ref.step_frame(w.frame_dt)
        sg, so = sim.solve_stats(), ref.solve_stats()
        assert sg.iterations == so.iterations, (label, sg.as_dict(), so.as_dict())
        assert sg.fell_back == so.fell_back == fell_back, label
        compare_state(sim, ref, f"external warm start ({label})")


@pytest.mark.parametrize("variant", [
    {},
    {"CS_RBGS_FUSE": "0"},
    {"CS_GRAPHS": "0"},
    {"CS_RBGS_HALO": "2", "CS_RBGS_THREADS": "256"},
    {"CS_RBGS_HALO": "1"},
], ids=["brick-h4-multi", "brick-h4-multi-unfused", "brick-h4-multi-nographs", "brick-h2", "brick-h1"])
def test_rescaled_solve_matches_oracle(ctx, orc, monkeypatch, variant):
    """rbgs_rescale_after: 32 untested iterations, the Galerkin rescale, then
    the tested iteration, from the zero, previous-step and an external start
    (the zero-start solution scaled by 0.95 plus noise), on every brick layout:
    pressure, iteration count and particles equal the oracle's bit for bit."""
    for k, v in variant.items():
        monkeypatch.setenv(k, v)
    w = cs.SeededWorkload("v60_working", seed=6, scale=4)
    nc = int(np.prod(list(w.settings.grid_dims)))
    for start in (cs.abi.CS_WARM_ZERO, cs.abi.CS_WARM_PREVIOUS, cs.abi.CS_WARM_EXTERNAL):
        s = w.settings.copy()
        s.warm_start = start
        s.rbgs_rescale_after = 32
        s.warm_fallback_residual = 1e30
        field = None
        if start == cs.abi.CS_WARM_EXTERNAL:
            probe = cs.MpmSim3D(ctx, w.settings, w.scene)
            probe.step_frame(w.frame_dt)
            p_fixed, _ = probe.read_pressure()
            probe.close()
            rng = np.random.default_rng(7)
            field = (0.95 * p_fixed.astype(np.float64) / cs.abi.FP_SCALE
                     + rng.normal(0.0, 0.5, nc)).astype(np.float32)
        sim, ref = _pair(ctx, orc, s, w.scene)
        for k in range(2 if start == cs.abi.CS_WARM_PREVIOUS else 1):
            if field is not None:
                sim.set_warm_start(field)
                ref.set_warm_start(field)
            sim.step_frame(w.frame_dt)
            ref.step_frame(w.frame_dt)
            sg, so = sim.solve_stats(), ref.solve_stats()
            assert sg.iterations == so.iterations, (variant, start, k, sg.as_dict(), so.as_dict())
            assert sg.iterations > 32 and sg.converged == so.converged == 1
            compare_state(sim, ref, f"rescale {variant} start {start} substep {k}")
        sim.close()
        ref.close()


def test_rescale_needs_the_brick_solver(ctx, monkeypatch):
    monkeypatch.setenv("CS_SOLVER", "dense")
    w = cs.SeededWorkload("v60_working", seed=6, scale=4)
    s = w.settings.copy()
    s.rbgs_rescale_after = 32
    sim = cs.MpmSim3D(ctx, s, w.scene)
    with pytest.raises(cs.CoffeeSimError, match="brick solver"):
        sim.step_frame(w.frame_dt)
    sim.close()


def test_save_load_round_trip_continues_bitwise(ctx):
    """A checkpoint taken mid-run and loaded into a fresh simulation continues
    exactly as the original does."""
    w = cs.SeededWorkload("v60_working", seed=1, scale=4)
    s = w.settings.copy()
    s.warm_start = cs.abi.CS_WARM_PREVIOUS
    a = cs.MpmSim3D(ctx, s, w.scene)
    for _ in range(3):
        a.step_frame(w.frame_dt)
    blob = a.save()
    b = cs.MpmSim3D(ctx, s, w.scene)
    b.load(blob)
    assert b.particle_count() == a.particle_count()
    for k in range(2):
        a.step_frame(w.frame_dt)
        b.step_frame(w.frame_dt)
        assert a.solve_stats().iterations == b.solve_stats().iterations
        n = a.particle_count()
        assert_bitwise(f"particles after load, substep {k}", b.read_particles(0, n), a.read_particles(0, n))
        assert_bitwise(f"affine after load, substep {k}", b.read_affine(0, n), a.read_affine(0, n))
    assert a.times() == b.times()


@pytest.mark.parametrize("variant", [{}, {"CS_RBGS_HALO": "4", "CS_RBGS_MULTI": "1"}],
                         ids=["default", "brick-h4-multi"])
def test_v60_working_full_size_fixed_pairs(ctx, orc, monkeypatch, variant):
    """The full 128^3 working grid (more active bricks than compute units, so
    every multi-brick workgroup walks several bricks): one substep with 13
    fixed RBGS pairs (an odd count, so the remainder path runs) matches the
    oracle bit for bit."""
    for k, v in variant.items():
        monkeypatch.setenv(k, v)
    w = cs.SeededWorkload("v60_working", seed=1)
    s = w.settings.copy()
    s.rbgs_mode = 0
    s.pressure_rbgs_pairs = 13
    sim, ref = _pair(ctx, orc, s, w.scene)
    sim.step_frame(w.frame_dt)
    ref.step_frame(w.frame_dt)
    compare_state(sim, ref, f"v60_working full size {variant}")


@pytest.mark.parametrize("variant", [{}, {"CS_RBGS_GROUPS": "8"}, {"CS_RBGS_FUSE": "0"}],
                         ids=["fused", "fused-8groups", "unfused"])
def test_residual_test_with_clamp_fires_matches_oracle(ctx, orc, monkeypatch, variant):
    """A short substep (dt x 0.2) fires the clamps while the solve runs to a
    loose tolerance, so the sweeps the device runs past the converged state
    must not count.  Pressure, iteration counts and the clamp-fire counter
    match the oracle."""
    for k, v in variant.items():
        monkeypatch.setenv(k, v)
    w = cs.SeededWorkload("v60_working", seed=3, scale=4)
    s = w.settings.copy()
    s.rbgs_max_iters = 400
    s.rbgs_tolerance = 0.2
    sim, ref = _pair(ctx, orc, s, w.scene)
    for k in range(2):
        sim.step_frame(w.frame_dt * 0.2)
        ref.step_frame(w.frame_dt * 0.2)
        sg, so = sim.solve_stats(), ref.solve_stats()
        assert sg.iterations == so.iterations, (k, sg.as_dict(), so.as_dict())
        assert sg.converged == so.converged
        assert ref.read_metrics()[cs.abi.METRIC_P_CLAMP] > 0
        compare_state(sim, ref, f"clamp fires {variant} substep {k}")


@pytest.mark.parametrize("variant", [{}, {"CS_P2G": "atomic"}], ids=["tiled", "atomic"])
def test_p2g_lanes_at_domain_edges_match_oracle(ctx, orc, monkeypatch, variant):
    """P2G over a 24^3 box (3 x 3 x 3 tiles) with particles spread over the
    whole domain including the half-cell margins the position clamp allows
    (base nodes -1 and n), random velocities and affine matrices, and bed-phase
    and below-threshold particles that must not scatter: the six lanes after
    the scatter, and the whole substep, match the oracle bit for bit."""
    from scenes import SceneArrays, box_settings

    for k, v in variant.items():
        monkeypatch.setenv(k, v)
    s = box_settings(n=24)
    n = 24
    side = float(s.bounds_size[0])
    dx = side / n
    rng = np.random.default_rng(11)
    m = 40000
    pos = rng.uniform(-0.5 * side - 0.5 * dx, 0.5 * side + 0.5 * dx, size=(m, 3))
    vel = rng.normal(0.0, 2.0, size=(m, 3))
    mass = np.full(m, 0.25)
    mass[::97] = 0.01  # below the active-mass threshold
    part = np.concatenate([pos, np.ones((m, 1)), vel, mass[:, None]], axis=1).astype(np.float32)
    aff = rng.normal(0.0, 3.0, size=(m, 12)).astype(np.float32)
    aff[:, 3] = 0.0
    aff[::53, 3] = 1.0  # bed phase
    aff[:, 7] = 0.0
    aff[:, 11] = 0.0
    scene = SceneArrays(part, aff)
    sim, ref = _pair(ctx, orc, s, scene.scene)
    sim.step_frame(1.0 / 600.0)
    ref.step_frame(1.0 / 600.0)
    assert_bitwise("p2g lanes", sim.read_p2g_lanes(), ref.read_p2g_lanes())
    compare_state(sim, ref, f"edge P2G {variant}")


def test_p2g_binning_with_scattered_particles_matches_oracle(ctx, orc):
    """Tiled P2G over a 96^3 box (1,728 tiles) with the particles in random
    order: a binning workgroup's 1,024 particles fall into far more tiles than
    its 64-slot LDS table holds, so most groups take their ranks from the tile
    counters directly, the rest through the table.  The lanes, and the whole
    substep, match the oracle bit for bit."""
    from scenes import SceneArrays, box_settings

    n = 96
    s = box_settings(n=n)
    side = float(s.bounds_size[0])
    rng = np.random.default_rng(12)
    m = 40000
    pos = rng.uniform(-0.45 * side, 0.45 * side, size=(m, 3))
    vel = rng.normal(0.0, 1.0, size=(m, 3))
    part = np.concatenate([pos, np.ones((m, 1)), vel, np.full((m, 1), 0.25)], axis=1).astype(np.float32)
    aff = np.zeros((m, 12), np.float32)
    scene = SceneArrays(part, aff)
    sim, ref = _pair(ctx, orc, s, scene.scene)
    sim.step_frame(1.0 / 600.0)
    ref.step_frame(1.0 / 600.0)
    assert_bitwise("p2g lanes", sim.read_p2g_lanes(), ref.read_p2g_lanes())
    compare_state(sim, ref, "scattered P2G")


@pytest.mark.parametrize("g2p", ["tiles", "global"])
def test_g2p_over_tile_bins_with_unbinned_and_outside_particles_matches_oracle(ctx, orc, monkeypatch, g2p):
    """G2P over P2G's tile bins (footprints staged in LDS) on a 90^3 box, whose
    last tiles are partial, with particles P2G leaves unbinned (not water, or at
    most a tenth of the nominal mass, including zero and negative masses) and
    particles outside the domain on every side: two substeps equal the oracle
    bit for bit, as does the per-particle G2P (CS_G2P=global)."""
    from scenes import SceneArrays, box_settings

    if g2p == "global":
        monkeypatch.setenv("CS_G2P", "global")
    n = 90
    s = box_settings(n=n)
    side = float(s.bounds_size[0])
    rng = np.random.default_rng(31)
    m = 30000
    pos = rng.uniform(-0.45 * side, 0.45 * side, size=(m, 3))
    out = rng.random(m) < 0.05
    pos[out] = rng.uniform(-0.6 * side, 0.6 * side, size=(int(out.sum()), 3))
    vel = rng.normal(0.0, 1.0, size=(m, 3))
    mass = np.full(m, 0.25)
    light = rng.random(m) < 0.03
    mass[light] = rng.choice([0.02, 0.0, -0.01, 0.025], size=int(light.sum()))
    part = np.concatenate([pos, np.ones((m, 1)), vel, mass[:, None]], axis=1).astype(np.float32)
    aff = (rng.normal(0.0, 0.5, size=(m, 12)) * np.tile([1, 1, 1, 0], 3)).astype(np.float32)
    aff[rng.random(m) < 0.03, 3] = 1.0  # not water
    scene = SceneArrays(part, aff)
    sim, ref = _pair(ctx, orc, s, scene.scene)
    for k in range(2):
        sim.step_frame(1.0 / 600.0)
        ref.step_frame(1.0 / 600.0)
        compare_state(sim, ref, f"G2P over tile bins ({g2p}), substep {k}")


@pytest.mark.parametrize("variant", [{}, {"CS_RBGS_FUSE": "0"}, {"CS_SOLVER": "dense"}],
                         ids=["fused", "unfused", "dense"])
def test_solve_with_cells_held_at_the_clamp_reaches_tolerance(ctx, orc, monkeypatch, variant):
    """A block falling at 20 units/s into a cup drives cells to the pressure
    clamp; the residual leaves those cells out, so both solves reach 1e-4 after
    the same iterations with the same state."""
    from scenes import SceneArrays, block_particles, box_settings

    for k, v in variant.items():
        monkeypatch.setenv(k, v)
    s = box_settings(n=16, cup=True, tolerance=1e-4)
    s.rbgs_max_iters = 4000
    sc = SceneArrays(block_particles(s, (4, 3, 4), (12, 9, 12), vel=(0.0, -20.0, 0.0)))
    sim, ref = _pair(ctx, orc, s, sc.scene)
    for k in range(3):
        sim.step_frame(1e-3)
        ref.step_frame(1e-3)
        sg, so = sim.solve_stats(), ref.solve_stats()
        assert (sg.iterations, sg.converged) == (so.iterations, so.converged), (k, sg.as_dict(), so.as_dict())
        assert so.converged == 1
        compare_state(sim, ref, f"clamp-held {variant} substep {k}")
    p, _ = ref.read_pressure()
    assert np.any(np.abs(p) == 8191 * 262144)
