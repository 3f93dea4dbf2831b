"""Batched scenes (cs_batch_*): many simulations stepped together, one launch
per pass, against the same simulations stepped one by one.

Each scene's state after a batched substep must equal, bit for bit, the state
of the same scene stepped alone with the residual test run as a pass of its
own (CS_RBGS_FUSE=0, whose per-brick partial sums and fixed reduction order
the batched residual shares); the single-scene path is itself checked against
the oracle in test_gpu_parity.py, and a batch of three scenes is checked
against the oracle directly here.  The full 64-scene set runs three batched
substeps with property checks (finite, mass, convergence)."""
from __future__ import annotations

import numpy as np
import pytest

import coffee_sim_amd as cs
from coffee_sim_amd import abi
from coffee_sim_amd.batch import PourSpec, make_batch, make_sims
from helpers import compare_state

pytestmark = pytest.mark.gpu

# uneven grids (three sides), particle counts and pour rates
SPECS = [
    PourSpec(48, 30_000, 900, 40.0, 11),
    PourSpec(32, 6_000, 300, 55.0, 12),
    PourSpec(64, 45_000, 1500, 32.0, 13),
    PourSpec(32, 9_000, 2000, 47.0, 14),
    PourSpec(48, 12_000, 600, 59.0, 15),
]


def _solo(ctx, monkeypatch, specs, **kw):
    monkeypatch.setenv("CS_RBGS_FUSE", "0")
    works, sims = make_sims(ctx, specs, **kw)
    monkeypatch.delenv("CS_RBGS_FUSE")
    return works, sims


def _check_same(batched, solo, tag):
    for i, (a, b) in enumerate(zip(batched, solo)):
        sa, sb = a.solve_stats(), b.solve_stats()
        assert (sa.iterations, sa.converged) == (sb.iterations, sb.converged), (tag, i, sa.as_dict(), sb.as_dict())
        assert sa.final_residual == sb.final_residual and sa.initial_residual == sb.initial_residual, (tag, i)
        compare_state(a, b, f"{tag} scene {i}")


@pytest.mark.parametrize("graphs", [1, 0], ids=["graphs", "no-graphs"])
def test_batched_tolerance_solves_match_single_scenes(ctx, monkeypatch, graphs):
    kw = dict(warm_start=abi.CS_WARM_PREVIOUS, max_iters=4000)
    works, solo = _solo(ctx, monkeypatch, SPECS, **kw)
    _, mine = make_sims(ctx, SPECS, **kw)
    cfg = cs.Batch.default_config()
    cfg.use_graphs = graphs
    cfg.sweep_groups = 64  # several entries per workgroup: exercises the prefetch across scenes
    batch = make_batch(ctx, mine, cfg, p2g_tiled={32: False})  # 32^3 scenes take the atomic P2G
    dt = works[0].frame_dt
    particle_substeps = 0
    for k in range(5):
        print(f"substep {k}", flush=True)
        batch.step_frame(dt)
        for s in solo:
            s.step_frame(dt)
        _check_same(mine, solo, f"substep {k}")
        particle_substeps += sum(s.particle_count() for s in mine)
    st = batch.stats()
    assert st.substeps == 5 and st.sweep_launches > 0
    assert st.particle_substeps == particle_substeps
    batch.close()
    for s in solo + mine:
        s.close()


def test_batched_zero_start_and_fixed_check_schedule(ctx, monkeypatch):
    """Zero start, a check every 6 iterations (three sweep launches per chunk,
    so the pressure buffers swap from one chunk to the next) and a cap of 60
    iterations that the scenes reach."""
    kw = dict(warm_start=abi.CS_WARM_ZERO, check_every=6, max_iters=60)
    works, solo = _solo(ctx, monkeypatch, SPECS[:3], **kw)
    _, mine = make_sims(ctx, SPECS[:3], **kw)
    batch = make_batch(ctx, mine)
    dt = works[0].frame_dt
    for k in range(3):
        print(f"substep {k}", flush=True)
        batch.step_frame(dt)
        for s in solo:
            s.step_frame(dt)
        _check_same(mine, solo, f"substep {k}")
    batch.close()
    for s in solo + mine:
        s.close()


def test_batched_fixed_pairs_match_single_scenes(ctx, monkeypatch):
    """The small V60 scene's fixed 40-pair mode (two seeds at half size)."""
    ws = [cs.SeededWorkload("v60_small", seed=sd, scale=2) for sd in (3, 4)]
    monkeypatch.setenv("CS_RBGS_FUSE", "0")
    solo = [cs.MpmSim3D(ctx, w.settings, w.scene) for w in ws]
    monkeypatch.delenv("CS_RBGS_FUSE")
    mine = [cs.MpmSim3D(ctx, w.settings, w.scene) for w in ws]
    batch = cs.Batch(ctx, mine)
    for k in range(4):
        print(f"substep {k}", flush=True)
        batch.step_frame(ws[0].frame_dt)
        for s in solo:
            s.step_frame(ws[0].frame_dt)
        for i, (a, b) in enumerate(zip(mine, solo)):
            assert a.solve_stats().iterations == b.solve_stats().iterations == 40
            assert a.solve_stats().final_residual == b.solve_stats().final_residual
            compare_state(a, b, f"fixed substep {k} scene {i}")
    batch.close()
    for s in solo + mine:
        s.close()


def test_batch_rejects_mixed_solver_settings(ctx):
    specs = SPECS[:2]
    _, a = make_sims(ctx, specs[:1], max_iters=400)
    _, b = make_sims(ctx, specs[1:], max_iters=800)
    with pytest.raises(cs.CoffeeSimError, match="rbgs_max_iters"):
        cs.Batch(ctx, a + b)
    w = cs.SeededWorkload("v60_small", seed=3, scale=2)
    fixed = cs.MpmSim3D(ctx, w.settings, w.scene)
    with pytest.raises(cs.CoffeeSimError, match="rbgs_mode"):
        cs.Batch(ctx, a + [fixed])
    with pytest.raises(cs.CoffeeSimError, match="twice"):
        cs.Batch(ctx, a + a)
    for s in a + b + [fixed]:
        s.close()


def test_batched_scenes_match_the_oracle(ctx, orc):
    """Batched scenes against the CPU oracle stepping each scene alone (uneven
    grids 32^3 and 48^3, the 32^3 scenes with the atomic P2G layout, the
    48^3 one tiled): iteration counts, residuals and the whole state bit for
    bit over three substeps of the pour."""
    specs = [PourSpec(32, 6_000, 300, 55.0, 21), PourSpec(48, 12_000, 600, 40.0, 22),
             PourSpec(32, 9_000, 2000, 47.0, 23)]
    kw = dict(warm_start=abi.CS_WARM_PREVIOUS, max_iters=4000)
    works, mine = make_sims(ctx, specs, **kw)
    batch = make_batch(ctx, mine, p2g_tiled={32: False})
    refs = []
    for w in works:
        s = w.settings.copy()
        s.warm_start = kw["warm_start"]
        s.rbgs_max_iters = kw["max_iters"]
        refs.append(orc.OracleSim(s, w.scene))
    dt = works[0].frame_dt
    for k in range(3):
        print(f"substep {k}", flush=True)
        batch.step_frame(dt)
        for r in refs:
            r.step_frame(dt)
        for i, (a, b) in enumerate(zip(mine, refs)):
            sa, sb = a.solve_stats(), b.solve_stats()
            assert (sa.iterations, sa.converged) == (sb.iterations, sb.converged), (k, i, sa.as_dict(), sb.as_dict())
            compare_state(a, b, f"batch vs oracle substep {k} scene {i}")
    batch.close()
    for s in mine:
        s.close()
    for r in refs:
        r.close()


def test_64_scene_batch_at_full_size(ctx):
    """The 64-scene set of the batched workload (grids 48-96, 20k-400k
    particles, kettle angles 30-60 degrees, pour rates 500-3,000 per substep;
    seed 0) for three batched substeps: every scene stays finite, holds its
    initial mass plus 0.25 per emitted particle, and converges its solve to
    the tolerance (or reaches the iteration cap only on the first substep,
    from the resting slab)."""
    from coffee_sim_amd.batch import pour_specs

    specs = pour_specs(64, seed=0)
    works, sims = make_sims(ctx, specs)
    m0 = []
    for sm in sims:
        n = sm.particle_count()
        m0.append((n, float(sm.read_particles(0, n)[:, 7].astype(np.float64).sum())))
    batch = make_batch(ctx, sims)
    dt = works[0].frame_dt
    for k in range(3):
        print(f"substep {k}", flush=True)
        batch.step_frame(dt)
        capped = [i for i, sm in enumerate(sims) if not sm.solve_stats().converged]
        if k > 0:
            assert not capped, (k, capped)
    for i, sm in enumerate(sims):
        n = sm.particle_count()
        p = sm.read_particles(0, n)
        assert np.isfinite(p).all(), i
        n0, mass0 = m0[i]
        assert n > n0, i  # every scene pours
        assert float(p[:, 7].astype(np.float64).sum()) == pytest.approx(mass0 + 0.25 * (n - n0), rel=1e-6), i
    st = batch.stats()
    assert st.substeps == 3
    batch.close()
    for s in sims:
        s.close()
