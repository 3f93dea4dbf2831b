"""The 256^3 workload (v60_large: dx 0.5 mm, 8,388,608 particles, porosity
0.4 +- 0.08 on a 16-cell noise lattice, 2.5e-4 s substeps).

CPU: the generator's scene at full size has the stated counts, grid, time
step, pour and porosity range, and a quarter-size build keeps the slab/pour
split.  GPU: the full-size scene advanced through the first pour substeps
with the default solve (multigrid start, then the red-black iteration to
1e-4 from the previous-step start; more than 1,024 active bricks, so the
launch-per-chunk iteration) stays finite, conserves particle mass up to the
emitted particles, and ends every substep with post-projection divergence at
or below the pre-solve figure.  The first two substeps from the resting slab
hold a large region at the fixed-point pressure clamp (+-8191,
state.rs:15-18) and stop at the 20,000-iteration cap; every solve from the
third substep on converges, and the test asserts that.  Parity at this size is not asserted
against the oracle (a 256^3 oracle substep takes minutes); the same kernels
are bitwise against it on the smaller grids of test_gpu_parity.py and
test_mg_start.py."""
from __future__ import annotations

import numpy as np
import pytest

import coffee_sim_amd as cs
from coffee_sim_amd import abi

FULL = 8_388_608
INITIAL = 5_872_026


def test_large_workload_inputs():
    w = cs.SeededWorkload("v60_large", seed=1)
    s = w.settings
    assert list(s.grid_dims) == [256, 256, 256]
    assert w.total_particles == FULL
    assert int(w.scene.num_particles) == INITIAL
    assert int(s.max_particles) == FULL
    assert int(s.emit_per_substep) == 16_000 and int(s.emit_total) == FULL - INITIAL
    assert w.frame_dt == pytest.approx(2.5e-4)
    assert s.rbgs_mode == abi.CS_RBGS_TOLERANCE and s.rbgs_tolerance == pytest.approx(1e-4)
    # 57.7 scene units per metre (units.rs:13): dx = 0.5 mm
    assert float(s.bounds_size[0]) / 256 == pytest.approx(0.5e-3 * 57.7, rel=1e-5)
    a = w.arrays()
    phi = a["bed_records"][:, 0]
    assert phi.min() >= 0.4 - 0.08 - 1e-6 and phi.max() <= 0.4 + 0.08 + 1e-6
    assert phi.max() - phi.min() > 0.1  # the noise spans most of its range
    assert np.isfinite(a["particles"]).all()
    w.close()


def test_large_workload_scales_down():
    w = cs.SeededWorkload("v60_large", seed=1, scale=4)
    assert list(w.settings.grid_dims) == [64, 64, 64]
    assert w.total_particles == FULL // 64
    assert int(w.scene.num_particles) == INITIAL // 64
    assert w.frame_dt == pytest.approx(1e-3)
    w.close()


@pytest.mark.gpu
def test_large_workload_steps_on_the_gpu(ctx):
    w = cs.SeededWorkload("v60_large", seed=1)
    s = w.settings.copy()
    s.warm_start = abi.CS_WARM_PREVIOUS
    s.rbgs_mg_cycles = 3
    sim = cs.MpmSim3D(ctx, s, w.scene)
    n0 = sim.particle_count()
    m0 = float(sim.read_particles(0, n0)[:, 7].astype(np.float64).sum())
    steps = 6
    for k in range(steps):
        sim.step_frame(w.frame_dt)
        st = sim.solve_stats()
        print(f"substep {k}: particles {sim.particle_count()} iterations {st.iterations} "
              f"residual {st.final_residual:.3g}", flush=True)
        if k >= 2:
            assert st.converged, st.as_dict()
        met = sim.read_metrics()
        assert met[abi.METRIC_POST_MAX_ABS_RES] <= met[abi.METRIC_PRE_MAX_ABS_RES]
    n = sim.particle_count()
    assert n == n0 + steps * 16_000
    p = sim.read_particles(0, n)
    assert np.isfinite(p).all()
    # water absorbed by no bed of particles here: mass is the initial plus 0.25 per emitted particle
    m = float(p[:, 7].astype(np.float64).sum())
    assert m == pytest.approx(m0 + 0.25 * (n - n0), rel=1e-6)
    sim.close()
    w.close()


def _emitted(orc, settings, scene, dt, substeps):
    """Positions of the particles each of `substeps` emissions writes, read
    right after it (the oracle's emitter alone, nothing advected)."""
    import ctypes as C

    o = orc.OracleSim(settings, scene)
    emit = o._L.orc_emit
    emit.argtypes = [C.c_void_p, C.c_float]
    emit.restype = None
    out = []
    for _ in range(substeps):
        n0 = o.particle_count()
        emit(o._h, dt)
        n1 = o.particle_count()
        out.append(o.read_particles(n0, n1 - n0)[:, :3].copy())
    o.close()
    return out


def test_emit_sample_period_restarts_the_disk(orc):
    """With the period equal to the per-substep count every substep restarts
    the golden-angle disk at sample 0, so a substep's emitted particles sit
    exactly where the previous substep's did when they were emitted (fixed
    spout and speed); with no period the running sample index moves them."""
    w = cs.SeededWorkload("v60_large", seed=1, scale=8)  # 32^3, 2,000 ... 250 poured per substep
    s = w.settings.copy()
    per = int(s.emit_per_substep)
    assert per > 8
    dt = w.frame_dt
    s.emit_sample_period = per
    first, second = _emitted(orc, s, w.scene, dt, 2)
    assert first.shape == second.shape == (per, 3)
    assert np.array_equal(first.view(np.uint32), second.view(np.uint32))
    s.emit_sample_period = 0
    first_running, second_running = _emitted(orc, s, w.scene, dt, 2)
    assert np.array_equal(first.view(np.uint32), first_running.view(np.uint32))
    assert not np.array_equal(first.view(np.uint32), second_running.view(np.uint32))
    w.close()


@pytest.mark.gpu
def test_emit_sample_period_matches_oracle(ctx, orc):
    """The device emitter restarts its samples as the oracle does (a period of
    1.5 substeps' emission, so the restart falls inside a substep's samples):
    the whole state is bitwise equal over three substeps."""
    from helpers import compare_state

    w = cs.SeededWorkload("v60_large", seed=1, scale=8)
    s = w.settings.copy()
    s.emit_sample_period = int(s.emit_per_substep) * 3 // 2
    sim, ref = cs.MpmSim3D(ctx, s, w.scene), orc.OracleSim(s, w.scene)
    for k in range(3):
        sim.step_frame(w.frame_dt)
        ref.step_frame(w.frame_dt)
        compare_state(sim, ref, f"sample period substep {k}")
    sim.close()
    ref.close()
    w.close()
