"""Recorded substeps: step_frame's passes recorded once and submitted as one
unit per substep (mod.rs:305-416), with the substep's uniforms written to a
device buffer before it (write_uniforms, mod.rs:569-670).  Here the whole
substep is one HIP graph whose kernels read the device-resident parameter
block; each substep is one parameter copy plus one graph launch.

GPU:
- replays equal launch-by-launch substeps bit for bit (particles, affine
  rows, grid lanes, pressure, kinds, metrics, solve statistics) over a pour
  whose emission width and particle count change every substep, with the
  multigrid start and the previous-step start, and with fixed pairs from the
  zero start;
- replays equal the CPU oracle bit for bit on the small V60 workload;
- eight replays between two waits of the host equal eight launch-by-launch
  substeps;
- with a recorded brick capacity below the scene's active bricks a replay
  halts at the brick scan mid-frame, the host finishes it and the frame's
  remaining substeps launch by launch, and the state still equals the
  launch-by-launch run's.
"""
from __future__ import annotations

import numpy as np
import pytest

import coffee_sim_amd as cs
from helpers import compare_state

pytestmark = pytest.mark.gpu


def _make(ctx, settings, scene, monkeypatch, env):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    sim = cs.MpmSim3D(ctx, settings, scene)
    for k in env:
        monkeypatch.delenv(k)
    return sim


def _advance(sim, dt, frames, tag):
    stats = []
    for f in range(frames):
        sim.step_frame(dt)
        st = sim.solve_stats()
        stats.append((int(st.iterations), int(st.converged), float(st.final_residual)))
        print(f"{tag} frame {f}: particles {sim.particle_count()} iterations {st.iterations}", flush=True)
    return stats


def _pair(ctx, settings, scene, dt, frames, monkeypatch, cap=None):
    env = {"CS_SUBSTEP_GRAPHS": "1"}
    if cap is not None:
        env["CS_SUBSTEP_CAP"] = str(cap)
    rec = _make(ctx, settings, scene, monkeypatch, env)
    eag = _make(ctx, settings, scene, monkeypatch, {"CS_SUBSTEP_GRAPHS": "0"})
    s_rec = _advance(rec, dt, frames, "recorded")
    s_eag = _advance(eag, dt, frames, "launch by launch")
    return rec, eag, s_rec, s_eag


@pytest.mark.parametrize("mode", ["mg_previous", "fixed_zero"])
def test_recorded_substeps_equal_launch_by_launch(ctx, monkeypatch, mode):
    w = cs.SeededWorkload("v60_working", seed=3, scale=4)
    s = w.settings.copy()
    if mode == "mg_previous":
        s.warm_start = cs.abi.CS_WARM_PREVIOUS
        s.rbgs_mg_cycles = 3
    else:
        s.warm_start = cs.abi.CS_WARM_ZERO
        s.rbgs_mode = cs.abi.CS_RBGS_FIXED_PAIRS
        s.pressure_rbgs_pairs = 24
    frames = 14
    rec, eag, s_rec, s_eag = _pair(ctx, s, w.scene, w.frame_dt, frames, monkeypatch)
    replayed, halted = rec.recorded_counts()
    assert replayed == frames * max(int(s.substeps), 1) and halted == 0, (replayed, halted)
    assert eag.recorded_counts() == (0, 0)
    assert s_rec == s_eag
    compare_state(rec, eag, f"recorded vs launch by launch ({mode})")
    assert np.array_equal(rec.read_metrics(), eag.read_metrics())
    rec.close()
    eag.close()


def test_recorded_substeps_equal_the_oracle(ctx, orc, monkeypatch):
    w = cs.SeededWorkload("v60_small", seed=0, scale=2)
    monkeypatch.setenv("CS_SUBSTEP_GRAPHS", "1")
    sim = cs.MpmSim3D(ctx, w.settings, w.scene)
    monkeypatch.delenv("CS_SUBSTEP_GRAPHS")
    ref = orc.OracleSim(w.settings, w.scene)
    for f in range(6):
        sim.step_frame(w.frame_dt)
        ref.step_frame(w.frame_dt)
        print(f"frame {f}", flush=True)
    assert sim.recorded_counts()[0] == 6 * max(int(w.settings.substeps), 1)
    compare_state(sim, ref, "recorded vs oracle")
    sim.close()


def test_recorded_substeps_eight_per_frame(ctx, monkeypatch):
    """Eight replays between two waits of the host, no halt."""
    w = cs.SeededWorkload("v60_working", seed=5, scale=4)
    s = w.settings.copy()
    s.warm_start = cs.abi.CS_WARM_PREVIOUS
    s.rbgs_mg_cycles = 2
    s.substeps = 8
    rec, eag, s_rec, s_eag = _pair(ctx, s, w.scene, w.frame_dt * 8, 3, monkeypatch)
    assert rec.recorded_counts() == (24, 0), rec.recorded_counts()
    assert s_rec == s_eag
    compare_state(rec, eag, "eight replays per wait vs launch by launch")
    rec.close()
    eag.close()


def test_recorded_substeps_halt_and_finish_launch_by_launch(ctx, monkeypatch):
    """benchmark_free_stream (mod.rs:145-152) pours into an empty domain, ten
    substeps a frame: with a recorded capacity of 4 active bricks the first
    substeps replay to the end and a later one halts mid-frame."""
    s = cs.preset("benchmark_free_stream")
    env = {"CS_SUBSTEP_GRAPHS": "1", "CS_SUBSTEP_CAP": "4"}
    rec = _make(ctx, s, None, monkeypatch, env)
    eag = _make(ctx, s, None, monkeypatch, {"CS_SUBSTEP_GRAPHS": "0"})
    for a in (rec, eag):
        a.set_kettle_angle(36.0)
    s_rec = _advance(rec, 1.0 / 60.0, 6, "recorded")
    s_eag = _advance(eag, 1.0 / 60.0, 6, "launch by launch")
    replayed, halted = rec.recorded_counts()
    print(f"replayed {replayed} halted {halted}", flush=True)
    assert halted >= 1 and replayed >= 1, (replayed, halted)
    assert s_rec == s_eag
    compare_state(rec, eag, "halted replays vs launch by launch")
    rec.close()
    eag.close()
