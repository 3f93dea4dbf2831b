"""Committed golden fixtures (tests/golden/*.npz, written by make_golden.py
from the CPU oracle): the oracle must keep reproducing them, and the HIP path
must reproduce them bit for bit."""
from __future__ import annotations

import hashlib
from pathlib import Path

import numpy as np
import pytest

import coffee_sim_amd as cs
from helpers import assert_bitwise

GOLD = Path(__file__).resolve().parent / "golden"


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load(name: str) -> dict:
    with np.load(GOLD / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def run_free_stream(sim) -> dict:
    sim.set_kettle_angle(36.0)
    for _ in range(4):
        sim.step_frame(1.0 / 60.0)
    return {
        "particles": sim.read_particles(),
        "affine": sim.read_affine(),
        "metrics": sim.read_metrics(),
        "p2g_lanes_sha256": digest(sim.read_p2g_lanes()),
        "grid_vel_sha256": digest(sim.read_grid_vel()),
        "final_residual": sim.solve_stats().final_residual,
        "times": np.array(sim.times(), dtype=np.float64),
    }


def check_free_stream(got: dict, want: dict) -> None:
    assert_bitwise("particles", got["particles"], want["particles"])
    assert_bitwise("affine", got["affine"], want["affine"])
    assert_bitwise("metrics", got["metrics"], want["metrics"])
    assert got["p2g_lanes_sha256"] == str(want["p2g_lanes_sha256"])
    assert got["grid_vel_sha256"] == str(want["grid_vel_sha256"])
    assert got["final_residual"] == pytest.approx(float(want["final_residual"]), rel=1e-9)
    assert np.array_equal(got["times"], want["times"])


def run_v60_small(make_sim) -> dict:
    w = cs.SeededWorkload("v60_small", seed=0, scale=2)
    sim = make_sim(w)
    sim.step_frame(w.frame_dt)
    lanes1 = sim.read_p2g_lanes()
    out = {
        "initial_particles_sha256": digest(w.arrays()["particles"]),
        "p2g_lanes_substep1_sha256": digest(lanes1),
        "p2g_mass_total_substep1": int(lanes1[0].astype(np.int64).sum()),
        "grid_vel_substep1_sha256": digest(sim.read_grid_vel()),
    }
    for _ in range(4):
        sim.step_frame(w.frame_dt)
    out["particles_substep5"] = sim.read_particles()
    out["affine_substep5_sha256"] = digest(sim.read_affine())
    return out


def check_v60_small(got: dict, want: dict) -> None:
    for k in ("initial_particles_sha256", "p2g_lanes_substep1_sha256", "grid_vel_substep1_sha256",
              "affine_substep5_sha256"):
        assert got[k] == str(want[k]), k
    assert got["p2g_mass_total_substep1"] == int(want["p2g_mass_total_substep1"])
    assert_bitwise("particles", got["particles_substep5"], want["particles_substep5"])


def run_v60_working(make_sim) -> dict:
    w = cs.SeededWorkload("v60_working", seed=1, scale=4)
    sim = make_sim(w)
    iters, res = [], []
    for _ in range(3):
        sim.step_frame(w.frame_dt)
        st = sim.solve_stats()
        iters.append(st.iterations)
        res.append(st.final_residual)
    p, k = sim.read_pressure()
    return {"iterations": iters, "final_residuals": res, "particles_sha256": digest(sim.read_particles()),
            "pressure_sha256": digest(p), "kind_sha256": digest(k), "num_particles": sim.particle_count()}


def check_v60_working(got: dict, want: dict) -> None:
    assert list(got["iterations"]) == list(want["iterations"])
    assert np.allclose(got["final_residuals"], want["final_residuals"], rtol=1e-9)
    for k in ("particles_sha256", "pressure_sha256", "kind_sha256"):
        assert got[k] == str(want[k]), k
    assert got["num_particles"] == int(want["num_particles"])


# ---- the oracle keeps reproducing its fixtures --------------------------------------
def test_oracle_free_stream_fixture(orc):
    check_free_stream(run_free_stream(orc.OracleSim(cs.preset("benchmark_free_stream"))),
                      load("free_stream_4frames"))


def test_oracle_v60_small_fixture(orc):
    check_v60_small(run_v60_small(lambda w: orc.OracleSim(w.settings, w.scene)), load("v60_small_s2"))


def test_oracle_v60_working_fixture(orc):
    check_v60_working(run_v60_working(lambda w: orc.OracleSim(w.settings, w.scene)), load("v60_working_s4"))


# ---- the HIP path reproduces them ------------------------------------------------------
@pytest.mark.gpu
def test_hip_free_stream_fixture(ctx):
    check_free_stream(run_free_stream(cs.MpmSim3D(ctx, cs.preset("benchmark_free_stream"))),
                      load("free_stream_4frames"))


@pytest.mark.gpu
def test_hip_v60_small_fixture(ctx):
    check_v60_small(run_v60_small(lambda w: cs.MpmSim3D(ctx, w.settings, w.scene)), load("v60_small_s2"))


@pytest.mark.gpu
def test_hip_v60_working_fixture(ctx):
    check_v60_working(run_v60_working(lambda w: cs.MpmSim3D(ctx, w.settings, w.scene)), load("v60_working_s4"))
