"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py

The reference ships no golden vectors for its solver (its tests assert
statistical invariants only), so these fixtures are produced by this
project's oracle on seeded inputs.  They pin the oracle against regressions
and give the HIP path a fixed artifact to reproduce bit for bit.
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import coffee_sim_amd as cs  # noqa: E402
import orc  # noqa: E402


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def free_stream() -> dict:
    s = cs.preset("benchmark_free_stream")
    o = orc.OracleSim(s)
    o.set_kettle_angle(36.0)
    for _ in range(4):
        o.step_frame(1.0 / 60.0)
    st = o.solve_stats()
    return {
        "particles": o.read_particles(),
        "affine": o.read_affine(),
        "metrics": o.read_metrics(),
        "p2g_lanes_sha256": np.array(digest(o.read_p2g_lanes())),
        "grid_vel_sha256": np.array(digest(o.read_grid_vel())),
        "final_residual": np.array(st.final_residual),
        "times": np.array(o.times(), dtype=np.float64),
    }


def v60_small() -> dict:
    w = cs.SeededWorkload("v60_small", seed=0, scale=2)
    o = orc.OracleSim(w.settings, w.scene)
    o.step_frame(w.frame_dt)
    lanes1 = o.read_p2g_lanes()
    gv1 = o.read_grid_vel()
    for _ in range(4):
        o.step_frame(w.frame_dt)
    return {
        "initial_particles_sha256": np.array(digest(w.arrays()["particles"])),
        "p2g_lanes_substep1_sha256": np.array(digest(lanes1)),
        "p2g_mass_total_substep1": np.array(int(lanes1[0].astype(np.int64).sum())),
        "grid_vel_substep1_sha256": np.array(digest(gv1)),
        "particles_substep5": o.read_particles(),
        "affine_substep5_sha256": np.array(digest(o.read_affine())),
    }


def v60_working() -> dict:
    w = cs.SeededWorkload("v60_working", seed=1, scale=4)
    o = orc.OracleSim(w.settings, w.scene)
    iters, res = [], []
    for _ in range(3):
        o.step_frame(w.frame_dt)
        st = o.solve_stats()
        iters.append(st.iterations)
        res.append(st.final_residual)
    p, k = o.read_pressure()
    return {
        "iterations": np.array(iters),
        "final_residuals": np.array(res),
        "particles_sha256": np.array(digest(o.read_particles())),
        "pressure_sha256": np.array(digest(p)),
        "kind_sha256": np.array(digest(k)),
        "num_particles": np.array(o.particle_count()),
    }


def main() -> None:
    for name, fn in [("free_stream_4frames", free_stream), ("v60_small_s2", v60_small),
                     ("v60_working_s4", v60_working)]:
        data = fn()
        np.savez_compressed(HERE / f"{name}.npz", **data)
        print(f"wrote {name}.npz")


if __name__ == "__main__":
    main()
