"""Run a seeded workload for a profiler: the pour (or a saved post-pour state),
then a number of substeps, with nothing else on the device (no timing events,
no probes).  Objects are left to the interpreter's shutdown.

    python -m coffee_sim_amd.profile_run [--workload v60_working] [--substeps 10]
        [--mg-cycles 3] [--warmup 2] [--state PATH]

Under `rocprofv3 --kernel-trace -- python3 -m coffee_sim_amd.profile_run ...`
the trace covers the pour too; the substeps after the line "profiled
substeps begin" are the ones to read (their dispatches come last)."""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import coffee_sim_amd as cs
from coffee_sim_amd.abi import CS_WARM_PREVIOUS


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", default="v60_working")
    ap.add_argument("--scale", type=int, default=1)
    ap.add_argument("--substeps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mg-cycles", type=int, default=3)
    ap.add_argument("--state", default=None, help="post-pour checkpoint: loaded when it exists, written otherwise")
    args = ap.parse_args()
    w = cs.SeededWorkload(args.workload, seed=1, scale=args.scale)
    s = w.settings.copy()
    s.warm_start = CS_WARM_PREVIOUS
    s.rbgs_mg_cycles = max(args.mg_cycles, 0)
    ctx = cs.Context(0)
    sim = cs.MpmSim3D(ctx, s, w.scene)
    state = Path(args.state) if args.state else None
    t0 = time.perf_counter()
    if state is not None and state.exists():
        sim.load(state.read_bytes())
    else:
        pour = (w.settings.emit_total + w.settings.emit_per_substep - 1) // max(w.settings.emit_per_substep, 1)
        for _ in range(pour):
            sim.step_frame(w.frame_dt)
        if state is not None:
            state.write_bytes(sim.save())
    for _ in range(args.warmup):
        sim.step_frame(w.frame_dt)
    sim.synchronize()
    print(f"profiled substeps begin ({sim.particle_count()} particles, {time.perf_counter() - t0:.1f} s)", flush=True)
    for _ in range(args.substeps):
        sim.step_frame(w.frame_dt)
    sim.synchronize()
    replayed, halted = sim.recorded_counts()
    print(f"profiled substeps end: RBGS iterations of the last {sim.solve_stats().iterations}; "
          f"substeps run as graph replays since creation {replayed}, halted {halted}", flush=True)
    _KEEP.extend([sim, ctx, w])
    return 0


_KEEP: list = []

if __name__ == "__main__":
    sys.exit(main())
