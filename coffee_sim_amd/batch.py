"""Batched pour scenes: the seeded scene set of the batched workload and the
helpers that build and step it (cs_batch_* in the native library).

The scene set: `count` independent pours whose grids are cubic with the side
drawn from {48, 64, 80, 96}, particle counts log-uniform in 20k-400k, kettle
angle uniform in 30-60 degrees and pour rate uniform in 500-3,000 particles per
substep, each solved to a 1e-4 relative residual (previous-step start).  The
draws come from one seeded generator, so a seed names the whole set."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import Batch, Context, MpmSim3D, SeededWorkload, abi

SIDES = (48, 64, 80, 96)


@dataclass(frozen=True)
class PourSpec:
    side: int
    particles: int
    pour_per_substep: int
    kettle_angle_deg: float
    seed: int

    def workload(self) -> SeededWorkload:
        return SeededWorkload.pour(self.side, self.particles, self.pour_per_substep, self.kettle_angle_deg,
                                   self.seed)


def pour_specs(count: int = 64, seed: int = 0, sides=SIDES, particles=(20_000, 400_000),
               angle_deg=(30.0, 60.0), pour=(500, 3000)) -> list[PourSpec]:
    """The seeded scene set (see the module docstring)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        side = int(rng.choice(np.asarray(sides)))
        n = int(round(float(np.exp(rng.uniform(np.log(particles[0]), np.log(particles[1]))))))
        a = float(rng.uniform(*angle_deg))
        rate = int(rng.integers(pour[0], pour[1] + 1))
        out.append(PourSpec(side, n, rate, a, seed * 1000 + i + 1))
    return out


def make_sims(ctx: Context, specs: list[PourSpec], warm_start: int = abi.CS_WARM_PREVIOUS,
              check_every: int | None = None, max_iters: int | None = None):
    """One simulation per spec; returns (workloads, sims)."""
    works, sims = [], []
    for sp in specs:
        w = sp.workload()
        s = w.settings.copy()
        s.warm_start = warm_start
        if check_every is not None:
            s.rbgs_check_every = check_every
        if max_iters is not None:
            s.rbgs_max_iters = max_iters
        works.append(w)
        sims.append(MpmSim3D(ctx, s, w.scene))
    return works, sims


def make_batch(ctx: Context, sims: list, config: abi.BatchConfig | None = None,
               p2g_tiled: dict | None = None) -> Batch:
    """A batch over `sims`; `p2g_tiled` maps a grid side to its P2G layout."""
    b = Batch(ctx, sims, config)
    if p2g_tiled:
        for i, sm in enumerate(sims):
            side = int(sm.dims[0])
            if side in p2g_tiled:
                b.set_scene_p2g(i, bool(p2g_tiled[side]))
    return b
