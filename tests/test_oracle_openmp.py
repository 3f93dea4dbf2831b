"""The OpenMP build of the oracle (oracle/liborc_omp.so) is only the CPU
timing baseline, but it must compute what the serial checker computes: every
parallel pass writes only its own cell or particle, and the P2G scatter and
counters use integer atomics.  Fixed-pairs mode, so no residual-driven
control flow (the parallel residual sums in another order)."""
from __future__ import annotations

import numpy as np

import coffee_sim_amd as cs


def test_openmp_oracle_matches_the_serial_oracle_bitwise(orc):
    w = cs.SeededWorkload("v60_small", seed=5, scale=2)
    a = orc.OracleSim(w.settings, w.scene)
    b = orc.OracleSim(w.settings, w.scene, parallel=True, threads=4)
    for k in range(3):
        a.step_frame(w.frame_dt)
        b.step_frame(w.frame_dt)
        n = a.particle_count()
        assert b.particle_count() == n
        for name in ("read_p2g_lanes", "read_grid_vel", "read_metrics"):
            np.testing.assert_array_equal(getattr(a, name)(), getattr(b, name)(), err_msg=f"{name} substep {k}")
        np.testing.assert_array_equal(a.read_pressure()[0], b.read_pressure()[0])
        np.testing.assert_array_equal(a.read_particles(0, n).view(np.uint32), b.read_particles(0, n).view(np.uint32))
        np.testing.assert_array_equal(a.read_affine(0, n).view(np.uint32), b.read_affine(0, n).view(np.uint32))
    a.close()
    b.close()
