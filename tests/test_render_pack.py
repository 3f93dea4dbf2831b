"""prepare_render (shader.rs:2242-2276): the RenderParticle layout packed from
the particle state.  The oracle is pinned by the closed forms of the pass; the
device packing is compared with the oracle bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

import coffee_sim_amd as cs
from helpers import assert_bitwise
from scenes import SceneArrays, block_particles, box_settings


def _crafted(s):
    parts = block_particles(s, (6, 6, 6), (8, 8, 8), vel=(3.0, -4.0, 12.0))  # |v| = 13: colour 1.3
    parts = np.array(parts, np.float32)
    parts[0, 7] = 0.0          # inactive: parked
    parts[1, 4:7] = (30.0, 0.0, 0.0)  # fast: colour clamped at 2
    aff = np.zeros((len(parts), 12), np.float32)
    aff[2, 3] = 1.0            # not water (no bed particles in field mode): colour -1
    return SceneArrays(parts, aff), parts


def test_oracle_packing_closed_forms(orc):
    s = box_settings(n=16)
    sc, parts = _crafted(s)
    o = orc.OracleSim(s, sc.scene)
    r = o.pack_render()
    dx = np.float32(s.bounds_size[0] / s.grid_dims[0])
    assert r.shape == (len(parts), 8)
    np.testing.assert_array_equal(r[0], np.array([0, -1e6, 0, -999, 0, 0, 0, 0], np.float32))
    assert r[1, 3] == 2.0
    assert r[2, 3] == -1.0
    np.testing.assert_array_equal(r[3:, :3], parts[3:, :3])
    speed = np.sqrt(np.float32(9.0) + np.float32(16.0) + np.float32(144.0))
    assert np.all(r[3:, 3] == np.float32(speed / np.float32(10.0)))
    assert np.all(r[1:, 4] == dx * np.float32(0.18)) and not r[:, 5:].any()
    o.close()


@pytest.mark.gpu
def test_device_packing_matches_oracle(ctx, orc):
    s = box_settings(n=16)
    sc, _ = _crafted(s)
    sim, ref = cs.MpmSim3D(ctx, s, sc.scene), orc.OracleSim(s, sc.scene)
    assert_bitwise("render crafted", sim.pack_render(), ref.pack_render())
    w = cs.SeededWorkload("v60_small", seed=2, scale=2)
    sim2, ref2 = cs.MpmSim3D(ctx, w.settings, w.scene), orc.OracleSim(w.settings, w.scene)
    for k in range(3):
        sim2.step_frame(w.frame_dt)
        ref2.step_frame(w.frame_dt)
        assert_bitwise(f"render substep {k}", sim2.pack_render(), ref2.pack_render())
    for h in (sim, ref, sim2, ref2):
        h.close()
