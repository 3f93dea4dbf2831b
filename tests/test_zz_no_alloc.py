"""cs_sim_pack_render_noalloc (the render packing staged in the library's
fixed device scratch array instead of an allocated buffer) against
cs_sim_pack_render, through the C ABI: bit for bit on two calls in a row, the
second with fewer particles, with and without a bed of particles; a call
beyond the array's capacity is a range error."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import coffee_sim_amd as cs
from helpers import assert_bitwise
from test_bed import bed_scene

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 1 << 18


def _noalloc(sim) -> np.ndarray:
    f = cs.load_library().cs_sim_pack_render_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
    out = np.full((sim.particle_count(), 8), np.nan, np.float32)
    rc = f(sim._h, out.ctypes.data_as(C.POINTER(C.c_float)))
    assert rc == 0, cs.load_library().cs_last_error()
    return out


def test_noalloc_render_matches_the_allocating_entry_point(ctx):
    w = cs.SeededWorkload("v60_small", seed=0, scale=2)
    sim = cs.MpmSim3D(ctx, w.settings, w.scene)
    for _ in range(2):
        sim.step_frame(w.frame_dt)
    n = sim.particle_count()
    assert n > 1000
    assert_bitwise("render, first call", _noalloc(sim), sim.pack_render())
    # the second call packs fewer particles into the same scratch array
    sim.set_particle_count(n // 3)
    got = _noalloc(sim)
    assert got.shape == (n // 3, 8)
    assert_bitwise("render, second call", got, sim.pack_render())
    sim.close()


def test_noalloc_render_with_a_bed_of_particles(ctx):
    s, sc, bed = bed_scene(substeps=1, n_water=3000)
    sim = cs.MpmSim3D(ctx, s, sc)
    sim.step_frame(1.0 / 600.0)
    assert_bitwise("bed render, first call", _noalloc(sim), sim.pack_render())
    sim.set_particle_count(1000)
    assert_bitwise("bed render, second call", _noalloc(sim), sim.pack_render())
    sim.close()


def test_noalloc_render_rejects_more_than_its_capacity(ctx):
    w = cs.SeededWorkload("v60_small", seed=0, scale=2)
    s = w.settings.copy()
    s.max_particles = NOALLOC_MAX + 16
    sim = cs.MpmSim3D(ctx, s, w.scene)
    sim.set_particle_count(NOALLOC_MAX + 1)
    f = cs.load_library().cs_sim_pack_render_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
    out = np.empty((NOALLOC_MAX + 1, 8), np.float32)
    assert f(sim._h, out.ctypes.data_as(C.POINTER(C.c_float))) == cs.abi.CS_ERR_RANGE
    assert b"262144" in cs.load_library().cs_last_error()
    sim.set_particle_count(NOALLOC_MAX)
    assert f(sim._h, out.ctypes.data_as(C.POINTER(C.c_float))) == 0
    sim.close()
