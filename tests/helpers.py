"""Shared helpers for the parity tests: run the HIP path and the CPU oracle on
identical inputs and compare their buffers."""
from __future__ import annotations

import numpy as np


def bits(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def assert_bitwise(name: str, got: np.ndarray, want: np.ndarray) -> None:
    assert got.shape == want.shape, f"{name}: shape {got.shape} != {want.shape}"
    g, w = bits(got), bits(want)
    diff = np.nonzero(g.reshape(-1) != w.reshape(-1))[0]
    if diff.size:
        i = int(diff[0])
        raise AssertionError(
            f"{name}: {diff.size} of {g.size} values differ; first at flat index {i}: "
            f"got {got.reshape(-1)[i]!r} want {want.reshape(-1)[i]!r}"
        )


def rel_l2(got: np.ndarray, want: np.ndarray) -> float:
    g = np.asarray(got, np.float64).reshape(-1)
    w = np.asarray(want, np.float64).reshape(-1)
    den = np.linalg.norm(w)
    return float(np.linalg.norm(g - w) / den) if den > 0 else float(np.linalg.norm(g - w))


def compare_state(sim, ref, tag: str, lanes: bool = True, pressure: bool = True) -> None:
    """Bitwise comparison of everything a substep leaves behind."""
    assert sim.counts() == ref.counts(), f"{tag}: counts {sim.counts()} != {ref.counts()}"
    n = sim.counts()[0]  # the water particles (a bed of particles stays as the scene gave it)
    assert_bitwise(f"{tag} particles", sim.read_particles(0, n), ref.read_particles(0, n))
    assert_bitwise(f"{tag} affine", sim.read_affine(0, n), ref.read_affine(0, n))
    if lanes:
        assert_bitwise(f"{tag} p2g lanes", sim.read_p2g_lanes(), ref.read_p2g_lanes())
        assert_bitwise(f"{tag} grid_vel", sim.read_grid_vel(), ref.read_grid_vel())
    if pressure:
        p1, k1 = sim.read_pressure()
        p2, k2 = ref.read_pressure()
        assert_bitwise(f"{tag} kind", k1, k2)
        assert_bitwise(f"{tag} pressure", p1, p2)
        assert_bitwise(f"{tag} metrics", sim.read_metrics(), ref.read_metrics())
