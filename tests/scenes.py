"""Small hand-built scenes for the oracle invariant tests and the GPU parity
tests (an open box, a cylindrical cup, optional porous bed field)."""
from __future__ import annotations

import ctypes as C

import numpy as np

import coffee_sim_amd as cs
from coffee_sim_amd import abi

DX = 0.175  # the reference default cell size (mod.rs:110-111)


def box_settings(n: int = 16, cup: bool = False, rbgs_pairs: int = 40, tolerance: float | None = None,
                 bulk: float = 900.0, viscosity: float = 1.2, pvol: float = 0.125) -> abi.Settings:
    s = cs.preset("benchmark_free_stream")
    side = n * DX
    for c in range(3):
        s.bounds_size[c] = side
        s.grid_dims[c] = n
    s.max_particles = 65536
    s.substeps = 1
    s.bulk_modulus = bulk
    s.viscosity = viscosity
    s.particle_vol_factor = pvol
    s.geometry.cone_enabled = 0
    s.geometry.cup_enabled = 0
    s.geometry.filter_enabled = 0
    s.emitter_mode = abi.CS_EMIT_FIXED_COUNT
    s.emit_per_substep = 0
    s.sdf_res = 2 * n
    s.num_obstacles = 0
    if cup:
        s.num_obstacles = 1
        o = s.obstacles[0]
        o.kind = abi.CS_OBSTACLE_CYLINDER
        o.radius = 0.42 * side
        o.top_y = 0.5 * side
        o.bot_y = -0.38 * side
        s.wall_thickness = 2.0 * DX
    s.pressure_rbgs_pairs = rbgs_pairs
    if tolerance is not None:
        s.rbgs_mode = abi.CS_RBGS_TOLERANCE
        s.rbgs_tolerance = tolerance
        s.rbgs_max_iters = 50000
    return s


def block_particles(s: abi.Settings, lo_cells, hi_cells, vel=(0.0, 0.0, 0.0), seed: int = 0,
                    inside=None) -> np.ndarray:
    """8 jittered particles per cell in the cell box [lo, hi)."""
    rng = np.random.default_rng(seed)
    dx = s.bounds_size[0] / s.grid_dims[0]
    origin = -0.5 * np.array(list(s.bounds_size), np.float64)
    out = []
    for iz in range(lo_cells[2], hi_cells[2]):
        for iy in range(lo_cells[1], hi_cells[1]):
            for ix in range(lo_cells[0], hi_cells[0]):
                if inside is not None and not inside(ix, iy, iz):
                    continue
                for sub in range(8):
                    off = np.array([sub & 1, (sub >> 1) & 1, (sub >> 2) & 1], np.float64)
                    p = origin + (np.array([ix, iy, iz]) + (off + rng.random(3)) * 0.5) * dx
                    out.append([p[0], p[1], p[2], 1.0, vel[0], vel[1], vel[2], 0.25])
    return np.array(out, np.float32).reshape(-1, 8)


class SceneArrays:
    """Owns numpy arrays and exposes them as a cs_scene."""

    def __init__(self, particles: np.ndarray, affine: np.ndarray | None = None,
                 bed_lookup: np.ndarray | None = None, bed_records: np.ndarray | None = None):
        self.p = np.ascontiguousarray(particles, np.float32)
        self.a = np.ascontiguousarray(affine if affine is not None else np.zeros((len(self.p), 12)), np.float32)
        self.scene = abi.Scene()
        self.scene.particles = self.p.ctypes.data_as(C.POINTER(C.c_float))
        self.scene.affine = self.a.ctypes.data_as(C.POINTER(C.c_float))
        self.scene.num_particles = len(self.p)
        if bed_lookup is not None:
            self.lookup = np.ascontiguousarray(bed_lookup, np.int32)
            self.records = np.ascontiguousarray(bed_records, np.float32).reshape(-1, 8)
            self.scene.bed_lookup = self.lookup.ctypes.data_as(C.POINTER(C.c_int32))
            self.scene.bed_records = C.cast(self.records.ctypes.data, C.POINTER(abi.BedRecord))
            self.scene.num_bed_records = len(self.records)


def bed_field(s: abi.Settings, lo_cells, hi_cells, porosity=0.4, permeability=1e-9):
    n = s.grid_dims
    lookup = -np.ones(n[0] * n[1] * n[2], np.int32)
    recs = []
    dx = s.bounds_size[0] / n[0]
    for iz in range(lo_cells[2], hi_cells[2]):
        for iy in range(lo_cells[1], hi_cells[1]):
            for ix in range(lo_cells[0], hi_cells[0]):
                y = -0.5 * s.bounds_size[1] + (iy + 0.5) * dx
                lookup[(iz * n[1] + iy) * n[0] + ix] = len(recs)
                recs.append([porosity, permeability, 0.0, 0.0, y + dx, 0.0, 0.0, 0.0])
    return lookup, np.array(recs, np.float32)


def cell_index(s, ix, iy, iz):
    return (iz * s.grid_dims[1] + iy) * s.grid_dims[0] + ix
