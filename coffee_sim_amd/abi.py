"""ctypes mirrors of the structs declared in include/coffee_sim.h.

Field order and types follow the C header exactly; `cs_settings` is the
superset of the reference's `MpmSettings` (crates/sim-wasm/src/mpm_3d/mod.rs:87-104).
Importing this module loads no native code.
"""
from __future__ import annotations

import ctypes as C

CS_OK = 0
CS_ERR_INVALID = -1
CS_ERR_DEVICE = -2
CS_ERR_NOMEM = -3
CS_ERR_RANGE = -4
CS_ERR_STATE = -5

CS_MAX_OBSTACLES = 4
CS_METRICS_SLOTS = 8

CS_OBSTACLE_NONE = 0
CS_OBSTACLE_TRUNCATED_CONE = 1
CS_OBSTACLE_CYLINDER = 2

CS_EMIT_FLOW_MODEL = 0
CS_EMIT_FIXED_COUNT = 1
CS_RBGS_FIXED_PAIRS = 0
CS_RBGS_TOLERANCE = 1
CS_WARM_ZERO = 0
CS_WARM_PREVIOUS = 1
CS_WARM_EXTERNAL = 2
CS_WARM_LEARNED = 3

CS_PRESET_DEFAULT_V60 = 0
CS_PRESET_FREE_STREAM = 1
CS_PRESET_CENTER_POUR = 2
PRESETS = {
    "default_v60": CS_PRESET_DEFAULT_V60,
    "benchmark_free_stream": CS_PRESET_FREE_STREAM,
    "benchmark_center_pour": CS_PRESET_CENTER_POUR,
}

CS_WORKLOAD_V60_SMALL = 0
CS_WORKLOAD_V60_WORKING = 1
CS_WORKLOAD_V60_LARGE = 2
WORKLOADS = {"v60_small": CS_WORKLOAD_V60_SMALL, "v60_working": CS_WORKLOAD_V60_WORKING,
             "v60_large": CS_WORKLOAD_V60_LARGE}

# cell kinds (shader.rs:174-178) and metric slots (shader.rs:67-71)
CELL_AIR, CELL_SURFACE, CELL_INTERIOR, CELL_BED, CELL_SOLID = range(5)
METRIC_MAX_ABS_DIV, METRIC_FLUID_CELLS, METRIC_DIV_CLAMP, METRIC_P_CLAMP, METRIC_OVERFLOW = range(5)
# added slots: max |div| and max |div - target| after projection, max |div - target| before the solve
METRIC_POST_MAX_ABS_DIV, METRIC_POST_MAX_ABS_RES, METRIC_PRE_MAX_ABS_RES = 5, 6, 7
FP_SCALE = 262144.0
METRICS_DIV_FP_SCALE = 1024.0


class Obstacle(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("center", C.c_float * 3),
        ("top_radius", C.c_float),
        ("bot_radius", C.c_float),
        ("radius", C.c_float),
        ("top_y", C.c_float),
        ("bot_y", C.c_float),
    ]


class Spout(C.Structure):
    _fields_ = [
        ("origin", C.c_float * 3),
        ("direction", C.c_float * 3),
        ("target", C.c_float * 3),
        ("nozzle_radius", C.c_float),
        ("stem_radius", C.c_float),
        ("activation_angle_deg", C.c_float),
        ("full_flow_angle_deg", C.c_float),
        ("head_at_activation", C.c_float),
        ("head_at_full_angle", C.c_float),
        ("discharge_coeff", C.c_float),
        ("volume_to_ml", C.c_float),
        ("max_flow_rate_ml_s", C.c_float),
        ("gentle_exit_speed", C.c_float),
        ("max_exit_speed", C.c_float),
        ("stem_length", C.c_float),
    ]


class Geometry(C.Structure):
    _fields_ = [
        ("cone_enabled", C.c_uint32),
        ("cone_bot_y", C.c_float),
        ("cone_top_y", C.c_float),
        ("cone_bot_radius", C.c_float),
        ("cone_top_radius", C.c_float),
        ("cup_enabled", C.c_uint32),
        ("cup_top_y", C.c_float),
        ("cup_radius", C.c_float),
        ("cup_bottom_y", C.c_float),
        ("filter_enabled", C.c_uint32),
        ("filter_center_y", C.c_float),
        ("filter_top_y", C.c_float),
        ("filter_bot_y", C.c_float),
        ("filter_top_radius", C.c_float),
        ("filter_thickness", C.c_float),
    ]


class Settings(C.Structure):
    _fields_ = [
        ("bounds_size", C.c_float * 3),
        ("grid_dims", C.c_uint32 * 3),
        ("max_particles", C.c_uint32),
        ("substeps", C.c_uint32),
        ("gravity", C.c_float),
        ("bulk_modulus", C.c_float),
        ("viscosity", C.c_float),
        ("render_radius", C.c_float),
        ("pressure_rbgs_pairs", C.c_uint32),
        ("use_sdf_cache", C.c_uint32),
        ("num_obstacles", C.c_uint32),
        ("obstacles", Obstacle * CS_MAX_OBSTACLES),
        ("spout", Spout),
        ("initial_kettle_angle_deg", C.c_float),
        ("sdf_res", C.c_uint32),
        ("particle_vol_factor", C.c_float),
        ("wall_thickness", C.c_float),
        ("water_viscosity_m2_s", C.c_float),
        ("min_bed_permeability", C.c_float),
        ("geometry", Geometry),
        ("emitter_mode", C.c_uint32),
        ("emit_per_substep", C.c_uint32),
        ("emit_total", C.c_uint32),
        ("emit_speed", C.c_float),
        ("emit_jet_radius", C.c_float),
        ("rbgs_mode", C.c_uint32),
        ("rbgs_tolerance", C.c_float),
        ("rbgs_check_every", C.c_uint32),
        ("rbgs_max_iters", C.c_uint32),
        ("warm_start", C.c_uint32),
        ("has_bed_field", C.c_uint32),
        ("warm_fallback_residual", C.c_float),
        ("rbgs_rescale_after", C.c_uint32),
        ("rbgs_mg_cycles", C.c_uint32),
        ("emit_sample_period", C.c_uint32),
    ]

    @property
    def num_cells(self) -> int:
        return int(self.grid_dims[0]) * int(self.grid_dims[1]) * int(self.grid_dims[2])

    def copy(self) -> "Settings":
        out = Settings()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(Settings))
        return out


class BedRecord(C.Structure):
    _fields_ = [
        ("porosity", C.c_float),
        ("permeability", C.c_float),
        ("compaction", C.c_float),
        ("saturation", C.c_float),
        ("sample_y", C.c_float),
        ("velocity", C.c_float * 3),
    ]


class Scene(C.Structure):
    _fields_ = [
        ("sdf", C.POINTER(C.c_float)),
        ("sdf_class", C.POINTER(C.c_uint8)),
        ("bed_lookup", C.POINTER(C.c_int32)),
        ("bed_records", C.POINTER(BedRecord)),
        ("num_bed_records", C.c_uint32),
        ("particles", C.POINTER(C.c_float)),
        ("affine", C.POINTER(C.c_float)),
        ("num_particles", C.c_uint32),
        ("bed_particles", C.POINTER(C.c_float)),
        ("bed_affine", C.POINTER(C.c_float)),
        ("bed_extracts", C.POINTER(C.c_float)),
        ("num_bed_particles", C.c_uint32),
    ]


class FilterConfig(C.Structure):
    """FilterConfig (filter.rs:3-26)."""
    _fields_ = [("center", C.c_float * 3), ("top_y", C.c_float), ("bot_y", C.c_float),
                ("top_radius", C.c_float), ("bot_radius", C.c_float), ("thickness", C.c_float),
                ("hole_radius", C.c_float)]


class BedConfig(C.Structure):
    """BedConfig (bed.rs:7-38)."""
    _fields_ = [("center", C.c_float * 3), ("top_y", C.c_float), ("bot_y", C.c_float),
                ("top_radius", C.c_float), ("bot_radius", C.c_float), ("num_particles", C.c_uint32),
                ("initial_porosity", C.c_float), ("initial_permeability", C.c_float),
                ("extractable_mass", C.c_float)]


class SolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_uint32),
        ("fluid_cells", C.c_uint32),
        ("initial_residual", C.c_double),
        ("final_residual", C.c_double),
        ("converged", C.c_uint32),
        ("fell_back", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class KernelTiming(C.Structure):
    _fields_ = [
        ("p2g_us", C.c_double),
        ("p2g_launches", C.c_uint64),
        ("p2g_particles", C.c_uint64),
        ("g2p_us", C.c_double),
        ("g2p_launches", C.c_uint64),
        ("g2p_particles", C.c_uint64),
        ("rbgs_us", C.c_double),
        ("rbgs_launches", C.c_uint64),
        ("rbgs_iterations", C.c_uint64),
        ("rbgs_cell_updates", C.c_uint64),
        ("rbgs_cells_swept", C.c_uint64),
        ("mg_us", C.c_double),
        ("mg_starts", C.c_uint64),
        ("mg_cycles", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Workload(C.Structure):
    _fields_ = [
        ("settings", Settings),
        ("scene", Scene),
        ("frame_dt", C.c_float),
        ("substeps_total", C.c_uint32),
        ("total_particles", C.c_uint32),
        ("_owner", C.c_void_p),
    ]


class PourParams(C.Structure):
    """cs_pour_params: one seeded pour scene on a cubic grid of `side` cells."""

    _fields_ = [
        ("side", C.c_uint32),
        ("particles", C.c_uint32),
        ("pour_per_substep", C.c_uint32),
        ("kettle_angle_deg", C.c_float),
        ("seed", C.c_uint64),
    ]


class BatchConfig(C.Structure):
    """cs_batch_config: launch shapes of a batch (zeros pick the defaults)."""

    _fields_ = [
        ("sweep_groups", C.c_uint32),
        ("residual_groups", C.c_uint32),
        ("check_batches", C.c_uint32),
        ("use_graphs", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchStats(C.Structure):
    _fields_ = [
        ("substeps", C.c_uint64),
        ("particle_substeps", C.c_uint64),
        ("sweep_launches", C.c_uint64),
        ("residual_checks", C.c_uint64),
        ("max_iterations", C.c_uint64),
        ("substep_us", C.c_double),
        ("solve_us", C.c_double),
        ("cell_updates", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}
