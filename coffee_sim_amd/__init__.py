"""coffee_sim_amd — MI355X-native MLS-MPM substep of craigxchen/coffee-sim.

Python host over the C ABI in include/coffee_sim.h.  `MpmSim3D` mirrors the
reference's `MpmSim3D` (crates/sim-wasm/src/mpm_3d/mod.rs:179-567): the same
constructor inputs (settings presets), `step_frame`, `reset`, kettle/spout
setters and counters, plus readbacks of the particle and grid buffers that the
reference's tests take through staging copies (physics_tests.rs:61-139).

All numerics run in the native library (coffee_sim_amd/lib/libcoffee_sim.so,
gfx950 kernels).  There is no Python or CPU fallback: if the library is not
built, `load_library()` raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
import os
from pathlib import Path

import numpy as np

from . import abi
from .abi import (  # noqa: F401  (re-exported)
    BedRecord,
    Geometry,
    Obstacle,
    Scene,
    Settings,
    SolveStats,
    Spout,
    Workload,
)

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "lib" / "libcoffee_sim.so"
_lib = None


class CoffeeSimError(RuntimeError):
    def __init__(self, code: int, where: str, message: str):
        super().__init__(f"{where} failed ({code}): {message}")
        self.code = code


def build(verbose: bool = False) -> Path:
    from .build import build as _build

    global _lib
    _lib = None
    return _build(verbose=verbose)


def _bind(lib: C.CDLL) -> None:
    P = C.POINTER
    sigs = {
        "cs_last_error": (C.c_char_p, []),
        "cs_settings_preset": (C.c_int, [C.c_int, P(Settings)]),
        "cs_workload_build": (C.c_int, [C.c_int, C.c_uint64, C.c_uint32, P(Workload)]),
        "cs_workload_free": (C.c_int, [P(Workload)]),
        "cs_generate_sdf": (C.c_int, [P(Settings), P(C.c_float)]),
        "cs_generate_sdf_class": (C.c_int, [P(Settings), P(C.c_float), P(C.c_uint8)]),
        "cs_device_count": (C.c_int, [P(C.c_int)]),
        "cs_ctx_create": (C.c_int, [C.c_int, P(C.c_void_p)]),
        "cs_ctx_destroy": (C.c_int, [C.c_void_p]),
        "cs_sim_create": (C.c_int, [C.c_void_p, P(Settings), P(Scene), P(C.c_void_p)]),
        "cs_sim_destroy": (C.c_int, [C.c_void_p]),
        "cs_sim_reset": (C.c_int, [C.c_void_p]),
        "cs_sim_step_frame": (C.c_int, [C.c_void_p, C.c_float]),
        "cs_sim_get_recorded_counts": (C.c_int, [C.c_void_p, P(C.c_uint64), P(C.c_uint64)]),
        "cs_sim_synchronize": (C.c_int, [C.c_void_p]),
        "cs_sim_set_kettle_angle": (C.c_int, [C.c_void_p, C.c_float]),
        "cs_sim_set_spout_position": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float]),
        "cs_sim_set_spout_target": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float]),
        "cs_sim_get_counts": (C.c_int, [C.c_void_p, P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
        "cs_sim_get_time": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float), P(C.c_uint32)]),
        "cs_sim_get_inflow": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float), P(C.c_float)]),
        "cs_sim_write_particles": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, P(C.c_float), P(C.c_float)]),
        "cs_sim_set_particle_count": (C.c_int, [C.c_void_p, C.c_uint32]),
        "cs_sim_read_particles": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, P(C.c_float)]),
        "cs_sim_read_affine": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, P(C.c_float)]),
        "cs_sim_read_p2g_lanes": (C.c_int, [C.c_void_p, P(C.c_int32)]),
        "cs_sim_read_grid_vel": (C.c_int, [C.c_void_p, P(C.c_float)]),
        "cs_sim_read_pressure": (C.c_int, [C.c_void_p, P(C.c_int32), P(C.c_int32)]),
        "cs_sim_read_metrics": (C.c_int, [C.c_void_p, P(C.c_uint32)]),
        "cs_sim_read_sdf": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_uint8)]),
        "cs_sim_read_bed_lookup": (C.c_int, [C.c_void_p, P(C.c_int32)]),
        "cs_sim_read_bed_extracts": (C.c_int, [C.c_void_p, P(C.c_float)]),
        "cs_sim_read_bed_particles": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float)]),
        "cs_filter_config_default": (None, [P(abi.FilterConfig)]),
        "cs_bed_config_default": (None, [P(abi.BedConfig)]),
        "cs_bed_config_seated_in_filter": (None, [P(abi.FilterConfig), P(abi.BedConfig)]),
        "cs_bed_init": (C.c_int, [P(abi.BedConfig), P(C.c_uint32), P(C.c_float), C.c_uint32, P(C.c_float),
                                  P(C.c_float), P(C.c_float), P(C.c_int32), P(C.c_uint32)]),
        "cs_sim_pack_render": (C.c_int, [C.c_void_p, P(C.c_float)]),
        "cs_sim_get_solve_stats": (C.c_int, [C.c_void_p, P(SolveStats)]),
        "cs_sim_set_warm_start": (C.c_int, [C.c_void_p, P(C.c_float)]),
        "cs_sim_save": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, P(C.c_size_t)]),
        "cs_sim_load": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
        "cs_sim_get_pass_times": (C.c_int, [C.c_void_p, P(C.c_float)]),
        "cs_sim_enable_timing": (C.c_int, [C.c_void_p, C.c_int]),
        "cs_device_expf": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float), C.c_uint32]),
        "cs_device_sincosf": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float), P(C.c_float), C.c_uint32]),
        "cs_sim_get_kernel_timing": (C.c_int, [C.c_void_p, P(abi.KernelTiming)]),
        "cs_sim_reset_kernel_timing": (C.c_int, [C.c_void_p]),
        "cs_unet_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, P(C.c_void_p)]),
        "cs_unet_destroy": (C.c_int, [C.c_void_p]),
        "cs_unet_param_count": (C.c_int, [C.c_void_p, P(C.c_uint32)]),
        "cs_unet_get_params": (C.c_int, [C.c_void_p, P(C.c_float)]),
        "cs_unet_set_params": (C.c_int, [C.c_void_p, P(C.c_float)]),
        "cs_unet_set_scales": (C.c_int, [C.c_void_p, C.c_float, C.c_float]),
        "cs_unet_get_scales": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float)]),
        "cs_unet_forward": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float)]),
        "cs_unet_gradient": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float), P(C.c_uint8), P(C.c_float),
                                       P(C.c_float)]),
        "cs_unet_train_step": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float), P(C.c_uint8), C.c_float,
                                         P(C.c_float)]),
        "cs_unet_capture": (C.c_int, [C.c_void_p, C.c_void_p]),
        "cs_unet_dataset_size": (C.c_int, [C.c_void_p, P(C.c_uint32)]),
        "cs_unet_dataset_clear": (C.c_int, [C.c_void_p]),
        "cs_unet_read_sample": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_float), P(C.c_float), P(C.c_uint8)]),
        "cs_unet_train_epoch": (C.c_int, [C.c_void_p, C.c_float, C.c_uint64, P(C.c_float)]),
        "cs_sim_set_warm_net": (C.c_int, [C.c_void_p, C.c_void_p]),
        "cs_unet_benchmark": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_double), P(C.c_double)]),
        "cs_sim_read_warm_start": (C.c_int, [C.c_void_p, P(C.c_float)]),
        "cs_workload_build_pour": (C.c_int, [P(abi.PourParams), P(Workload)]),
        "cs_batch_default_config": (C.c_int, [P(abi.BatchConfig)]),
        "cs_batch_create": (C.c_int, [C.c_void_p, P(C.c_void_p), C.c_uint32, P(abi.BatchConfig), P(C.c_void_p)]),
        "cs_batch_destroy": (C.c_int, [C.c_void_p]),
        "cs_batch_set_scene_p2g": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int]),
        "cs_batch_step_frame": (C.c_int, [C.c_void_p, C.c_float]),
        "cs_batch_get_stats": (C.c_int, [C.c_void_p, P(abi.BatchStats)]),
        "cs_batch_reset_stats": (C.c_int, [C.c_void_p]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def load_library() -> C.CDLL:
    """Loads the native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m coffee_sim_amd.build` "
                "(there is no fallback implementation)"
            )
        lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_LOCAL | getattr(os, "RTLD_NOW", 2))
        _bind(lib)
        _lib = lib
    return _lib


def _check(rc: int, where: str) -> None:
    if rc != abi.CS_OK:
        msg = load_library().cs_last_error()
        raise CoffeeSimError(rc, where, msg.decode() if msg else "")


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def preset(name: str = "default_v60") -> Settings:
    """`MpmSettings::{default_v60, benchmark_free_stream, benchmark_center_pour}` (mod.rs:106-161)."""
    s = Settings()
    _check(load_library().cs_settings_preset(abi.PRESETS[name], C.byref(s)), "cs_settings_preset")
    return s


class SeededWorkload:
    """A seeded synthetic workload (settings + scene arrays) owned by the library."""

    def __init__(self, name: str, seed: int = 1, scale: int = 1):
        self._w = Workload()
        self.name = name
        if name == "pour":  # built by SeededWorkload.pour
            return
        _check(load_library().cs_workload_build(abi.WORKLOADS[name], seed, scale, C.byref(self._w)),
               "cs_workload_build")

    @classmethod
    def pour(cls, side: int, particles: int, pour_per_substep: int, kettle_angle_deg: float,
             seed: int = 1) -> "SeededWorkload":
        """A pour scene on a cubic grid of `side` cells (a multiple of 16 in
        32..256) over the working scene's 128 mm domain: 70% of `particles` as a
        pre-wetted slab in the bed, the rest poured at `pour_per_substep` from a
        spout tilted `kettle_angle_deg` below horizontal."""
        w = cls("pour")
        p = abi.PourParams(side, particles, pour_per_substep, kettle_angle_deg, seed)
        _check(load_library().cs_workload_build_pour(C.byref(p), C.byref(w._w)), "cs_workload_build_pour")
        return w

    @property
    def settings(self) -> Settings:
        return self._w.settings

    @property
    def scene(self) -> Scene:
        return self._w.scene

    @property
    def frame_dt(self) -> float:
        return float(self._w.frame_dt)

    @property
    def substeps_total(self) -> int:
        return int(self._w.substeps_total)

    @property
    def total_particles(self) -> int:
        return int(self._w.total_particles)

    def arrays(self) -> dict:
        s, sc = self.settings, self.scene
        nc = s.num_cells
        ns = int(s.sdf_res) ** 3
        n = int(sc.num_particles)
        out = {
            "sdf": np.ctypeslib.as_array(sc.sdf, shape=(ns,)).copy(),
            "sdf_class": np.ctypeslib.as_array(sc.sdf_class, shape=(nc,)).copy(),
            "particles": np.ctypeslib.as_array(sc.particles, shape=(n * 8,)).reshape(n, 8).copy(),
        }
        if sc.bed_lookup:
            out["bed_lookup"] = np.ctypeslib.as_array(sc.bed_lookup, shape=(nc,)).copy()
            nr = int(sc.num_bed_records)
            rec = np.ctypeslib.as_array(C.cast(sc.bed_records, C.POINTER(C.c_float)), shape=(nr * 8,))
            out["bed_records"] = rec.reshape(nr, 8).copy()
        return out

    def close(self) -> None:
        if self._w._owner:
            load_library().cs_workload_free(C.byref(self._w))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Handles still open at interpreter exit are closed before the HIP runtime
# tears itself down (simulations first: each holds its context).
_live_sims: "weakref.WeakSet" = weakref.WeakSet()
_live_ctxs: "weakref.WeakSet" = weakref.WeakSet()


@atexit.register
def _close_live_handles() -> None:
    for sim in list(_live_sims):
        sim.close()
    for ctx in list(_live_ctxs):
        ctx.close()


class Context:
    """Device + stream (the wgpu Device/Queue pair the reference borrows, mod.rs:196)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(load_library().cs_ctx_create(device, C.byref(self._h)), "cs_ctx_create")
        self._children: "weakref.WeakSet" = weakref.WeakSet()  # simulations and networks on this context
        _live_ctxs.add(self)

    @property
    def handle(self):
        return self._h

    def device_expf(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty_like(x)
        _check(load_library().cs_device_expf(self._h, _fptr(x), _fptr(out), x.size), "cs_device_expf")
        return out

    def device_sincosf(self, x: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        x = np.ascontiguousarray(x, dtype=np.float32)
        sn, cs_ = np.empty_like(x), np.empty_like(x)
        _check(load_library().cs_device_sincosf(self._h, _fptr(x), _fptr(sn), _fptr(cs_), x.size),
               "cs_device_sincosf")
        return sn, cs_

    def close(self) -> None:
        if self._h:
            # everything created on this context goes first: it uses the context's stream
            for child in list(self._children):
                child.close()
            load_library().cs_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _u8ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


class UNet:
    """The learned pressure warm start (include/coffee_sim.h, cs_unet_*): a 3-D
    U-Net on a cubic grid of `side` cells predicting the converged pressure
    (in units of the pressure scale) from four features per cell."""

    def __init__(self, ctx: Context, side: int, seed: int = 0):
        self.ctx = ctx
        self.side = int(side)
        self.nv = self.side ** 3
        self._h = C.c_void_p()
        _check(load_library().cs_unet_create(ctx.handle, self.side, seed, C.byref(self._h)), "cs_unet_create")
        n = C.c_uint32()
        _check(load_library().cs_unet_param_count(self._h, C.byref(n)), "cs_unet_param_count")
        self.param_count = int(n.value)
        _live_sims.add(self)
        ctx._children.add(self)

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            load_library().cs_unet_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        out = np.empty(self.param_count, np.float32)
        _check(load_library().cs_unet_get_params(self._h, _fptr(out)), "cs_unet_get_params")
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32).reshape(-1)
        if p.size != self.param_count:
            raise ValueError(f"expected {self.param_count} parameters")
        _check(load_library().cs_unet_set_params(self._h, _fptr(p)), "cs_unet_set_params")

    def set_scales(self, rhs_scale: float, pressure_scale: float) -> None:
        _check(load_library().cs_unet_set_scales(self._h, rhs_scale, pressure_scale), "cs_unet_set_scales")

    def scales(self) -> tuple[float, float]:
        a, b = C.c_float(), C.c_float()
        _check(load_library().cs_unet_get_scales(self._h, C.byref(a), C.byref(b)), "cs_unet_get_scales")
        return float(a.value), float(b.value)

    def _features(self, f4: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(f4, np.float32).reshape(-1)
        if f.size != self.nv * 4:
            raise ValueError("features must hold 4 values per cell")
        return f

    def forward(self, f4: np.ndarray) -> np.ndarray:
        out = np.empty(self.nv, np.float32)
        _check(load_library().cs_unet_forward(self._h, _fptr(self._features(f4)), _fptr(out)), "cs_unet_forward")
        return out.reshape((self.side,) * 3)

    def gradient(self, f4: np.ndarray, target: np.ndarray, mask: np.ndarray) -> tuple[float, np.ndarray]:
        t = np.ascontiguousarray(target, np.float32).reshape(-1)
        m = np.ascontiguousarray(mask, np.uint8).reshape(-1)
        loss = C.c_float()
        grad = np.empty(self.param_count, np.float32)
        _check(load_library().cs_unet_gradient(self._h, _fptr(self._features(f4)), _fptr(t), _u8ptr(m),
                                               C.byref(loss), _fptr(grad)), "cs_unet_gradient")
        return float(loss.value), grad

    def train_step(self, f4: np.ndarray, target: np.ndarray, mask: np.ndarray, lr: float) -> float:
        t = np.ascontiguousarray(target, np.float32).reshape(-1)
        m = np.ascontiguousarray(mask, np.uint8).reshape(-1)
        loss = C.c_float()
        _check(load_library().cs_unet_train_step(self._h, _fptr(self._features(f4)), _fptr(t), _u8ptr(m), lr,
                                                 C.byref(loss)), "cs_unet_train_step")
        return float(loss.value)

    def capture(self, sim: "MpmSim3D") -> None:
        _check(load_library().cs_unet_capture(self._h, sim._h), "cs_unet_capture")

    def dataset_size(self) -> int:
        n = C.c_uint32()
        _check(load_library().cs_unet_dataset_size(self._h, C.byref(n)), "cs_unet_dataset_size")
        return int(n.value)

    def clear_dataset(self) -> None:
        _check(load_library().cs_unet_dataset_clear(self._h), "cs_unet_dataset_clear")

    def read_sample(self, i: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        f = np.empty((self.nv, 4), np.float32)
        t = np.empty(self.nv, np.float32)
        m = np.empty(self.nv, np.uint8)
        _check(load_library().cs_unet_read_sample(self._h, i, _fptr(f), _fptr(t), _u8ptr(m)), "cs_unet_read_sample")
        return f, t, m

    def benchmark(self, reps: int = 10) -> tuple[float, float]:
        """Device microseconds of one forward and one forward + backward pass on the last input."""
        f, g = C.c_double(), C.c_double()
        _check(load_library().cs_unet_benchmark(self._h, reps, C.byref(f), C.byref(g)), "cs_unet_benchmark")
        return float(f.value), float(g.value)

    def train_epoch(self, lr: float, seed: int = 0) -> float:
        loss = C.c_float()
        _check(load_library().cs_unet_train_epoch(self._h, lr, seed, C.byref(loss)), "cs_unet_train_epoch")
        return float(loss.value)


class MpmSim3D:
    """The reference's `MpmSim3D` over the C ABI (mod.rs:179-567)."""

    def __init__(self, ctx: Context, settings: Settings, scene: Scene | None = None):
        self.ctx = ctx
        self.settings = settings.copy()
        self._h = C.c_void_p()
        sc = C.byref(scene) if scene is not None else None
        _check(load_library().cs_sim_create(ctx.handle, C.byref(self.settings), sc, C.byref(self._h)),
               "cs_sim_create")
        self.nc = self.settings.num_cells
        self.dims = tuple(int(d) for d in self.settings.grid_dims)
        _live_sims.add(self)
        ctx._children.add(self)

    # ---- lifecycle -----------------------------------------------------------------
    def close(self) -> None:
        if self._h:
            load_library().cs_sim_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self) -> None:
        _check(load_library().cs_sim_reset(self._h), "cs_sim_reset")

    def step_frame(self, dt: float) -> None:
        _check(load_library().cs_sim_step_frame(self._h, C.c_float(dt)), "cs_sim_step_frame")

    def synchronize(self) -> None:
        _check(load_library().cs_sim_synchronize(self._h), "cs_sim_synchronize")

    def recorded_counts(self) -> tuple[int, int]:
        """(substeps run as graph replays, replays finished launch by launch after
        halting at the brick scan) since creation."""
        a, b = C.c_uint64(0), C.c_uint64(0)
        _check(load_library().cs_sim_get_recorded_counts(self._h, C.byref(a), C.byref(b)), "get_recorded_counts")
        return int(a.value), int(b.value)

    # ---- setters / getters (mod.rs:448-538) --------------------------------------------
    def set_kettle_angle(self, angle_deg: float) -> None:
        _check(load_library().cs_sim_set_kettle_angle(self._h, angle_deg), "set_kettle_angle")

    def set_spout_position(self, x: float, y: float, z: float) -> None:
        _check(load_library().cs_sim_set_spout_position(self._h, x, y, z), "set_spout_position")

    def set_spout_target(self, x: float, y: float, z: float) -> None:
        _check(load_library().cs_sim_set_spout_target(self._h, x, y, z), "set_spout_target")

    def counts(self) -> tuple[int, int, int]:
        w, b, m = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(load_library().cs_sim_get_counts(self._h, C.byref(w), C.byref(b), C.byref(m)), "get_counts")
        return w.value, b.value, m.value

    def particle_count(self) -> int:
        w, b, _ = self.counts()
        return w + b

    def water_slots_used(self) -> int:
        return self.counts()[0]

    def times(self) -> tuple[float, float, int]:
        t, m, d = C.c_float(), C.c_float(), C.c_uint32()
        _check(load_library().cs_sim_get_time(self._h, C.byref(t), C.byref(m), C.byref(d)), "get_time")
        return t.value, m.value, d.value

    def total_time(self) -> float:
        return self.times()[0]

    def inflow(self) -> tuple[float, float, float]:
        a, f, e = C.c_float(), C.c_float(), C.c_float()
        _check(load_library().cs_sim_get_inflow(self._h, C.byref(a), C.byref(f), C.byref(e)), "get_inflow")
        return a.value, f.value, e.value

    # ---- data in / out ----------------------------------------------------------------
    def write_particles(self, offset: int, particles: np.ndarray, affine: np.ndarray | None = None) -> None:
        p = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1, 8)
        a = None if affine is None else np.ascontiguousarray(affine, dtype=np.float32).reshape(-1, 12)
        _check(load_library().cs_sim_write_particles(self._h, offset, p.shape[0], _fptr(p),
                                                     None if a is None else _fptr(a)), "write_particles")

    def set_particle_count(self, n: int) -> None:
        _check(load_library().cs_sim_set_particle_count(self._h, n), "set_particle_count")

    def read_particles(self, offset: int = 0, count: int | None = None) -> np.ndarray:
        """Water particles [offset, offset + count) (bed particles live in the scene's own arrays)."""
        if count is None:
            count = self.water_slots_used() - offset
        out = np.empty((count, 8), np.float32)
        _check(load_library().cs_sim_read_particles(self._h, offset, count, _fptr(out)), "read_particles")
        return out

    def read_affine(self, offset: int = 0, count: int | None = None) -> np.ndarray:
        if count is None:
            count = self.water_slots_used() - offset
        out = np.empty((count, 12), np.float32)
        _check(load_library().cs_sim_read_affine(self._h, offset, count, _fptr(out)), "read_affine")
        return out

    def read_p2g_lanes(self) -> np.ndarray:
        out = np.empty((6, self.nc), np.int32)
        _check(load_library().cs_sim_read_p2g_lanes(self._h, _iptr(out)), "read_p2g_lanes")
        return out

    def read_grid_vel(self) -> np.ndarray:
        out = np.empty((self.nc, 4), np.float32)
        _check(load_library().cs_sim_read_grid_vel(self._h, _fptr(out)), "read_grid_vel")
        return out

    def read_pressure(self) -> tuple[np.ndarray, np.ndarray]:
        p = np.empty(self.nc, np.int32)
        k = np.empty(self.nc, np.int32)
        _check(load_library().cs_sim_read_pressure(self._h, _iptr(p), _iptr(k)), "read_pressure")
        return p, k

    def read_bed_lookup(self) -> np.ndarray:
        """Per cell, the highest bed particle that stamped it in the last substep, or -1."""
        out = np.empty(self.nc, np.int32)
        _check(load_library().cs_sim_read_bed_lookup(self._h, out.ctypes.data_as(C.POINTER(C.c_int32))),
               "cs_sim_read_bed_lookup")
        return out

    def read_bed_extracts(self) -> np.ndarray:
        """The bed particles' BedExtracts after the last substep: (n_bed, 8) float32."""
        out = np.zeros((self.counts()[1], 8), np.float32)
        _check(load_library().cs_sim_read_bed_extracts(self._h, _fptr(out)), "cs_sim_read_bed_extracts")
        return out

    def read_bed_particles(self) -> tuple[np.ndarray, np.ndarray]:
        """The bed grains and their affine rows after the last substep: (n_bed, 8) and (n_bed, 12) float32."""
        nb = self.counts()[1]
        part = np.zeros((nb, 8), np.float32)
        aff = np.zeros((nb, 12), np.float32)
        _check(load_library().cs_sim_read_bed_particles(self._h, _fptr(part), _fptr(aff)), "cs_sim_read_bed_particles")
        return part, aff

    def read_sdf(self) -> tuple[np.ndarray, np.ndarray]:
        """The SDF (sdf_res^3) and class mask (one byte per cell) the simulation holds."""
        r = int(self.settings.sdf_res)
        sdf = np.empty(r * r * r, np.float32)
        cls = np.empty(self.nc, np.uint8)
        _check(load_library().cs_sim_read_sdf(self._h, _fptr(sdf), cls.ctypes.data_as(C.POINTER(C.c_uint8))),
               "cs_sim_read_sdf")
        return sdf, cls

    def read_metrics(self) -> np.ndarray:
        out = np.empty(abi.CS_METRICS_SLOTS, np.uint32)
        _check(load_library().cs_sim_read_metrics(self._h, out.ctypes.data_as(C.POINTER(C.c_uint32))),
               "read_metrics")
        return out

    def pack_render(self) -> np.ndarray:
        """prepare_render's RenderParticle per particle in use, bed particles first: (n, 8) float32."""
        out = np.empty((self.particle_count(), 8), np.float32)
        _check(load_library().cs_sim_pack_render(self._h, _fptr(out)), "cs_sim_pack_render")
        return out

    def solve_stats(self) -> SolveStats:
        st = SolveStats()
        _check(load_library().cs_sim_get_solve_stats(self._h, C.byref(st)), "get_solve_stats")
        return st

    def set_warm_start(self, pressure: np.ndarray) -> None:
        p = np.ascontiguousarray(pressure, dtype=np.float32).reshape(-1)
        if p.size != self.nc:
            raise ValueError("warm start must hold one value per cell")
        _check(load_library().cs_sim_set_warm_start(self._h, _fptr(p)), "set_warm_start")

    def set_warm_net(self, net: "UNet | None") -> None:
        """Network used when the settings ask for CS_WARM_LEARNED."""
        self._warm_net = net
        _check(load_library().cs_sim_set_warm_net(self._h, net.handle if net is not None else None),
               "cs_sim_set_warm_net")

    def read_warm_start(self) -> np.ndarray:
        out = np.empty(self.nc, np.float32)
        _check(load_library().cs_sim_read_warm_start(self._h, _fptr(out)), "cs_sim_read_warm_start")
        return out

    def save(self) -> bytes:
        n = C.c_size_t()
        _check(load_library().cs_sim_save(self._h, None, 0, C.byref(n)), "save(size)")
        buf = C.create_string_buffer(n.value)
        _check(load_library().cs_sim_save(self._h, buf, n.value, C.byref(n)), "save")
        return buf.raw[: n.value]

    def load(self, blob: bytes) -> None:
        buf = C.create_string_buffer(blob, len(blob))
        _check(load_library().cs_sim_load(self._h, buf, len(blob)), "load")

    def enable_timing(self, on: bool = True) -> None:
        _check(load_library().cs_sim_enable_timing(self._h, int(on)), "enable_timing")

    def pass_times_us(self) -> np.ndarray:
        out = np.zeros(7, np.float32)
        _check(load_library().cs_sim_get_pass_times(self._h, _fptr(out)), "get_pass_times")
        return out

    def kernel_timing(self) -> abi.KernelTiming:
        kt = abi.KernelTiming()
        _check(load_library().cs_sim_get_kernel_timing(self._h, C.byref(kt)), "get_kernel_timing")
        return kt

    def reset_kernel_timing(self) -> None:
        _check(load_library().cs_sim_reset_kernel_timing(self._h), "reset_kernel_timing")


class Batch:
    """Many simulations advanced together, one launch per pass over every
    scene (cs_batch_*).  The simulations stay owned by the caller and are not
    stepped on their own while the batch exists."""

    def __init__(self, ctx: Context, sims: list, config: abi.BatchConfig | None = None):
        self.ctx = ctx
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[sm._h.value for sm in self.sims])
        self._h = C.c_void_p()
        cfg = C.byref(config) if config is not None else None
        _check(load_library().cs_batch_create(ctx.handle, arr, len(self.sims), cfg, C.byref(self._h)),
               "cs_batch_create")
        _live_sims.add(self)
        ctx._children.add(self)

    @staticmethod
    def default_config() -> abi.BatchConfig:
        c = abi.BatchConfig()
        _check(load_library().cs_batch_default_config(C.byref(c)), "cs_batch_default_config")
        return c

    def set_scene_p2g(self, scene: int, tiled: bool) -> None:
        _check(load_library().cs_batch_set_scene_p2g(self._h, scene, 1 if tiled else 0), "cs_batch_set_scene_p2g")

    def step_frame(self, dt: float) -> None:
        _check(load_library().cs_batch_step_frame(self._h, dt), "cs_batch_step_frame")

    def stats(self) -> abi.BatchStats:
        st = abi.BatchStats()
        _check(load_library().cs_batch_get_stats(self._h, C.byref(st)), "cs_batch_get_stats")
        return st

    def reset_stats(self) -> None:
        _check(load_library().cs_batch_reset_stats(self._h), "cs_batch_reset_stats")

    def close(self) -> None:
        if self._h:
            _check(load_library().cs_batch_destroy(self._h), "cs_batch_destroy")
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def filter_config_default() -> abi.FilterConfig:
    """FilterConfig::default (filter.rs:14-26)."""
    f = abi.FilterConfig()
    load_library().cs_filter_config_default(C.byref(f))
    return f


def bed_config_default() -> abi.BedConfig:
    """BedConfig::default (bed.rs:20-38)."""
    b = abi.BedConfig()
    load_library().cs_bed_config_default(C.byref(b))
    return b


def bed_config_seated_in_filter(filt: abi.FilterConfig | None = None) -> abi.BedConfig:
    """BedConfig::seated_in_filter (bed.rs:41-81); the default filter when none is given."""
    f = filt if filt is not None else filter_config_default()
    b = abi.BedConfig()
    load_library().cs_bed_config_seated_in_filter(C.byref(f), C.byref(b))
    return b


def bed_init(config: abi.BedConfig, grid_dims, bounds_size) -> dict:
    """init_bed_particles (bed.rs:107-206) with build_cell_lookup (bed.rs:208-288):
    'particles' (n, 8), 'affine' (n, 12), 'extracts' (n, 8) and 'cell_lookup' (gx*gy*gz)."""
    L = load_library()
    dims = (C.c_uint32 * 3)(*[int(d) for d in grid_dims])
    bounds = (C.c_float * 3)(*[float(b) for b in bounds_size])
    n = C.c_uint32()
    _check(L.cs_bed_init(C.byref(config), dims, bounds, 0, None, None, None, None, C.byref(n)), "cs_bed_init")
    count = int(n.value)
    part = np.zeros((count, 8), np.float32)
    aff = np.zeros((count, 12), np.float32)
    ext = np.zeros((count, 8), np.float32)
    lookup = np.empty(int(np.prod([int(d) for d in grid_dims])), np.int32)
    _check(L.cs_bed_init(C.byref(config), dims, bounds, max(count, 1), _fptr(part), _fptr(aff), _fptr(ext),
                         lookup.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(n)), "cs_bed_init")
    return {"particles": part, "affine": aff, "extracts": ext, "cell_lookup": lookup}


def attach_bed(scene: Scene | None, bed: dict) -> Scene:
    """A scene carrying `bed` (from bed_init) as its bed of particles; the arrays are
    kept alive on the returned scene."""
    sc = scene if scene is not None else Scene()
    keep = [np.ascontiguousarray(bed["particles"], np.float32), np.ascontiguousarray(bed["affine"], np.float32),
            np.ascontiguousarray(bed["extracts"], np.float32)]
    sc.bed_particles = _fptr(keep[0])
    sc.bed_affine = _fptr(keep[1])
    sc.bed_extracts = _fptr(keep[2])
    sc.num_bed_particles = len(keep[0])
    sc._bed_keep = keep
    return sc


def generate_sdf(settings: Settings) -> np.ndarray:
    out = np.empty(int(settings.sdf_res) ** 3, np.float32)
    _check(load_library().cs_generate_sdf(C.byref(settings), _fptr(out)), "cs_generate_sdf")
    return out


def generate_sdf_class(settings: Settings, sdf: np.ndarray) -> np.ndarray:
    sdf = np.ascontiguousarray(sdf, dtype=np.float32)
    out = np.empty(settings.num_cells, np.uint8)
    _check(load_library().cs_generate_sdf_class(C.byref(settings), _fptr(sdf),
                                                out.ctypes.data_as(C.POINTER(C.c_uint8))),
           "cs_generate_sdf_class")
    return out


def device_count() -> int:
    n = C.c_int()
    _check(load_library().cs_device_count(C.byref(n)), "cs_device_count")
    return n.value
