"""The C ABI: the library loads, exports every entry point include/coffee_sim.h
declares, and the ctypes mirrors have the C layout.  No device needed."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

import coffee_sim_amd as cs
from coffee_sim_amd import abi

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "coffee_sim.h"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(cs_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "cs_sim_step_frame" in names and "cs_sim_create" in names
    assert len(names) >= 35


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


STRUCTS = {
    "cs_obstacle": abi.Obstacle,
    "cs_spout": abi.Spout,
    "cs_geometry": abi.Geometry,
    "cs_settings": abi.Settings,
    "cs_bed_record": abi.BedRecord,
    "cs_scene": abi.Scene,
    "cs_solve_stats": abi.SolveStats,
    "cs_workload": abi.Workload,
    "cs_kernel_timing": abi.KernelTiming,
}


def test_ctypes_layout_matches_c(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'printf("{cname}.{field} %zu\\n", offsetof({cname}, {field}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    c_vals = dict(line.rsplit(" ", 1) for line in out if line)
    for cname, py in STRUCTS.items():
        assert int(c_vals[cname]) == C.sizeof(py), cname
        for field, _ in py._fields_:
            assert int(c_vals[f"{cname}.{field}"]) == getattr(py, field).offset, f"{cname}.{field}"


def _build_caller(tmp_path) -> Path:
    """tests/abi_caller.c, a plain C host, linked against the built library."""
    exe = tmp_path / "abi_caller"
    libdir = cs.LIB_PATH.parent
    subprocess.run(["gcc", "-std=c11", "-O1", "-Wall", "-Werror", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "abi_caller.c"), "-o", str(exe), f"-L{libdir}", "-lcoffee_sim",
                    f"-Wl,-rpath,{libdir}", "-lm"], check=True)
    return exe


def test_c_program_drives_the_host_entry_points(lib, tmp_path):
    """A C program built against include/coffee_sim.h alone calls the presets,
    the seeded workload, the bed lattice and the SDF generator."""
    out = subprocess.run([str(_build_caller(tmp_path))], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "abi caller ok: host part" in out.stdout


@pytest.mark.gpu
def test_c_program_steps_a_simulation(lib, tmp_path):
    """The same C program creates a context and a simulation, steps it,
    checkpoints and restores it, and reads the solve statistics."""
    out = subprocess.run([str(_build_caller(tmp_path)), "run"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert "fluid cells" in out.stdout, out.stdout


def test_preset_errors_are_reported(lib):
    s = abi.Settings()
    rc = lib.cs_settings_preset(7, C.byref(s))
    assert rc == abi.CS_ERR_INVALID
    assert lib.cs_last_error()


def test_workload_rejects_bad_arguments(lib):
    w = abi.Workload()
    assert lib.cs_workload_build(5, 1, 1, C.byref(w)) == abi.CS_ERR_INVALID
    assert lib.cs_workload_build(0, 1, 0, C.byref(w)) == abi.CS_ERR_INVALID


def test_null_handles_are_rejected_without_a_device(lib):
    assert lib.cs_sim_step_frame(None, C.c_float(0.01)) == abi.CS_ERR_INVALID
    assert lib.cs_sim_read_metrics(None, None) == abi.CS_ERR_INVALID
    assert lib.cs_ctx_destroy(None) == abi.CS_OK


def test_context_creation_fails_cleanly_without_device(lib):
    n = C.c_int()
    lib.cs_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a device is present")
    h = C.c_void_p()
    assert lib.cs_ctx_create(0, C.byref(h)) == abi.CS_ERR_DEVICE
    assert b"device" in lib.cs_last_error().lower()
