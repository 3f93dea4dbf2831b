"""Builds the native library of the MI355X MLS-MPM substep in-tree.

Host code (scene generation, C ABI, orchestration) is compiled with g++; the
kernels with hipcc for gfx950.  Both use binary32 without FMA contraction so
host-side precomputation and device passes round exactly like the reference's
arithmetic.  Output: coffee_sim_amd/lib/libcoffee_sim.so.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB_DIR = HERE / "lib"
LIB = LIB_DIR / "libcoffee_sim.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = os.environ.get("CS_OFFLOAD_ARCH", "gfx950")

HOST_SOURCES = ["cs_scene.cpp", "cs_bedinit.cpp", "cs_workload.cpp", "cs_sim.cpp", "cs_unet.cpp", "cs_batch.cpp"]
DEVICE_SOURCES = ["cs_kernels.hip", "cs_solver.hip", "cs_rbgs_persist.hip", "cs_p2g.hip", "cs_unet_kernels.hip",
                  "cs_noalloc.hip", "cs_mg.hip"]
HEADERS = ["cs_common.h", "cs_bed.h", "cs_device.h", "cs_host.h", "cs_kernels.h", "cs_mg.h", "cs_persist.h", "cs_solver.h"]

HOST_FLAGS = [
    "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}", "-Wall", "-Wno-unused-function",
]
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line
    )
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def build(verbose: bool = False) -> Path:
    """Compiles every source (always, even when the library exists) and links it."""
    obj_dir = HERE / "_build"
    obj_dir.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    jobs = []
    for src in HOST_SOURCES:
        jobs.append(["g++", *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(obj_dir / (src + ".o"))])
    for src in DEVICE_SOURCES:
        jobs.append([str(ROCM / "bin" / "hipcc"), *DEVICE_FLAGS, "-c", str(CSRC / src), "-o",
                     str(obj_dir / (src + ".o"))])
    objs = [obj_dir / (src + ".o") for src in HOST_SOURCES + DEVICE_SOURCES]
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "8"))))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for f in [pool.submit(_run, cmd) for cmd in jobs]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([str(ROCM / "bin" / "hipcc"), f"--offload-arch={ARCH}", "-shared", "-fPIC",
          *[str(o) for o in objs], "-o", str(tmp), f"-L{ROCM / 'lib'}", "-lamdhip64",
          f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
