"""Build libfmd.so (gfx950 HIP kernels + C ABI) and the C++ drivers, in-tree.

    python -m fuller_md_en_amd.build        # or fuller_md_en_amd.build.build()

Every HIP source is compiled with hipcc for gfx950 only, packed-FP32 code
generation disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``).
Outputs: ``fuller_md_en_amd/libfmd.so`` and ``fuller_md_en_amd/bin/*``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
BIN = os.path.join(PKG, "bin")
LIB = os.path.join(PKG, "libfmd.so")

ARCH = os.environ.get("FMD_OFFLOAD_ARCH", "gfx950")
HIP_SOURCES = ["rigid_kernels.hip", "fmd_api.hip", "fmd_atomic_api.hip", "atomic_kernels.hip",
               "surrogate_kernels.hip", "fmd_surrogate_api.hip", "mm_kernels.hip",
               "fmd_mm_api.hip"]
HOST_SOURCES = ["fmd_host.cpp"]
DRIVERS = {"fuller_LJ_npt_md": "driver_lj.cpp", "fuller_airebo_npt_md": "driver_airebo.cpp",
           "fuller_LJ_npt_mmmd": "driver_mmmd.cpp", "fuller_LJ_core": "driver_core.cpp"}


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm required to build libfmd.so)")


def hip_flags() -> list[str]:
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall",
            "-Wno-unused-function", "-Xclang", "-target-feature", "-Xclang",
            "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        jobs.append([hipcc, *hip_flags(), "-c", path, "-o", obj])
    for src in HOST_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        jobs.append(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-I",
                     os.path.join(ROOT, "include"), "-c", path, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for f in [ex.submit(_run, j) for j in jobs]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    drv_jobs = []
    for name, src in DRIVERS.items():
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        drv_jobs.append(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
                         path, "-o", os.path.join(BIN, name), "-L", PKG, "-lfmd",
                         "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"])
    with cf.ThreadPoolExecutor(max_workers=max(1, len(drv_jobs))) as ex:
        for f in [ex.submit(_run, j) for j in drv_jobs]:
            f.result()
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
