"""Builds the native library of the package in-tree.

``libm3dmt.so`` holds the host-side octree / staggered-grid / operator
producer (C++, g++) and the HIP solve with its C-ABI (hipcc, gfx950), linked
into one shared object under ``maple3dmt_amd/lib/``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
BUILDDIR = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")
LIBNAME = "libm3dmt.so"

HOST_SOURCES = [
    "host/octree_mesh.cpp",
    "host/staggered_grid.cpp",
    "host/operators.cpp",
    "host/fv_problem.cpp",
    "host/forward_solver_gpu.cpp",
    "host/host_capi.cpp",
    "host/edi_io.cpp",
]
DEVICE_SOURCES = [
    "device/kernels.hip",
    "device/solver.hip",
    "device/nn.hip",
    "device/stations.hip",
]

ARCH = os.environ.get("M3D_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HOST_FLAGS = ["-std=c++17", "-O3", "-fPIC", "-Wall", "-Wextra", "-I", INCLUDE]
# Packed FP32 instructions are disabled for every HIP object.
DEVICE_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE, "-Wno-unused-result",
]


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n"
                           f"{proc.stdout}\n{proc.stderr}")
    return proc


def build(verbose: bool = False) -> str:
    """Compile every source and link ``lib/libm3dmt.so``; returns its path."""
    os.makedirs(BUILDDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = []
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILDDIR, os.path.basename(src) + ".o")
        jobs.append([CXX, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in DEVICE_SOURCES:
        obj = os.path.join(BUILDDIR, os.path.basename(src) + ".o")
        jobs.append([HIPCC, *DEVICE_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for proc in ex.map(_run, jobs):
            if verbose and (proc.stdout or proc.stderr):
                sys.stdout.write(proc.stdout + proc.stderr)
    out_tmp = os.path.join(LIBDIR, LIBNAME + ".tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", out_tmp])
    out = os.path.join(LIBDIR, LIBNAME)
    os.replace(out_tmp, out)
    return out


def clean() -> None:
    shutil.rmtree(BUILDDIR, ignore_errors=True)


if __name__ == "__main__":
    print(build(verbose=True))
