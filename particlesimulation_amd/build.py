"""Build recipe of the native libraries (hipcc for gfx950, no CMake).

    python -m particlesimulation_amd.build      # or __graft_entry__.build()

Produces, in-tree:
    particlesimulation_amd/libpmgpu.so   HIP kernels + the C ABI of include/pmgpu.h
    particlesimulation_amd/libpmsim.so   C++ drop-in classes (include/pmsim/*.h)
    tests/dropin/dropin_check            programs using the drop-in classes (test inputs)
    tests/dropin/galaxy_timing
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")
HIP_SOURCES = ["api.hip", "sort.hip", "deposit.hip", "fft.hip", "greens.hip", "interp.hip", "p3m.hip",
               "mlp.hip", "diag.hip", "train.hip", "batch.hip", "dft.hip"]
ARCH = os.environ.get("PMGPU_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 (v_pk_*_f32) is disabled for all device code.
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
             "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
             "-Wno-unused-command-line-argument", "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_pmgpu(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for s in HIP_SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(BUILD, s.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if out.strip():
                sys.stderr.write(out)
    lib = os.path.join(HERE, "libpmgpu.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, lib)
    return lib


def build_pmsim() -> str | None:
    """C++ drop-in host classes, linked against libpmgpu.so."""
    if not os.path.isdir(HOST):
        return None
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    if not srcs:
        return None
    lib = os.path.join(HERE, "libpmsim.so")
    tmp = lib + ".tmp"
    cxx = shutil.which("g++") or "g++"
    _run([cxx, "-O2", "-fPIC", "-shared", "-std=c++20", "-I", os.path.join(ROOT, "include"), *srcs,
          "-L", HERE, "-lpmgpu", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, lib)
    return lib


def build_dropin_check(name: str = "dropin_check") -> str | None:
    """A C++ program the drop-in tests run (tests/dropin/<name>.cpp)."""
    src = os.path.join(ROOT, "tests", "dropin", name + ".cpp")
    if not os.path.isfile(src):
        return None
    exe = src[:-4]
    cxx = shutil.which("g++") or "g++"
    _run([cxx, "-O2", "-std=c++20", "-I", os.path.join(ROOT, "include"), src, "-L", HERE, "-lpmsim",
          "-lpmgpu", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/../../particlesimulation_amd", "-o",
          exe + ".tmp"])
    os.replace(exe + ".tmp", exe)
    return exe


def build_all() -> None:
    build_pmgpu()
    build_pmsim()
    build_dropin_check()
    build_dropin_check("galaxy_timing")


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(HERE, "libpmgpu.so"))
