"""Builds the HIP library of the step pipeline in-tree.

  box2c_amd/libb2mi.so         C-ABI library (host driver + gfx950 kernels), built by hipcc
  box2c_amd/libb2mi_scenes.so  benchmark/test scenes (csrc/scenes.c) linked against it

Floating point: no contraction into FMA (-ffp-contract=off), correctly rounded fp32 division
and square root, packed-fp32 instructions disabled; together these make every kernel
expression round exactly like the reference's scalar and 8-wide AVX2 float code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libb2mi.so")
SCENES = os.path.join(HERE, "libb2mi_scenes.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("B2MI_OFFLOAD_ARCH", "gfx950")

SOURCES = ["kernels.hip", "adjoint.hip", "broadphase.hip", "world.cpp", "joint.cpp", "step.cpp", "batch.cpp", "tape.cpp", "tree.cpp",
           "geometry.cpp", "collide_api.cpp"]

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-disable-promote-alloca-to-lds",
    "-fvisibility=hidden", "-DB2MI_BUILDING_LIBRARY",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose: bool = True, jobs: int | None = None) -> str:
    from concurrent.futures import ThreadPoolExecutor

    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".", "_") + ".o")
        cmd = [HIPCC, *HIP_FLAGS, "-I", CSRC, "-I", os.path.join(ROOT, "include")]
        if src.endswith(".cpp"):
            cmd += ["-x", "hip"]
        cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
        cmds.append(cmd)
        objs.append(obj)
    # the objects compile independently (one hipcc per source, at most 8 at a time)
    with ThreadPoolExecutor(max_workers=jobs or min(8, os.cpu_count() or 1)) as pool:
        for f in [pool.submit(_run, c) for c in cmds]:
            f.result()
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-Wl,-Bsymbolic", "-o", LIB, *objs])
    _run(["gcc", "-O2", "-std=gnu17", "-fPIC", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
          "-shared", "-o", SCENES, os.path.join(CSRC, "scenes.c"), "-L", HERE, "-lb2mi",
          "-Wl,-rpath,$ORIGIN", "-lm"])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
