"""Builds the CPU emulation of the library for tests without a GPU.

  tests/emu/_build/libb2mi.so         the library's sources (host driver and kernels) compiled
                                      as host C++ against the HIP stand-in in tests/emu/hip/
  tests/emu/_build/libb2mi_scenes.so  the scenes, linked against it

With asan=True both are built with AddressSanitizer; the Python process that loads them then
needs the sanitizer runtime preloaded (tests/emu/run_emu.py does that). Test infrastructure
only: the product never loads these files.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "box2c_amd", "csrc")
OUT = os.path.join(HERE, "_build")
sys.path.insert(0, ROOT)
from box2c_amd.build import SOURCES  # noqa: E402


def build(asan: bool = True, opt: str = "-O1", out: str = OUT) -> str:
    os.makedirs(out, exist_ok=True)
    flags = ["-std=c++17", "-fPIC", opt, "-g", "-ffp-contract=off", "-fno-fast-math", "-fvisibility=hidden",
             "-DB2MI_BUILDING_LIBRARY", "-Wno-unknown-pragmas", "-Wno-ignored-attributes"]
    if asan:
        flags += ["-fsanitize=address", "-fno-omit-frame-pointer"]
    inc = ["-I", HERE, "-I", CSRC, "-I", os.path.join(ROOT, "include")]
    objs = []
    for src in SOURCES + ["emu_runtime.cpp"]:
        path = os.path.join(HERE, src) if src == "emu_runtime.cpp" else os.path.join(CSRC, src)
        obj = os.path.join(out, src.replace(".", "_") + ".o")
        cmd = ["g++", "-x", "c++", *flags, *inc, "-c", path, "-o", obj]
        print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    lib = os.path.join(out, "libb2mi.so")
    link = ["g++", "-shared", "-o", lib, *objs, "-lpthread"]
    if asan:
        link.insert(1, "-fsanitize=address")
    subprocess.check_call(link)
    scenes = os.path.join(out, "libb2mi_scenes.so")
    cmd = ["gcc", "-O1", "-g", "-std=gnu17", "-fPIC", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"), "-shared",
           "-o", scenes, os.path.join(CSRC, "scenes.c"), "-L", out, "-lb2mi", "-Wl,-rpath,$ORIGIN", "-lm"]
    if asan:
        cmd.insert(1, "-fsanitize=address")
    subprocess.check_call(cmd)
    return lib


if __name__ == "__main__":
    out = sys.argv[sys.argv.index("--out") + 1] if "--out" in sys.argv else OUT
    build(asan="--no-asan" not in sys.argv, out=out)
