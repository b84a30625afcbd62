"""Shared test utilities: the three implementations of the public C API and comparisons.

  product()   box2c_amd/libb2mi.so (HIP path; needs a GPU to step)
  oracle()    oracle/liboracle.so (CPU restatement; the checker)
  oracle64()  oracle/liboracle_f64.so (same, solver in double)
  reference() oracle/_ref/libbox2d_ref.so (the reference built from its own sources), or None
"""
from __future__ import annotations

import os

import numpy as np

from box2c_amd.capi import B2Api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
_cache: dict[str, B2Api | None] = {}


def _get(name, lib, scenes):
    if name not in _cache:
        _cache[name] = B2Api(lib, scenes) if os.path.exists(lib) and os.path.exists(scenes) else None
    return _cache[name]


def oracle() -> B2Api:
    api = _get("oracle", os.path.join(ORACLE_DIR, "liboracle.so"), os.path.join(ORACLE_DIR, "liboracle_scenes.so"))
    assert api is not None, "oracle not built (make -C oracle)"
    return api


def oracle64() -> B2Api:
    api = _get("oracle64", os.path.join(ORACLE_DIR, "liboracle_f64.so"),
               os.path.join(ORACLE_DIR, "liboracle_f64_scenes.so"))
    assert api is not None, "oracle not built (make -C oracle)"
    return api


def reference() -> B2Api | None:
    return _get("ref", os.path.join(ORACLE_DIR, "_ref", "libbox2d_ref.so"),
                os.path.join(ORACLE_DIR, "_ref", "libref_scenes.so"))


_emu_api = None


def product() -> B2Api:
    """The library under test: the HIP build, or with B2MI_TEST_LIBRARY=emu its CPU emulation
    (tests/emu/build_emu.py), which runs the same host driver and kernels on the host."""
    global _emu_api
    if os.environ.get("B2MI_TEST_LIBRARY") == "emu":
        if _emu_api is None:
            emu = os.environ.get("B2MI_EMU_DIR") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu", "_build")
            _emu_api = B2Api(os.path.join(emu, "libb2mi.so"), os.path.join(emu, "libb2mi_scenes.so"))
            _emu_api.bind_product()
        return _emu_api
    import box2c_amd
    return box2c_amd.load()


def masked(contacts: np.ndarray) -> np.ndarray:
    """Contact records with the struct padding byte cleared (it is uninitialized memory in
    the reference's manifold points)."""
    c = contacts.copy()
    c["points"]["_pad"] = 0
    return c


def first_contact_mismatch(a: np.ndarray, b: np.ndarray):
    """None when both contact dumps are bit-identical, else a description of the first
    difference."""
    a, b = masked(a), masked(b)
    if len(a) != len(b):
        return f"contact count {len(a)} != {len(b)}"
    if np.array_equal(a.view(np.uint8), b.view(np.uint8)):
        return None
    for i in range(len(a)):
        if a[i].tobytes() != b[i].tobytes():
            return f"record {i}:\n  {a[i]}\n  {b[i]}"
    return "differs"


def first_body_mismatch(a: np.ndarray, b: np.ndarray):
    if len(a) != len(b):
        return f"body count {len(a)} != {len(b)}"
    if np.array_equal(a.view(np.uint8), b.view(np.uint8)):
        return None
    for i in range(len(a)):
        if a[i].tobytes() != b[i].tobytes():
            return f"body {i}:\n  {a[i]}\n  {b[i]}"
    return "differs"


# scene name, scene kwargs, world-def overrides, steps, substeps
PARITY_SCENES = [
    ("hello", {}, {}, 120, 4),
    ("large_pyramid", {"base": 20}, {}, 60, 4),
    ("pyramid_grid", {"rows": 2, "cols": 3, "base": 10, "seed": 3}, {}, 40, 4),
    ("tumbler", {"count": 400, "seed": 5}, {}, 120, 4),
    ("shape_zoo", {"copies": 8, "seed": 2}, {"enableContinous": False}, 100, 4),
    ("drop_boxes", {"count": 16, "seed": 7, "index": 3}, {"enableContinous": False}, 120, 4),
    ("batch_pyramid", {"seed": 1, "index": 5}, {}, 60, 4),
    ("tilted_stacks", {}, {}, 100, 3),
    ("overflow", {"count": 30}, {}, 60, 4),
    ("overflow_negzero", {"count": 30}, {}, 60, 4),
    ("polygon_zoo", {"copies": 6, "seed": 4}, {"enableContinous": False}, 90, 4),
]


def manifold_bytes(m) -> np.ndarray:
    """Raw bytes of a b2Manifold with the two manifold points' padding bytes cleared."""
    import ctypes
    b = np.frombuffer(ctypes.string_at(ctypes.addressof(m), ctypes.sizeof(m)), dtype=np.uint8).copy()
    return mask_manifold_bytes(b)


def mask_manifold_bytes(b: np.ndarray) -> np.ndarray:
    """Clears the padding byte of each manifold point (offset 47 of each 48-byte point)."""
    b = b.copy()
    b[..., 47] = 0
    b[..., 48 + 47] = 0
    return b


def autotune_steps(api: B2Api, batch: int, trial_steps: int) -> int:
    """The steps b2mi_Batch_Autotune took: one rebuild step and `trial_steps` measured steps for
    each of its 60 shapes, and for each of the all-in-LDS footprint limits (96, 64, 48, 32 KiB)
    below the LDS budget it chose."""
    import ctypes
    b, t, l, f = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    api.lib.b2mi_Batch_GetLaunchShape(batch, ctypes.byref(b), ctypes.byref(t), ctypes.byref(l), ctypes.byref(f))
    fulls = sum(1 for x in (96, 64, 48, 32) if x * 1024 < l.value)
    return (60 + fulls) * (1 + max(trial_steps, 1))


def make_world(api: B2Api, name: str, kw: dict, overrides: dict):
    d = api.default_world_def(enableSleep=False)
    for k, v in overrides.items():
        setattr(d, k, v)
    return api.scene(name, d, **kw)


def stability(bodies: np.ndarray, contacts: np.ndarray, start: np.ndarray | None) -> dict:
    """Aggregate measures of a stack from body and contact dumps: the highest dynamic body
    center, the largest horizontal drift of a dynamic body from `start` (body records of the
    same world at its start; skipped when None), the deepest penetration over the manifold points
    of touching contacts, and the kinetic energy of unit-density half-extent-0.5 boxes
    (m = 1, I = 1/6)."""
    dyn = bodies["type"] == 2
    out = {"top": float(bodies["center"][dyn, 1].max()) if dyn.any() else 0.0}
    if start is not None:
        sd = start["type"] == 2
        out["drift"] = float(np.abs(bodies["center"][dyn, 0] - start["center"][sd, 0]).max())
    touching = contacts["colorIndex"] >= 0
    pc = contacts["pointCount"][touching]
    sep = contacts["points"]["separation"][touching]
    live = np.arange(2)[None, :] < pc[:, None]
    out["penetration"] = float(max(0.0, -sep[live].min())) if live.any() else 0.0
    v, w = bodies["v"][dyn].astype(np.float64), bodies["w"][dyn].astype(np.float64)
    out["kinetic"] = float(0.5 * (v * v).sum() + 0.5 * (w * w).sum() / 6.0)
    return out
