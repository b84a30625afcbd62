"""MI355X-native step pipeline for the Box2D v3 C API.

The product is the C-ABI library ``libb2mi.so`` (host driver + gfx950 HIP kernels) in this
directory. ``load()`` opens it through ctypes; it raises if the library has not been built,
and every step raises inside the library (assert hook) when no HIP device is present: there
is no CPU fallback.
"""
from __future__ import annotations

import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libb2mi.so")
SCENES_PATH = os.path.join(HERE, "libb2mi_scenes.so")

_api = None


def load():
    """Returns the ctypes binding (capi.B2Api) of libb2mi.so and libb2mi_scenes.so."""
    global _api
    if _api is None:
        if not (os.path.exists(LIB_PATH) and os.path.exists(SCENES_PATH)):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() first")
        from .capi import B2Api
        _api = B2Api(LIB_PATH, SCENES_PATH)
        _api.bind_product()
    return _api
