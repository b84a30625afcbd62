#!/usr/bin/env python3
"""Writes the golden fixtures of this directory from the reference built from its own sources
(oracle/_ref/libbox2d_ref.so, made by `make -C oracle` where the reference tree is present).

Fixtures (inputs and the reference's outputs, as raw little-endian struct bytes):
  manifolds.npz   seeded random shape pairs for each public manifold function of
                  include/box2d/collision.h (circles, capsules, segments, boxes, offset and
                  rounded boxes at random poses, from separated to deeply overlapping), and
                  seeded random segment pairs for b2SegmentDistance;
  scenes.npz      for each stepped scene: the body records after N steps, and for the small
                  scenes the contact records (manifold points with feature ids, anchors,
                  separations and accumulated impulses, graph colour and flags);
  tilted_stacks.npz  the reference's determinism test scene (test/test_determinism.c):
                  final positions and rotations after 100 steps of 3 substeps.

Run: python tests/golden/make_golden.py   (never at test time; the fixtures are committed)
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from box2c_amd import capi  # noqa: E402

# scene, scene kwargs, world-def overrides, steps, substeps, keep contacts
SCENES = [
    ("hello", {}, {}, 120, 4, True),
    ("large_pyramid", {"base": 20}, {}, 60, 4, False),
    ("pyramid_grid", {"rows": 2, "cols": 3, "base": 10, "seed": 3}, {}, 40, 4, False),
    ("tumbler", {"count": 400, "seed": 5}, {}, 120, 4, False),
    ("shape_zoo", {"copies": 8, "seed": 2}, {"enableContinous": False}, 100, 4, True),
    ("drop_boxes", {"count": 16, "seed": 7, "index": 3}, {"enableContinous": False}, 120, 4, True),
    ("batch_pyramid", {"seed": 1, "index": 5}, {}, 60, 4, False),
    ("tilted_stacks", {}, {}, 100, 3, True),
    ("overflow", {"count": 30}, {}, 60, 4, True),
    ("overflow_negzero", {"count": 30}, {}, 60, 4, True),
    ("polygon_zoo", {"copies": 6, "seed": 4}, {"enableContinous": False}, 90, 4, True),
]

PAIRS_PER_FUNCTION = 96
SEGMENT_PAIRS = 256


def struct_bytes(s) -> np.ndarray:
    return np.frombuffer(ctypes.string_at(ctypes.addressof(s), ctypes.sizeof(s)), dtype=np.uint8).copy()


def random_shape(api, kind, rng):
    u = lambda lo, hi: float(np.float32(rng.uniform(lo, hi)))  # noqa: E731
    if kind is capi.Circle:
        return capi.Circle(capi.Vec2(u(-0.5, 0.5), u(-0.5, 0.5)), u(0.05, 1.0))
    if kind is capi.Capsule:
        return capi.Capsule(capi.Vec2(u(-1, 0), u(-0.5, 0.5)), capi.Vec2(u(0, 1), u(-0.5, 0.5)), u(0.05, 0.5))
    if kind is capi.Segment:
        return capi.Segment(capi.Vec2(u(-1.5, 0), u(-1, 1)), capi.Vec2(u(0, 1.5), u(-1, 1)))
    r = rng.integers(3)
    if r == 0:
        return api.lib.b2MakeBox(u(0.1, 1.0), u(0.1, 1.0))
    if r == 1:
        return api.lib.b2MakeOffsetBox(u(0.1, 1.0), u(0.1, 1.0), capi.Vec2(u(-0.5, 0.5), u(-0.5, 0.5)), u(-3, 3))
    return api.lib.b2MakeRoundedBox(u(0.1, 0.8), u(0.1, 0.8), u(0.02, 0.3))


def random_transform(rng, spread):
    a = np.float32(rng.uniform(-np.pi, np.pi))
    return capi.Transform(capi.Vec2(float(np.float32(rng.uniform(-spread, spread))),
                                    float(np.float32(rng.uniform(-spread, spread)))),
                          capi.Rot(float(np.cos(a, dtype=np.float32)), float(np.sin(a, dtype=np.float32))))


def transform_bytes(xf) -> np.ndarray:
    return struct_bytes(xf)


def make_manifolds(api):
    rng = np.random.default_rng(20240601)
    out = {}
    for name, ka, kb in capi.COLLIDE_FUNCTIONS:
        fn = getattr(api.lib, name)
        sa, xa, sb, xb, mf = [], [], [], [], []
        for i in range(PAIRS_PER_FUNCTION):
            a, b = random_shape(api, ka, rng), random_shape(api, kb, rng)
            xfA = random_transform(rng, 3.0)
            # B placed near A: from deep overlap to a little beyond the speculative distance
            xfB = random_transform(rng, 0.0)
            xfB.p = capi.Vec2(float(np.float32(xfA.p.x + rng.uniform(-1.3, 1.3))),
                              float(np.float32(xfA.p.y + rng.uniform(-1.3, 1.3))))
            m = fn(ctypes.byref(a), xfA, ctypes.byref(b), xfB)
            sa.append(struct_bytes(a))
            sb.append(struct_bytes(b))
            xa.append(transform_bytes(xfA))
            xb.append(transform_bytes(xfB))
            mf.append(struct_bytes(m))
        out[f"{name}__shapeA"] = np.stack(sa)
        out[f"{name}__xfA"] = np.stack(xa)
        out[f"{name}__shapeB"] = np.stack(sb)
        out[f"{name}__xfB"] = np.stack(xb)
        out[f"{name}__manifold"] = np.stack(mf)
    pts = rng.uniform(-2, 2, size=(SEGMENT_PAIRS, 4, 2)).astype(np.float32)
    # a few degenerate and parallel cases
    pts[0, 1] = pts[0, 0]
    pts[1, 3] = pts[1, 2]
    pts[2, 1] = pts[2, 0]
    pts[2, 3] = pts[2, 2]
    pts[3, 2:] = pts[3, :2] + np.float32(0.5)
    res = []
    for p in pts:
        r = api.lib.b2SegmentDistance(*[capi.Vec2(float(x), float(y)) for x, y in p])
        res.append(struct_bytes(r))
    out["segment_distance__points"] = pts
    out["segment_distance__result"] = np.stack(res)
    np.savez_compressed(os.path.join(HERE, "manifolds.npz"), **out)


def make_scenes(api):
    out = {}
    for name, kw, overrides, steps, substeps, keep_contacts in SCENES:
        d = api.default_world_def(enableSleep=False)
        for k, v in overrides.items():
            setattr(d, k, v)
        w, _ = api.scene(name, d, **kw)
        api.step(w, substeps=substeps, n=steps)
        bodies = api.dump_bodies(w)
        out[f"{name}__bodies"] = bodies.view(np.uint8).reshape(len(bodies), -1)
        if keep_contacts:
            c = api.dump_contacts(w)
            c["points"]["_pad"] = 0
            out[f"{name}__contacts"] = c.view(np.uint8).reshape(len(c), -1)
        api.destroy(w)
    np.savez_compressed(os.path.join(HERE, "scenes.npz"), **out)


def make_tilted_stacks(api):
    d = api.default_world_def(enableSleep=False)
    w, ids = api.scene("tilted_stacks", d)
    api.step(w, substeps=3, n=100)
    p = np.array([[api.lib.b2Body_GetPosition(i).x, api.lib.b2Body_GetPosition(i).y] for i in ids], np.float32)
    q = np.array([[api.lib.b2Body_GetRotation(i).c, api.lib.b2Body_GetRotation(i).s] for i in ids], np.float32)
    api.destroy(w)
    np.savez_compressed(os.path.join(HERE, "tilted_stacks.npz"), positions=p, rotations=q)


def main():
    ref_dir = os.path.join(ROOT, "oracle", "_ref")
    api = capi.B2Api(os.path.join(ref_dir, "libbox2d_ref.so"), os.path.join(ref_dir, "libref_scenes.so"))
    make_manifolds(api)
    make_scenes(api)
    make_tilted_stacks(api)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
