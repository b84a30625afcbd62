"""Step events: contact begin/end touch, hit and body move events (src/world.c:506-667 and
1366-1413, src/solver.c:226-231 and 1670-1723, include/box2d/types.h:876-1005).

The oracle's events are pinned against the reference built from its own sources, step by step,
and against what the contact and body dumps fix on their own: the begin (end) events of a step
are exactly the shape pairs that started (stopped) touching, a move event per awake body with its
transform, and each hit event's pair touching with that point, normal and a speed above the
threshold. The HIP path's events must equal the oracle's bit for bit."""
import numpy as np
import pytest

from box2c_amd.capi import Vec2
from tests import helpers

# scene, kwargs, world-def overrides, steps, hit threshold
EVENT_SCENES = [
    ("drop_boxes", {"count": 16, "seed": 7, "index": 3}, {"enableContinous": False}, 90, 0.5),
    ("shape_zoo", {"copies": 8, "seed": 2}, {"enableContinous": False}, 80, 1.0),
    ("tumbler", {"count": 400, "seed": 5}, {}, 90, 1.0),
    ("pyramid_grid", {"rows": 2, "cols": 3, "base": 10, "seed": 3}, {}, 40, 0.05),
]
IDS = [s[0] for s in EVENT_SCENES]


def make(api, name, kw, overrides, threshold, mixed=True):
    """The scene with hit events on two bodies of three and, when `mixed`, contact events left on
    only for the shapes of every third body (a pair reports when either shape has them)."""
    w, ids = helpers.make_world(api, name, kw, overrides)
    api.lib.b2World_SetHitEventThreshold(w, threshold)
    for i, b in enumerate(ids):
        if i % 3 != 2:
            api.lib.b2Body_EnableHitEvents(b, True)
        if mixed and i % 3 != 0:
            for s in api.body_shapes(b):
                api.lib.b2Shape_EnableContactEvents(s, False)
    return w, ids


def normalized(ev: dict) -> dict:
    """Events with the world index of the ids cleared (each library numbers its worlds)."""
    out = {}
    for k, e in ev.items():
        # field by field: a structured copy leaves struct padding undefined
        a = np.zeros(len(e), dtype=e.dtype)
        for f in e.dtype.names:
            a[f] = e[f]
        for f in ("shapeIdA", "shapeIdB", "bodyId"):
            if a.dtype.names and f in a.dtype.names:
                a[f]["world0"] = 0
        out[k] = a
    return out


def first_event_mismatch(a: dict, b: dict):
    for k in a:
        if len(a[k]) != len(b[k]):
            return f"{k}: {len(a[k])} events != {len(b[k])}"
        if not np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)):
            i = next(i for i in range(len(a[k])) if a[k][i].tobytes() != b[k][i].tobytes())
            return f"{k}[{i}]: {a[k][i]} != {b[k][i]}"
    return None


def run_pair(api_a, api_b, name, kw, overrides, steps, threshold):
    """Steps the scene in both libraries and compares every step's events; returns the totals."""
    wa, _ = make(api_a, name, kw, overrides, threshold)
    wb, _ = make(api_b, name, kw, overrides, threshold)
    totals = {"begin": 0, "end": 0, "hit": 0, "move": 0}
    try:
        for s in range(steps):
            api_a.step(wa)
            api_b.step(wb)
            ea, eb = normalized(api_a.events(wa)), normalized(api_b.events(wb))
            m = first_event_mismatch(ea, eb)
            assert m is None, f"{name} step {s}: {m}"
            for k in totals:
                totals[k] += len(ea[k])
    finally:
        api_a.destroy(wa)
        api_b.destroy(wb)
    return totals


@pytest.mark.parametrize("name,kw,overrides,steps,threshold", EVENT_SCENES, ids=IDS)
def test_oracle_events_match_reference(name, kw, overrides, steps, threshold):
    ref = helpers.reference()
    if ref is None:
        pytest.skip("reference not built (oracle/Makefile needs the reference sources)")
    totals = run_pair(helpers.oracle(), ref, name, kw, overrides, steps, threshold)
    assert totals["begin"] > 0 and totals["move"] > 0, totals
    if name in ("drop_boxes", "pyramid_grid"):
        assert totals["hit"] > 0, totals
    if name in ("tumbler", "drop_boxes", "shape_zoo"):
        assert totals["end"] > 0, totals


def _touching(api, w):
    c = api.dump_contacts(w)
    t = c[(c["simFlags"] & 0x00010000) != 0]
    return {(int(a), int(b)) for a, b in zip(t["shapeIdA"], t["shapeIdB"])}, c


@pytest.mark.parametrize("name,kw,overrides,steps,threshold", EVENT_SCENES, ids=IDS)
def test_oracle_events_follow_contacts(name, kw, overrides, steps, threshold):
    """With contact events on for every shape: begin/end events = the pairs that started/stopped
    touching over the step, as the contact dumps show them; move events = the awake bodies with
    their dumped transforms; hit events on touching pairs with the manifold's normal and one of its
    points, faster than the threshold."""
    api = helpers.oracle()
    w, _ = make(api, name, kw, overrides, threshold, mixed=False)
    try:
        before = set()
        for s in range(steps):
            api.step(w)
            now, contacts = _touching(api, w)
            ev = api.events(w)
            pair = lambda e: (int(e["shapeIdA"]["index1"]) - 1, int(e["shapeIdB"]["index1"]) - 1)
            assert {pair(e) for e in ev["begin"]} == now - before, f"step {s}"
            assert {pair(e) for e in ev["end"]} == before - now, f"step {s}"
            assert len(ev["begin"]) == len(now - before) and len(ev["end"]) == len(before - now)
            bodies = api.dump_bodies(w)
            awake = bodies[bodies["setIndex"] == 2]
            assert len(ev["move"]) == len(awake)
            by_id = {int(b["bodyId"]): b for b in awake}
            for e in ev["move"]:
                b = by_id[int(e["bodyId"]["index1"]) - 1]
                assert np.array_equal(e["transform"].view(np.uint32), np.concatenate([b["p"], b["q"]]).view(np.uint32))
            by_pair = {(int(c["shapeIdA"]), int(c["shapeIdB"])): c for c in contacts}
            for e in ev["hit"]:
                c = by_pair[pair(e)]
                assert pair(e) in now and e["approachSpeed"] > threshold
                assert np.array_equal(e["normal"], c["normal"])
                pts = [tuple(c["points"][k]["point"]) for k in range(c["pointCount"])]
                assert tuple(e["point"]) in pts
            before = now
    finally:
        api.destroy(w)


@pytest.mark.gpu
@pytest.mark.parametrize("name,kw,overrides,steps,threshold", EVENT_SCENES, ids=IDS)
def test_events_match_oracle(name, kw, overrides, steps, threshold):
    totals = run_pair(helpers.product(), helpers.oracle(), name, kw, overrides, steps, threshold)
    assert totals["begin"] > 0 and totals["move"] > 0, totals


@pytest.mark.gpu
def test_events_taken_before_an_edit():
    """The move events are those of the step even when a body is moved before they are read, and
    a step with a zero time step clears them (src/world.c:685-699)."""
    gpu, cpu = helpers.product(), helpers.oracle()
    name, kw, overrides, steps, threshold = EVENT_SCENES[0]
    wg, ig = make(gpu, name, kw, overrides, threshold)
    wc, ic = make(cpu, name, kw, overrides, threshold)
    try:
        for _ in range(20):
            gpu.step(wg)
            cpu.step(wc)
        for api, w, ids in ((gpu, wg, ig), (cpu, wc, ic)):
            api.lib.b2Body_SetLinearVelocity(ids[-1], Vec2(0.0, 5.0))
        m = first_event_mismatch(normalized(gpu.events(wg)), normalized(cpu.events(wc)))
        assert m is None, m
        gpu.step(wg, dt=0.0)
        cpu.step(wc, dt=0.0)
        eg, ec = gpu.events(wg), cpu.events(wc)
        assert all(len(eg[k]) == 0 and len(ec[k]) == 0 for k in eg)
    finally:
        gpu.destroy(wg)
        cpu.destroy(wc)


@pytest.mark.gpu
def test_batch_events_match_oracle():
    """Each world of a batch keeps its own events, as if it had been stepped alone."""
    gpu, cpu = helpers.product(), helpers.oracle()
    batch = gpu.create_batch()
    gpu.use_batch(batch)
    worlds_g, worlds_c = [], []
    try:
        for i in range(6):
            wg, ig = gpu.scene("batch_pyramid", seed=11, index=i)
            wc, ic = cpu.scene("batch_pyramid", seed=11, index=i)
            for api, w, ids in ((gpu, wg, ig), (cpu, wc, ic)):
                api.lib.b2World_SetHitEventThreshold(w, 0.05)
                for b in ids[::2]:
                    api.lib.b2Body_EnableHitEvents(b, True)
            worlds_g.append(wg)
            worlds_c.append(wc)
    finally:
        gpu.use_batch(None)
    try:
        totals = {"begin": 0, "hit": 0}
        for s in range(40):
            gpu.batch_step(batch)
            for wc in worlds_c:
                cpu.step(wc)
            for i, (wg, wc) in enumerate(zip(worlds_g, worlds_c)):
                eg, ec = normalized(gpu.events(wg)), normalized(cpu.events(wc))
                m = first_event_mismatch(eg, ec)
                assert m is None, f"world {i} step {s}: {m}"
                totals["begin"] += len(eg["begin"])
                totals["hit"] += len(eg["hit"])
        assert totals["begin"] > 0 and totals["hit"] > 0, totals
    finally:
        gpu.lib.b2mi_DestroyBatch(batch)
        for wc in worlds_c:
            cpu.destroy(wc)
