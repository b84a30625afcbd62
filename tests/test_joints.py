"""Joints (revolute and weld; the graph solver's joint stages).

CPU: the oracle's joint restatement against the reference built from its own sources, step by
step on the reference's own joint benchmarks (benchmark/tumbler.c, benchmark/joint_grid.c) and a
joint zoo (limits, springs, motors, static and kinematic partners, collideConnected, overflow
joints); properties a plausible mistake breaks (anchors held together, a motor reaching its speed);
and the kernels' arc tangent (fdlibm) against the C library's atan2f.

GPU: the HIP path against the oracle, bit for bit after every step, in every solver store and
large-island mode, with joint edits between steps."""
import ctypes
import math

import numpy as np
import pytest

from tests import helpers

REVOLUTE, WELD = 4, 5
ZOO_TYPES = 0x7F  # every joint type

# scene, scene kwargs, world-def overrides, steps
JOINT_SCENES = [
    pytest.param("joint_tumbler", {"grid": 12}, {}, 150, id="tumbler12"),
    pytest.param("joint_zoo", {"spokes": 16, "types": ZOO_TYPES, "seed": 3}, {"enableContinous": False}, 150, id="zoo16"),
    pytest.param("joint_grid", {"N": 12, "gravity": -10.0}, {}, 120, id="grid12g"),
    pytest.param("joint_grid", {"N": 10, "gravity": 0.0}, {}, 30, id="grid10"),
]


def _joint_bits(api, joints):
    return api.joint_dump(joints).view(np.uint8)


def _compare_run(a, b, name, kw, ov, steps):
    wa, ba = helpers.make_world(a, name, kw, ov)
    wb, bb = helpers.make_world(b, name, kw, ov)
    ja, jb = a.joints_of(ba), b.joints_of(bb)
    assert len(ja) == len(jb) > 0
    assert list(a.counters(wa).colorCounts) == list(b.counters(wb).colorCounts)
    for s in range(steps):
        a.step(wa)
        b.step(wb)
        m = helpers.first_body_mismatch(a.dump_bodies(wa), b.dump_bodies(wb))
        assert m is None, f"{name} step {s}: {m}"
        m = helpers.first_contact_mismatch(a.dump_contacts(wa), b.dump_contacts(wb))
        assert m is None, f"{name} step {s}: {m}"
        assert np.array_equal(_joint_bits(a, ja), _joint_bits(b, jb)), f"{name} step {s}: joint impulses differ"
    assert list(a.counters(wa).colorCounts) == list(b.counters(wb).colorCounts)
    a.destroy(wa)
    b.destroy(wb)


# ---------------------------------------------------------------------------------------------
# CPU: pinning the oracle

@pytest.mark.parametrize("name,kw,ov,steps", JOINT_SCENES)
def test_oracle_joints_match_reference(name, kw, ov, steps):
    ref = helpers.reference()
    if ref is None:
        pytest.skip("reference build not available")
    _compare_run(helpers.oracle(), ref, name, kw, ov, steps)


def test_zoo_has_overflow_joints_and_all_partners():
    """The zoo exercises what the joint path must handle: overflow joints (the hub's), joints to
    static and kinematic bodies and joints among dynamic bodies."""
    o = helpers.oracle()
    w, bodies = helpers.make_world(o, "joint_zoo", {"spokes": 16, "types": ZOO_TYPES, "seed": 3}, {"enableContinous": False})
    c = o.counters(w)
    assert c.colorCounts[11] >= 5  # overflow
    types = set()
    for j in o.joints_of(bodies):
        for b in (o.lib.b2Joint_GetBodyA(j), o.lib.b2Joint_GetBodyB(j)):
            types.add(o.lib.b2Body_GetType(b))
    assert types == {0, 1, 2}
    o.destroy(w)


def _anchor_gap(api, joint):
    xa = api.lib.b2Body_GetTransform(api.lib.b2Joint_GetBodyA(joint))
    xb = api.lib.b2Body_GetTransform(api.lib.b2Joint_GetBodyB(joint))
    la, lb = api.lib.b2Joint_GetLocalAnchorA(joint), api.lib.b2Joint_GetLocalAnchorB(joint)
    pa = (xa.q.c * la.x - xa.q.s * la.y + xa.p.x, xa.q.s * la.x + xa.q.c * la.y + xa.p.y)
    pb = (xb.q.c * lb.x - xb.q.s * lb.y + xb.p.x, xb.q.s * lb.x + xb.q.c * lb.y + xb.p.y)
    return math.hypot(pa[0] - pb[0], pa[1] - pb[1])


def test_oracle_joints_hold_their_anchors():
    """Revolute and rigid weld joints keep their two anchors together (a sign or index mistake in
    the point constraint pulls them apart within a few steps): the tumbler's pivot, a hanging
    net of 264 joints, and the zoo's chains (its one soft 5 Hz weld sags further)."""
    o = helpers.oracle()
    for name, kw, ov, limit, softLimit in (
            ("joint_tumbler", {"grid": 12}, {}, 1e-3, 1e-3),
            ("joint_grid", {"N": 12, "gravity": -10.0}, {}, 0.03, 0.03),
            ("joint_zoo", {"spokes": 0, "types": ZOO_TYPES, "seed": 3}, {"enableContinous": False}, 0.03, 0.3)):
        w, bodies = helpers.make_world(o, name, kw, ov)
        joints = [j for j in o.joints_of(bodies) if o.lib.b2Joint_GetType(j) in (REVOLUTE, WELD)]
        o.step(w, n=120)
        gaps = sorted(_anchor_gap(o, j) for j in joints)
        assert gaps[max(0, len(gaps) - 2)] < limit and gaps[-1] < softLimit, (name, gaps[-3:])
        o.destroy(w)


def test_oracle_motor_reaches_its_speed():
    """The joint tumbler's motor (1e8 N m limit) turns the container at its set speed of 25 deg/s."""
    o = helpers.oracle()
    w, bodies = helpers.make_world(o, "joint_tumbler", {"grid": 4}, {})
    o.step(w, n=30)
    container = bodies[0]
    wz = o.lib.b2Body_GetAngularVelocity(container)
    assert abs(wz - math.radians(25.0)) < 1e-4, wz
    o.destroy(w)


def _libm_and_kernel_atan2(y, x):
    o, p = helpers.oracle(), helpers.product()
    y = np.ascontiguousarray(y, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    a = np.empty_like(y)
    b = np.empty_like(y)
    fp = ctypes.POINTER(ctypes.c_float)
    o.lib.oracle_Atan2.argtypes = [fp, fp, fp, ctypes.c_int]
    p.lib.b2mi_Atan2.argtypes = [fp, fp, fp, ctypes.c_int]
    args = [v.ctypes.data_as(fp) for v in (y, x)]
    o.lib.oracle_Atan2(*args, a.ctypes.data_as(fp), len(y))
    p.lib.b2mi_Atan2(*args, b.ctypes.data_as(fp), len(y))
    return a, b


def test_kernel_atan2_matches_c_library():
    """The joint kernels' arc tangent gives the host C library's atan2f bits: on rotation-like
    arguments (the sines and cosines b2RelativeAngle forms), on quotients across every branch of
    the argument reduction and the |y/x| cut-offs, and on random bit patterns and special values."""
    rng = np.random.default_rng(12)
    n = 2_000_000
    t = rng.uniform(-4.0, 4.0, n).astype(np.float32)
    scale = rng.choice(np.array([1.0, 0.5, 3.0, 1e-3, 1e3], dtype=np.float32), n)
    ys = [np.sin(t) * scale, rng.uniform(-1, 1, n).astype(np.float32)]
    xs = [np.cos(t), rng.uniform(-1, 1, n).astype(np.float32)]
    e = rng.integers(-40, 41, n)
    xq = rng.uniform(0.01, 100.0, n).astype(np.float32) * rng.choice(np.array([-1, 1], dtype=np.float32), n)
    ys.append((xq * np.ldexp(1.0, e) * rng.uniform(1.0, 2.0, n)).astype(np.float32) * rng.choice(np.array([-1, 1], dtype=np.float32), n))
    xs.append(xq)
    bits = rng.integers(0, 2**32, size=(2, n), dtype=np.uint64).astype(np.uint32)
    ys.append(bits[0].view(np.float32))
    xs.append(bits[1].view(np.float32))
    special = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 3.4e38], dtype=np.float32)
    sy, sx = np.meshgrid(special, special)
    ys.append(sy.ravel())
    xs.append(sx.ravel())
    a, b = _libm_and_kernel_atan2(np.concatenate(ys), np.concatenate(xs))
    same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
    assert same.all(), f"{(~same).sum()} mismatches, first at y={np.concatenate(ys)[~same][0]!r} x={np.concatenate(xs)[~same][0]!r}"


# ---------------------------------------------------------------------------------------------
# GPU: the HIP path against the oracle

@pytest.mark.gpu
@pytest.mark.parametrize("name,kw,ov,steps", JOINT_SCENES)
def test_joint_parity(name, kw, ov, steps):
    _compare_run(helpers.product(), helpers.oracle(), name, kw, ov, steps)


# LDS budgets: 0 keeps every item in HBM, 2 KiB only the bodies of the larger items, the default
# everything in LDS
STORES = [pytest.param(0, id="hbm"), pytest.param(2048, id="bodies_lds"), pytest.param(-1, id="lds")]


@pytest.mark.gpu
@pytest.mark.parametrize("budget", STORES)
def test_joint_parity_stores(budget):
    p = helpers.product()
    o = helpers.oracle()
    kw, ov = {"spokes": 16, "types": ZOO_TYPES, "seed": 5}, {"enableContinous": False}
    wp, bp = helpers.make_world(p, "joint_zoo", kw, ov)
    wo, bo = helpers.make_world(o, "joint_zoo", kw, ov)
    if budget >= 0:
        p.lib.b2mi_World_SetLdsBudget(wp, budget)
    jp, jo = p.joints_of(bp), o.joints_of(bo)
    for s in range(80):
        p.step(wp)
        o.step(wo)
        assert helpers.first_body_mismatch(p.dump_bodies(wp), o.dump_bodies(wo)) is None, s
        assert np.array_equal(_joint_bits(p, jp), _joint_bits(o, jo)), s
    p.destroy(wp)
    o.destroy(wo)


# the tumbler's one large island: one workgroup, one cooperative grid launch (automatic size, or
# three workgroups), one launch per phase
ISLAND_MODES = [pytest.param(None, id="workgroup"), pytest.param(0, id="grid_auto"), pytest.param(3, id="grid3"),
                pytest.param(-1, id="per_phase")]


@pytest.mark.gpu
@pytest.mark.parametrize("grid", ISLAND_MODES)
def test_joint_tumbler_island_modes(grid):
    p = helpers.product()
    o = helpers.oracle()
    wp, bp = helpers.make_world(p, "joint_tumbler", {"grid": 20}, {})
    wo, bo = helpers.make_world(o, "joint_tumbler", {"grid": 20}, {})
    if grid is not None:
        p.lib.b2mi_World_SetLdsBudget(wp, 0)
        p.lib.b2mi_World_SetSplitSlots(wp, 1)
        p.lib.b2mi_World_SetGridBlocks(wp, grid)
    jp, jo = p.joints_of(bp), o.joints_of(bo)
    for s in range(100):
        p.step(wp)
        o.step(wo)
        assert helpers.first_body_mismatch(p.dump_bodies(wp), o.dump_bodies(wo)) is None, s
        assert np.array_equal(_joint_bits(p, jp), _joint_bits(o, jo)), s
    p.destroy(wp)
    o.destroy(wo)


# every zoo item over the grid (a split threshold of one slot): the hub's overflow joints in level
# runs of one workgroup (the default), or each level over the whole grid
ZOO_MODES = [pytest.param(-1, -1, id="per_phase"), pytest.param(-1, 0, id="per_phase-levels_over_grid"),
             pytest.param(0, -1, id="grid_auto"), pytest.param(3, 0, id="grid3-levels_over_grid")]


@pytest.mark.gpu
@pytest.mark.parametrize("grid,run", ZOO_MODES)
def test_joint_zoo_island_modes(grid, run):
    p = helpers.product()
    o = helpers.oracle()
    kw, ov = {"spokes": 16, "types": ZOO_TYPES, "seed": 3}, {"enableContinous": False}
    wp, bp = helpers.make_world(p, "joint_zoo", kw, ov)
    wo, bo = helpers.make_world(o, "joint_zoo", kw, ov)
    p.lib.b2mi_World_SetLdsBudget(wp, 0)
    p.lib.b2mi_World_SetSplitSlots(wp, 1)
    p.lib.b2mi_World_SetGridBlocks(wp, grid)
    p.lib.b2mi_World_SetLevelRunSlots(wp, run)
    jp, jo = p.joints_of(bp), o.joints_of(bo)
    for s in range(60):
        p.step(wp)
        o.step(wo)
        assert helpers.first_body_mismatch(p.dump_bodies(wp), o.dump_bodies(wo)) is None, s
        assert np.array_equal(_joint_bits(p, jp), _joint_bits(o, jo)), s
    items = p.work_items(world=wp)
    modes = set(items[items[:, 0] > 0, 3].tolist())
    assert modes == ({1} if grid < 0 else {2}), modes
    p.destroy(wp)
    o.destroy(wo)


@pytest.mark.gpu
def test_joint_net_runs_over_the_grid():
    """A one-island net of 4,900 joints and no contacts, its state in HBM, counts its joint slots
    toward the large-island threshold of a world stepped alone (4,096 slots), so it runs over the
    grid, one launch per phase, and gives the oracle's bits."""
    p = helpers.product()
    o = helpers.oracle()
    wp, bp = helpers.make_world(p, "joint_grid", {"N": 50, "gravity": -10.0}, {})
    wo, bo = helpers.make_world(o, "joint_grid", {"N": 50, "gravity": -10.0}, {})
    p.lib.b2mi_World_SetLdsBudget(wp, 0)
    jp, jo = p.joints_of(bp), o.joints_of(bo)
    assert len(jp) >= 4096
    for s in range(12):
        p.step(wp)
        o.step(wo)
        assert helpers.first_body_mismatch(p.dump_bodies(wp), o.dump_bodies(wo)) is None, s
        assert np.array_equal(_joint_bits(p, jp), _joint_bits(o, jo)), s
    items = p.work_items(world=wp)
    assert items.shape[0] == 1 and items[0, 0] == len(jp) and items[0, 3] == 1, items
    p.destroy(wp)
    o.destroy(wo)


@pytest.mark.gpu
def test_joint_edits_between_steps():
    """Motor speed, limits and joint destruction between steps, on both implementations."""
    p = helpers.product()
    o = helpers.oracle()
    kw, ov = {"spokes": 14, "types": ZOO_TYPES, "seed": 8}, {"enableContinous": False}
    wp, bp = helpers.make_world(p, "joint_zoo", kw, ov)
    wo, bo = helpers.make_world(o, "joint_zoo", kw, ov)
    jp, jo = p.joints_of(bp), o.joints_of(bo)
    revolute = [i for i, j in enumerate(jp) if p.lib.b2Joint_GetType(j) == REVOLUTE]
    for s in range(90):
        if s == 20:
            for api, js in ((p, jp), (o, jo)):
                api.lib.b2RevoluteJoint_SetMotorSpeed(js[revolute[6]], ctypes.c_float(-3.0))
                api.lib.b2RevoluteJoint_EnableLimit(js[revolute[3]], True)
                api.lib.b2RevoluteJoint_SetLimits(js[revolute[3]], ctypes.c_float(-0.1), ctypes.c_float(0.2))
        if s == 45:
            for api, js in ((p, jp), (o, jo)):
                api.lib.b2DestroyJoint(js[revolute[-1]])
                api.lib.b2DestroyJoint(js[2])
            jp = [j for i, j in enumerate(jp) if i not in (revolute[-1], 2)]
            jo = [j for i, j in enumerate(jo) if i not in (revolute[-1], 2)]
        p.step(wp)
        o.step(wo)
        assert helpers.first_body_mismatch(p.dump_bodies(wp), o.dump_bodies(wo)) is None, s
        assert helpers.first_contact_mismatch(p.dump_contacts(wp), o.dump_contacts(wo)) is None, s
        assert np.array_equal(_joint_bits(p, jp), _joint_bits(o, jo)), s
        assert p.lib.b2RevoluteJoint_GetMotorTorque(jp[revolute[6]]) == o.lib.b2RevoluteJoint_GetMotorTorque(jo[revolute[6]])
    p.destroy(wp)
    o.destroy(wo)
