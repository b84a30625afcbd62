"""ctypes declarations of the public C ABI (include/box2d/*.h, include/b2mi.h).

`B2Api(lib_path, scenes_path)` binds one shared library that exports the reference's world /
body / shape / step API (this package's libb2mi.so, or any other library with the same ABI)
together with the scene library linked against it. Structures mirror the C layouts exactly
(reference include/box2d/types.h, collision.h, id.h). Libraries are opened RTLD_LOCAL so that
several implementations of the same symbol names can live in one process.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_bool, c_char_p, c_float, c_int, c_int32, c_uint16,
                    c_uint32, c_uint64, c_void_p)

import numpy as np


class Vec2(Structure):
    _fields_ = [("x", c_float), ("y", c_float)]


class Rot(Structure):
    _fields_ = [("c", c_float), ("s", c_float)]


class Transform(Structure):
    _fields_ = [("p", Vec2), ("q", Rot)]


class AABB(Structure):
    _fields_ = [("lowerBound", Vec2), ("upperBound", Vec2)]


class WorldId(Structure):
    _fields_ = [("index1", c_uint16), ("revision", c_uint16)]


class BodyId(Structure):
    _fields_ = [("index1", c_int32), ("world0", c_uint16), ("revision", c_uint16)]


class ShapeId(Structure):
    _fields_ = [("index1", c_int32), ("world0", c_uint16), ("revision", c_uint16)]


class JointId(Structure):
    _fields_ = [("index1", c_int32), ("world0", c_uint16), ("revision", c_uint16)]


class WorldDef(Structure):
    _fields_ = [("gravity", Vec2), ("restitutionThreshold", c_float),
                ("contactPushoutVelocity", c_float), ("hitEventThreshold", c_float),
                ("contactHertz", c_float), ("contactDampingRatio", c_float),
                ("jointHertz", c_float), ("jointDampingRatio", c_float),
                ("enableSleep", c_bool), ("enableContinous", c_bool), ("workerCount", c_int32),
                ("enqueueTask", c_void_p), ("finishTask", c_void_p),
                ("userTaskContext", c_void_p), ("internalValue", c_int32)]


class BodyDef(Structure):
    _fields_ = [("type", c_int32), ("position", Vec2), ("rotation", Rot), ("linearVelocity", Vec2),
                ("angularVelocity", c_float), ("linearDamping", c_float), ("angularDamping", c_float),
                ("gravityScale", c_float), ("sleepThreshold", c_float), ("userData", c_void_p),
                ("enableSleep", c_bool), ("isAwake", c_bool), ("fixedRotation", c_bool),
                ("isBullet", c_bool), ("isEnabled", c_bool), ("automaticMass", c_bool),
                ("internalValue", c_int32)]


class Filter(Structure):
    _fields_ = [("categoryBits", c_uint32), ("maskBits", c_uint32), ("groupIndex", c_int32)]


class ShapeDef(Structure):
    _fields_ = [("userData", c_void_p), ("friction", c_float), ("restitution", c_float),
                ("density", c_float), ("filter", Filter), ("customColor", c_uint32),
                ("isSensor", c_bool), ("enableSensorEvents", c_bool), ("enableContactEvents", c_bool),
                ("enableHitEvents", c_bool), ("enablePreSolveEvents", c_bool),
                ("forceContactCreation", c_bool), ("internalValue", c_int32)]


class Polygon(Structure):
    _fields_ = [("vertices", Vec2 * 8), ("normals", Vec2 * 8), ("centroid", Vec2), ("radius", c_float),
                ("count", c_int32)]


class Circle(Structure):
    _fields_ = [("center", Vec2), ("radius", c_float)]


class Capsule(Structure):
    _fields_ = [("center1", Vec2), ("center2", Vec2), ("radius", c_float)]


class Segment(Structure):
    _fields_ = [("point1", Vec2), ("point2", Vec2)]


class MassData(Structure):
    _fields_ = [("mass", c_float), ("center", Vec2), ("rotationalInertia", c_float)]


class SegmentDistanceResult(Structure):
    _fields_ = [("closest1", Vec2), ("closest2", Vec2), ("fraction1", c_float), ("fraction2", c_float),
                ("distanceSquared", c_float)]


class DistanceProxy(Structure):
    _fields_ = [("points", Vec2 * 8), ("count", c_int), ("radius", c_float)]


class DistanceCache(Structure):
    _fields_ = [("count", ctypes.c_uint16), ("indexA", ctypes.c_uint8 * 3), ("indexB", ctypes.c_uint8 * 3)]


class DistanceInput(Structure):
    _fields_ = [("proxyA", DistanceProxy), ("proxyB", DistanceProxy), ("transformA", Transform), ("transformB", Transform),
                ("useRadii", ctypes.c_bool)]


class DistanceOutput(Structure):
    _fields_ = [("pointA", Vec2), ("pointB", Vec2), ("distance", c_float), ("iterations", c_int), ("simplexCount", c_int)]


class Sweep(Structure):
    _fields_ = [("localCenter", Vec2), ("c1", Vec2), ("c2", Vec2), ("q1", Rot), ("q2", Rot)]


class TOIInput(Structure):
    _fields_ = [("proxyA", DistanceProxy), ("proxyB", DistanceProxy), ("sweepA", Sweep), ("sweepB", Sweep), ("tMax", c_float)]


class TOIOutput(Structure):
    _fields_ = [("state", c_int), ("t", c_float)]


class ManifoldPoint(Structure):
    _fields_ = [("point", Vec2), ("anchorA", Vec2), ("anchorB", Vec2), ("separation", c_float),
                ("normalImpulse", c_float), ("tangentImpulse", c_float),
                ("maxNormalImpulse", c_float), ("normalVelocity", c_float), ("id", c_uint16),
                ("persisted", c_bool)]


class Manifold(Structure):
    _fields_ = [("points", ManifoldPoint * 2), ("normal", Vec2), ("pointCount", c_int32)]


class Counters(Structure):
    _fields_ = [(n, c_int32) for n in (
        "staticBodyCount bodyCount shapeCount contactCount jointCount islandCount stackUsed "
        "staticTreeHeight treeHeight byteCount taskCount").split()] + [("colorCounts", c_int32 * 12)]


class Profile(Structure):
    _fields_ = [(n, c_float) for n in (
        "step pairs collide solve buildIslands solveConstraints prepareTasks solverTasks "
        "prepareConstraints integrateVelocities warmStart solveVelocities integratePositions "
        "relaxVelocities applyRestitution storeImpulses finalizeBodies splitIslands sleepIslands "
        "hitEvents broadphase continuous").split()]


class BodyEvents(Structure):
    _fields_ = [("moveEvents", c_void_p), ("moveCount", c_int32)]


class SensorEvents(Structure):
    _fields_ = [("beginEvents", c_void_p), ("endEvents", c_void_p), ("beginCount", c_int32), ("endCount", c_int32)]


class ContactEvents(Structure):
    _fields_ = [("beginEvents", c_void_p), ("endEvents", c_void_p), ("hitEvents", c_void_p),
                ("beginCount", c_int32), ("endCount", c_int32), ("hitCount", c_int32)]


_ID_DTYPE = np.dtype([("index1", "<i4"), ("world0", "<u2"), ("revision", "<u2")])
TOUCH_EVENT_DTYPE = np.dtype([("shapeIdA", _ID_DTYPE), ("shapeIdB", _ID_DTYPE)])
HIT_EVENT_DTYPE = np.dtype([("shapeIdA", _ID_DTYPE), ("shapeIdB", _ID_DTYPE), ("point", "<f4", 2), ("normal", "<f4", 2),
                            ("approachSpeed", "<f4")])
MOVE_EVENT_DTYPE = np.dtype({"names": ["transform", "bodyId", "userData", "fellAsleep"],
                             "formats": [("<f4", 4), _ID_DTYPE, "<u8", "u1"], "offsets": [0, 16, 24, 32], "itemsize": 40})


def _event_array(ptr, count, dtype):
    """A copy of `count` events at `ptr`, field by field (struct padding reads as zero)."""
    raw = np.zeros(count, dtype=dtype)
    if count > 0:
        ctypes.memmove(raw.ctypes.data, ptr, count * dtype.itemsize)
    out = np.zeros(count, dtype=dtype)
    for f in dtype.names:
        out[f] = raw[f]
    return out


class ContactRecord(Structure):
    _fields_ = [("contactId", c_int32), ("shapeIdA", c_int32), ("shapeIdB", c_int32),
                ("colorIndex", c_int32), ("localIndex", c_int32), ("simFlags", c_uint32),
                ("friction", c_float), ("restitution", c_float), ("manifold", Manifold)]


class BodyRecord(Structure):
    _fields_ = [("bodyId", c_int32), ("type", c_int32), ("setIndex", c_int32),
                ("localIndex", c_int32), ("transform", Transform), ("center", Vec2),
                ("linearVelocity", Vec2), ("angularVelocity", c_float), ("pad", c_float)]


assert ctypes.sizeof(WorldDef) == 80
assert ctypes.sizeof(BodyDef) == 72
assert ctypes.sizeof(ShapeDef) == 48
assert ctypes.sizeof(Polygon) == 144
assert ctypes.sizeof(Manifold) == 108
assert ctypes.sizeof(BodyRecord) == 56

# numpy views of the records (same byte layout)
MANIFOLD_POINT_DTYPE = np.dtype([("point", "<f4", 2), ("anchorA", "<f4", 2), ("anchorB", "<f4", 2),
                                 ("separation", "<f4"), ("normalImpulse", "<f4"),
                                 ("tangentImpulse", "<f4"), ("maxNormalImpulse", "<f4"),
                                 ("normalVelocity", "<f4"), ("id", "<u2"), ("persisted", "u1"),
                                 ("_pad", "u1")])
CONTACT_DTYPE = np.dtype([("contactId", "<i4"), ("shapeIdA", "<i4"), ("shapeIdB", "<i4"),
                          ("colorIndex", "<i4"), ("localIndex", "<i4"), ("simFlags", "<u4"),
                          ("friction", "<f4"), ("restitution", "<f4"),
                          ("points", MANIFOLD_POINT_DTYPE, 2), ("normal", "<f4", 2),
                          ("pointCount", "<i4")])
BODY_DTYPE = np.dtype([("bodyId", "<i4"), ("type", "<i4"), ("setIndex", "<i4"),
                       ("localIndex", "<i4"), ("p", "<f4", 2), ("q", "<f4", 2),
                       ("center", "<f4", 2), ("v", "<f4", 2), ("w", "<f4"), ("pad", "<f4")])
JOINT_DTYPE = np.dtype([("jointId", "<i4"), ("type", "<i4"), ("force", "<f4", 2), ("torque", "<f4")])
assert CONTACT_DTYPE.itemsize == ctypes.sizeof(ContactRecord)
assert BODY_DTYPE.itemsize == ctypes.sizeof(BodyRecord)

# the public manifold functions (include/box2d/collision.h): name, shape A type, shape B type
COLLIDE_FUNCTIONS = [
    ("b2CollideCircles", Circle, Circle),
    ("b2CollideCapsuleAndCircle", Capsule, Circle),
    ("b2CollideSegmentAndCircle", Segment, Circle),
    ("b2CollidePolygonAndCircle", Polygon, Circle),
    ("b2CollideCapsules", Capsule, Capsule),
    ("b2CollideSegmentAndCapsule", Segment, Capsule),
    ("b2CollidePolygonAndCapsule", Polygon, Capsule),
    ("b2CollidePolygons", Polygon, Polygon),
    ("b2CollideSegmentAndPolygon", Segment, Polygon),
]

WORLD_FACTORY = ctypes.CFUNCTYPE(c_uint32, POINTER(WorldDef), c_void_p)


def _open(path: str) -> ctypes.CDLL:
    return ctypes.CDLL(path, mode=os.RTLD_LOCAL)


class B2Api:
    """One implementation of the public C API plus its scene library."""

    def __init__(self, lib_path: str, scenes_path: str | None = None):
        self.lib_path = lib_path
        self.lib = _open(lib_path)
        self.scenes = _open(scenes_path) if scenes_path else None
        L = self.lib
        L.b2DefaultWorldDef.restype = WorldDef
        L.b2CreateWorld.restype = WorldId
        L.b2CreateWorld.argtypes = [POINTER(WorldDef)]
        L.b2DestroyWorld.argtypes = [WorldId]
        L.b2World_IsValid.argtypes = [WorldId]
        L.b2World_IsValid.restype = c_bool
        L.b2World_Step.argtypes = [WorldId, c_float, c_int]
        L.b2World_GetCounters.argtypes = [WorldId]
        L.b2World_GetCounters.restype = Counters
        L.b2World_GetProfile.argtypes = [WorldId]
        L.b2World_GetProfile.restype = Profile
        L.b2World_EnableWarmStarting.argtypes = [WorldId, c_bool]
        L.b2World_EnableContinuous.argtypes = [WorldId, c_bool]
        L.b2Body_GetPosition.argtypes = [BodyId]
        L.b2Body_GetPosition.restype = Vec2
        L.b2Body_GetRotation.argtypes = [BodyId]
        L.b2Body_GetRotation.restype = Rot
        L.b2Body_GetLinearVelocity.argtypes = [BodyId]
        L.b2Body_GetLinearVelocity.restype = Vec2
        L.b2Body_GetAngularVelocity.argtypes = [BodyId]
        L.b2Body_GetAngularVelocity.restype = c_float
        L.b2Body_SetLinearVelocity.argtypes = [BodyId, Vec2]
        L.b2Body_SetAngularVelocity.argtypes = [BodyId, c_float]
        L.b2Body_GetType.argtypes = [BodyId]
        L.b2Body_GetType.restype = c_int
        L.b2Body_GetMass.argtypes = [BodyId]
        L.b2Body_GetMass.restype = c_float
        L.b2Body_GetInertiaTensor.argtypes = [BodyId]
        L.b2Body_GetInertiaTensor.restype = c_float
        L.b2Body_GetContactCapacity.argtypes = [BodyId]
        L.b2Body_GetContactCapacity.restype = c_int
        L.b2DefaultBodyDef.restype = BodyDef
        L.b2DefaultShapeDef.restype = ShapeDef
        L.b2CreateBody.argtypes = [WorldId, POINTER(BodyDef)]
        L.b2CreateBody.restype = BodyId
        L.b2DestroyBody.argtypes = [BodyId]
        L.b2Body_IsValid.argtypes = [BodyId]
        L.b2Body_IsValid.restype = c_bool
        for name, shape in (("Polygon", Polygon), ("Circle", Circle), ("Capsule", Capsule), ("Segment", Segment)):
            f = getattr(L, f"b2Create{name}Shape")
            f.argtypes = [BodyId, POINTER(ShapeDef), POINTER(shape)]
            f.restype = ShapeId
        L.b2MakeBox.argtypes = [c_float, c_float]
        L.b2MakeBox.restype = Polygon
        L.b2MakeSquare.argtypes = [c_float]
        L.b2MakeSquare.restype = Polygon
        L.b2MakeRoundedBox.argtypes = [c_float, c_float, c_float]
        L.b2MakeRoundedBox.restype = Polygon
        L.b2MakeOffsetBox.argtypes = [c_float, c_float, Vec2, c_float]
        L.b2MakeOffsetBox.restype = Polygon
        for name, shape in (("Polygon", Polygon), ("Circle", Circle), ("Capsule", Capsule)):
            f = getattr(L, f"b2Compute{name}Mass")
            f.argtypes = [POINTER(shape), c_float]
            f.restype = MassData
        for name, shape in (("Polygon", Polygon), ("Circle", Circle), ("Capsule", Capsule), ("Segment", Segment)):
            f = getattr(L, f"b2Compute{name}AABB")
            f.argtypes = [POINTER(shape), Transform]
            f.restype = AABB
        L.b2SegmentDistance.argtypes = [Vec2, Vec2, Vec2, Vec2]
        L.b2SegmentDistance.restype = SegmentDistanceResult
        L.b2MakeProxy.argtypes = [POINTER(Vec2), c_int, c_float]
        L.b2MakeProxy.restype = DistanceProxy
        L.b2GetSweepTransform.argtypes = [POINTER(Sweep), c_float]
        L.b2GetSweepTransform.restype = Transform
        L.b2ShapeDistance.argtypes = [POINTER(DistanceCache), POINTER(DistanceInput), ctypes.c_void_p, c_int]
        L.b2ShapeDistance.restype = DistanceOutput
        L.b2TimeOfImpact.argtypes = [POINTER(TOIInput)]
        L.b2TimeOfImpact.restype = TOIOutput
        for name, sa, sb in COLLIDE_FUNCTIONS:
            f = getattr(L, name)
            f.argtypes = [POINTER(sa), Transform, POINTER(sb), Transform]
            f.restype = Manifold
        L.b2World_GetBodyEvents.argtypes = [WorldId]
        L.b2World_GetBodyEvents.restype = BodyEvents
        L.b2World_GetSensorEvents.argtypes = [WorldId]
        L.b2World_GetSensorEvents.restype = SensorEvents
        L.b2World_GetContactEvents.argtypes = [WorldId]
        L.b2World_GetContactEvents.restype = ContactEvents
        L.b2Shape_EnableContactEvents.argtypes = [ShapeId, c_bool]
        L.b2Shape_AreContactEventsEnabled.argtypes = [ShapeId]
        L.b2Shape_AreContactEventsEnabled.restype = c_bool
        L.b2Shape_EnableHitEvents.argtypes = [ShapeId, c_bool]
        L.b2Shape_AreHitEventsEnabled.argtypes = [ShapeId]
        L.b2Shape_AreHitEventsEnabled.restype = c_bool
        L.b2Body_EnableHitEvents.argtypes = [BodyId, c_bool]
        L.b2Body_GetShapeCount.argtypes = [BodyId]
        L.b2Body_GetShapeCount.restype = c_int
        L.b2Body_GetShapes.argtypes = [BodyId, POINTER(ShapeId), c_int]
        L.b2Body_GetShapes.restype = c_int
        L.b2World_SetHitEventThreshold.argtypes = [WorldId, c_float]
        # joints
        L.b2Body_GetJointCount.argtypes = [BodyId]
        L.b2Body_GetJointCount.restype = c_int
        L.b2Body_GetJoints.argtypes = [BodyId, POINTER(JointId), c_int]
        L.b2Body_GetJoints.restype = c_int
        L.b2DestroyJoint.argtypes = [JointId]
        L.b2Joint_IsValid.argtypes = [JointId]
        L.b2Joint_IsValid.restype = c_bool
        L.b2Joint_GetType.argtypes = [JointId]
        L.b2Joint_GetType.restype = c_int
        L.b2Joint_GetConstraintForce.argtypes = [JointId]
        L.b2Joint_GetConstraintForce.restype = Vec2
        L.b2Joint_GetConstraintTorque.argtypes = [JointId]
        L.b2Joint_GetConstraintTorque.restype = c_float
        L.b2Joint_GetBodyA.argtypes = [JointId]
        L.b2Joint_GetBodyA.restype = BodyId
        L.b2Joint_GetBodyB.argtypes = [JointId]
        L.b2Joint_GetBodyB.restype = BodyId
        L.b2Joint_GetLocalAnchorA.argtypes = [JointId]
        L.b2Joint_GetLocalAnchorA.restype = Vec2
        L.b2Joint_GetLocalAnchorB.argtypes = [JointId]
        L.b2Joint_GetLocalAnchorB.restype = Vec2
        L.b2Body_GetTransform.argtypes = [BodyId]
        L.b2Body_GetTransform.restype = Transform
        L.b2RevoluteJoint_SetMotorSpeed.argtypes = [JointId, c_float]
        L.b2RevoluteJoint_GetMotorTorque.argtypes = [JointId]
        L.b2RevoluteJoint_GetMotorTorque.restype = c_float
        L.b2RevoluteJoint_GetAngle.argtypes = [JointId]
        L.b2RevoluteJoint_GetAngle.restype = c_float
        L.b2RevoluteJoint_EnableLimit.argtypes = [JointId, c_bool]
        L.b2RevoluteJoint_SetLimits.argtypes = [JointId, c_float, c_float]
        L.b2mi_World_DumpContacts.argtypes = [WorldId, c_void_p, c_int]
        L.b2mi_World_DumpContacts.restype = c_int
        L.b2mi_World_DumpBodies.argtypes = [WorldId, c_void_p, c_int]
        L.b2mi_World_DumpBodies.restype = c_int
        if self.scenes is not None:
            S = self.scenes
            S.scene_set_world_factory.argtypes = [c_void_p, c_void_p]
            S.scene_large_pyramid.argtypes = [POINTER(WorldDef), c_int, POINTER(WorldId), c_void_p, c_int]
            S.scene_pyramid_grid.argtypes = [POINTER(WorldDef), c_int, c_int, c_int, c_uint64,
                                             POINTER(WorldId), c_void_p, c_int]
            S.scene_tumbler.argtypes = [POINTER(WorldDef), c_int, c_uint64, POINTER(WorldId), c_void_p, c_int]
            S.scene_batch_pyramid.argtypes = [POINTER(WorldDef), c_uint64, c_int, POINTER(WorldId),
                                              c_void_p, c_int, POINTER(c_float)]
            S.scene_drop_boxes.argtypes = [POINTER(WorldDef), c_uint64, c_int, c_int, POINTER(WorldId),
                                           c_void_p, c_int, POINTER(c_float)]
            S.scene_shape_zoo.argtypes = [POINTER(WorldDef), c_int, c_uint64, POINTER(WorldId), c_void_p, c_int]
            S.scene_hello.argtypes = [POINTER(WorldDef), POINTER(WorldId), c_void_p, c_int]
            S.scene_tilted_stacks.argtypes = [POINTER(WorldDef), POINTER(WorldId), c_void_p, c_int]
            S.scene_overflow.argtypes = [POINTER(WorldDef), c_int, c_float, POINTER(WorldId), c_void_p, c_int]
            S.scene_polygon_zoo.argtypes = [POINTER(WorldDef), c_int, c_uint64, POINTER(WorldId), c_void_p, c_int]
            S.scene_single.argtypes = [POINTER(WorldDef), c_int, c_float, c_float, c_float, c_float, c_float,
                                       c_float, c_float, c_float, c_float, c_float, c_int,
                                       POINTER(WorldId), c_void_p, c_int]
            S.scene_step_n.argtypes = [WorldId, c_float, c_int, c_int]
            S.scene_joint_tumbler.argtypes = [POINTER(WorldDef), c_int, POINTER(WorldId), c_void_p, c_int]
            S.scene_joint_grid.argtypes = [POINTER(WorldDef), c_int, c_float, POINTER(WorldId), c_void_p, c_int]
            S.scene_joint_zoo.argtypes = [POINTER(WorldDef), c_int, c_uint32, c_uint64, POINTER(WorldId), c_void_p, c_int]
            S.scene_continuous.argtypes = [POINTER(WorldDef), c_int, c_uint64, c_int, POINTER(WorldId), c_void_p, c_int]

    # ---- worlds ----
    def default_world_def(self, **overrides) -> WorldDef:
        d = self.lib.b2DefaultWorldDef()
        for k, v in overrides.items():
            setattr(d, k, v)
        return d

    def step(self, world: WorldId, dt: float = 1.0 / 60.0, substeps: int = 4, n: int = 1):
        if self.scenes is not None:
            self.scenes.scene_step_n(world, c_float(dt), substeps, n)
        else:
            for _ in range(n):
                self.lib.b2World_Step(world, c_float(dt), substeps)

    def counters(self, world: WorldId) -> Counters:
        return self.lib.b2World_GetCounters(world)

    def destroy(self, world: WorldId):
        self.lib.b2DestroyWorld(world)

    # ---- scenes (return world id and the dynamic body ids) ----
    def _bodies(self, n_max: int):
        return (BodyId * max(n_max, 1))()

    def scene(self, name: str, world_def: WorldDef | None = None, **kw):
        d = world_def if world_def is not None else self.default_world_def(enableSleep=False)
        w = WorldId()
        S = self.scenes
        if name == "large_pyramid":
            base = kw.get("base", 100)
            ids = self._bodies(base * (base + 1) // 2)
            n = S.scene_large_pyramid(ctypes.byref(d), base, ctypes.byref(w), ids, len(ids))
        elif name == "pyramid_grid":
            rows, cols, base = kw.get("rows", 2), kw.get("cols", 2), kw.get("base", 10)
            ids = self._bodies(rows * cols * base * (base + 1) // 2)
            n = S.scene_pyramid_grid(ctypes.byref(d), rows, cols, base, kw.get("seed", 1),
                                     ctypes.byref(w), ids, len(ids))
        elif name == "tumbler":
            count = kw.get("count", 500)
            ids = self._bodies(count)
            n = S.scene_tumbler(ctypes.byref(d), count, kw.get("seed", 1), ctypes.byref(w), ids, len(ids))
        elif name == "batch_pyramid":
            ids = self._bodies(820)
            f = c_float()
            n = S.scene_batch_pyramid(ctypes.byref(d), kw.get("seed", 1), kw.get("index", 0),
                                      ctypes.byref(w), ids, len(ids), ctypes.byref(f))
        elif name == "drop_boxes":
            count = kw.get("count", 16)
            ids = self._bodies(count)
            params = (c_float * 2)()
            n = S.scene_drop_boxes(ctypes.byref(d), kw.get("seed", 1), kw.get("index", 0), count,
                                   ctypes.byref(w), ids, len(ids), params)
        elif name == "shape_zoo":
            copies = kw.get("copies", 4)
            ids = self._bodies(copies * 5)
            n = S.scene_shape_zoo(ctypes.byref(d), copies, kw.get("seed", 1), ctypes.byref(w), ids, len(ids))
        elif name == "hello":
            ids = self._bodies(1)
            n = S.scene_hello(ctypes.byref(d), ctypes.byref(w), ids, len(ids))
        elif name == "polygon_zoo":
            copies = kw.get("copies", 4)
            ids = self._bodies(copies * 8)
            n = S.scene_polygon_zoo(ctypes.byref(d), copies, kw.get("seed", 1), ctypes.byref(w), ids, len(ids))
        elif name in ("overflow", "overflow_negzero"):
            # overflow_negzero: the kinematic rake's velocity x is -0.0
            count = kw.get("count", 30)
            ids = self._bodies(1 + count)
            vx = -0.0 if name == "overflow_negzero" else 0.0
            n = S.scene_overflow(ctypes.byref(d), count, vx, ctypes.byref(w), ids, len(ids))
        elif name == "tilted_stacks":
            ids = self._bodies(100)
            n = S.scene_tilted_stacks(ctypes.byref(d), ctypes.byref(w), ids, len(ids))
        elif name == "single":
            ids = self._bodies(1)
            shape = {"circle": 0, "box": 3}[kw.get("shape", "box")]
            n = S.scene_single(ctypes.byref(d), shape, kw.get("size", 0.5), kw.get("x", 0.0), kw.get("y", 0.5),
                               kw.get("vx", 0.0), kw.get("vy", 0.0), kw.get("w", 0.0), kw.get("friction", 0.6),
                               kw.get("restitution", 0.0), kw.get("linearDamping", 0.0),
                               kw.get("angularDamping", 0.0), int(kw.get("ground", True)),
                               ctypes.byref(w), ids, len(ids))
        elif name == "joint_tumbler":
            grid = kw.get("grid", 45)
            ids = self._bodies(1 + grid * grid)
            n = S.scene_joint_tumbler(ctypes.byref(d), grid, ctypes.byref(w), ids, len(ids))
        elif name == "joint_grid":
            N = kw.get("N", 100)
            ids = self._bodies(N * N)
            n = S.scene_joint_grid(ctypes.byref(d), N, kw.get("gravity", 0.0), ctypes.byref(w), ids, len(ids))
        elif name == "joint_zoo":
            spokes = kw.get("spokes", 16)
            ids = self._bodies(64 + spokes)
            n = S.scene_joint_zoo(ctypes.byref(d), spokes, kw.get("types", 0x7F), kw.get("seed", 1),
                                  ctypes.byref(w), ids, len(ids))
        elif name == "continuous":
            count = kw.get("count", 40)
            ids = self._bodies(count + 1)
            n = S.scene_continuous(ctypes.byref(d), count, kw.get("seed", 1), kw.get("bullets", 0), ctypes.byref(w), ids, len(ids))
        else:
            raise ValueError(f"unknown scene {name}")
        return w, list(ids)[:n]

    # ---- joints ----
    def joints_of(self, bodies) -> list:
        """The distinct joints attached to `bodies`, in ascending joint id."""
        seen = {}
        for b in bodies:
            n = self.lib.b2Body_GetJointCount(b)
            if n == 0:
                continue
            arr = (JointId * n)()
            m = self.lib.b2Body_GetJoints(b, arr, n)
            for i in range(m):
                seen[arr[i].index1] = JointId(arr[i].index1, arr[i].world0, arr[i].revision)
        return [seen[k] for k in sorted(seen)]

    def joint_dump(self, joints) -> np.ndarray:
        """Per joint: type, constraint force (x, y) and torque, as float32 bits comparable across
        implementations (the force and torque are the accumulated impulses times 1/h)."""
        out = np.zeros(len(joints), dtype=JOINT_DTYPE)
        for i, j in enumerate(joints):
            f = self.lib.b2Joint_GetConstraintForce(j)
            out[i] = (j.index1, self.lib.b2Joint_GetType(j), (f.x, f.y), self.lib.b2Joint_GetConstraintTorque(j))
        return out

    def body_shapes(self, body: BodyId) -> list:
        n = self.lib.b2Body_GetShapeCount(body)
        arr = (ShapeId * max(n, 1))()
        m = self.lib.b2Body_GetShapes(body, arr, n)
        return [arr[i] for i in range(m)]

    # ---- events of the last step ----
    def events(self, world: WorldId) -> dict:
        """The last step's events as arrays: contact begin/end/hit, sensor begin/end, body moves."""
        c = self.lib.b2World_GetContactEvents(world)
        se = self.lib.b2World_GetSensorEvents(world)
        b = self.lib.b2World_GetBodyEvents(world)
        return {"begin": _event_array(c.beginEvents, c.beginCount, TOUCH_EVENT_DTYPE),
                "end": _event_array(c.endEvents, c.endCount, TOUCH_EVENT_DTYPE),
                "hit": _event_array(c.hitEvents, c.hitCount, HIT_EVENT_DTYPE),
                "sensor_begin": _event_array(se.beginEvents, se.beginCount, TOUCH_EVENT_DTYPE),
                "sensor_end": _event_array(se.endEvents, se.endCount, TOUCH_EVENT_DTYPE),
                "move": _event_array(b.moveEvents, b.moveCount, MOVE_EVENT_DTYPE)}

    # ---- inspection ----
    def dump_contacts(self, world: WorldId) -> np.ndarray:
        n = self.lib.b2mi_World_DumpContacts(world, None, 0)
        out = np.zeros(n, dtype=CONTACT_DTYPE)
        if n:
            self.lib.b2mi_World_DumpContacts(world, out.ctypes.data, n)
        return out

    def dump_bodies(self, world: WorldId) -> np.ndarray:
        n = self.lib.b2mi_World_DumpBodies(world, None, 0)
        out = np.zeros(n, dtype=BODY_DTYPE)
        if n:
            self.lib.b2mi_World_DumpBodies(world, out.ctypes.data, n)
        return out

    # ---- extensions of libb2mi.so (batching, launch shape, timing) ----
    def bind_product(self):
        L = self.lib
        L.b2mi_CreateBatch.restype = c_int
        L.b2mi_DestroyBatch.argtypes = [c_int]
        L.b2mi_Batch_CreateWorld.argtypes = [c_int, POINTER(WorldDef)]
        L.b2mi_Batch_CreateWorld.restype = WorldId
        L.b2mi_Batch_GetWorldCount.argtypes = [c_int]
        L.b2mi_Batch_GetWorldCount.restype = c_int
        L.b2mi_Batch_Step.argtypes = [c_int, c_float, c_int]
        L.b2mi_Batch_Configure.argtypes = [c_int, c_int, c_int]
        L.b2mi_World_Configure.argtypes = [WorldId, c_int, c_int]
        L.b2mi_Batch_Autotune.argtypes = [c_int, c_float, c_int, c_int]
        L.b2mi_Batch_Autotune.restype = c_float
        L.b2mi_World_Autotune.argtypes = [WorldId, c_float, c_int, c_int]
        L.b2mi_World_Autotune.restype = c_float
        L.b2mi_Batch_GetLaunchShape.argtypes = [c_int, POINTER(c_int), POINTER(c_int), POINTER(c_int), POINTER(c_int)]
        L.b2mi_Batch_SetLdsBudget.argtypes = [c_int, c_int]
        L.b2mi_World_SetLdsBudget.argtypes = [WorldId, c_int]
        L.b2mi_Batch_SetLdsFullBudget.argtypes = [c_int, c_int]
        L.b2mi_Batch_GetSolverLaunches.argtypes = [c_int, ctypes.c_void_p, c_int]
        L.b2mi_Batch_GetSolverLaunches.restype = c_int
        L.b2mi_Batch_GetHostTimeline.argtypes = [c_int, ctypes.c_void_p]
        L.b2mi_Batch_SetWorldChunks.argtypes = [c_int, c_int]
        L.b2mi_Batch_SetSplitSlots.argtypes = [c_int, c_int]
        L.b2mi_World_SetSplitSlots.argtypes = [WorldId, c_int]
        L.b2mi_Batch_SetGridBlocks.argtypes = [c_int, c_int]
        L.b2mi_World_SetGridBlocks.argtypes = [WorldId, c_int]
        L.b2mi_Batch_SetLevelRunSlots.argtypes = [c_int, c_int]
        L.b2mi_World_GetTimeOfImpactCount.argtypes = [WorldId]
        L.b2mi_World_SetLevelRunSlots.argtypes = [WorldId, c_int]
        L.b2mi_World_SetDevicePairs.argtypes = [WorldId, c_int]
        L.b2mi_World_GetDevicePairStats.argtypes = [WorldId, POINTER(c_float)]
        L.b2mi_Batch_SetLaunchGraphs.argtypes = [c_int, c_int]
        L.b2mi_World_SetLaunchGraphs.argtypes = [WorldId, c_int]
        L.b2mi_Batch_GetLaunchGraphStats.argtypes = [c_int, POINTER(c_int), POINTER(c_int)]
        L.b2mi_World_GetLaunchGraphStats.argtypes = [WorldId, POINTER(c_int), POINTER(c_int)]
        L.b2mi_Batch_GetWorkItemCount.argtypes = [c_int]
        L.b2mi_Batch_GetWorkItems.argtypes = [c_int, ctypes.c_void_p, c_int]
        L.b2mi_Batch_GetWorkItems.restype = c_int
        L.b2mi_World_GetWorkItems.argtypes = [WorldId, ctypes.c_void_p, c_int]
        L.b2mi_World_GetWorkItems.restype = c_int
        L.b2mi_Batch_GetWorkItemCount.restype = c_int
        L.b2mi_Batch_GetTiming.argtypes = [c_int, POINTER(c_float), POINTER(c_float), POINTER(c_float)]
        L.b2mi_World_GetTiming.argtypes = [WorldId, POINTER(c_float), POINTER(c_float), POINTER(c_float)]
        L.b2mi_DeviceCount.restype = c_int
        L.b2mi_GetLastError.restype = c_char_p
        self._factories = []

    def create_batch(self) -> int:
        b = self.lib.b2mi_CreateBatch()
        if b <= 0:
            raise RuntimeError("no free batch")
        return b

    def use_batch(self, batch: int | None):
        """Routes the scene builders' world creation into a batch (None = b2CreateWorld)."""
        if batch is None:
            self.scenes.scene_set_world_factory(None, None)
            return
        lib = self.lib

        def factory(defp, _ctx):
            w = lib.b2mi_Batch_CreateWorld(batch, defp)
            return w.index1 | (w.revision << 16)

        cb = WORLD_FACTORY(factory)
        self._factories.append(cb)
        self.scenes.scene_set_world_factory(ctypes.cast(cb, c_void_p), None)

    def batch_step(self, batch: int, dt: float = 1.0 / 60.0, substeps: int = 4, n: int = 1):
        for _ in range(n):
            self.lib.b2mi_Batch_Step(batch, c_float(dt), substeps)

    def batch_timing(self, batch: int):
        a, b, c = c_float(), c_float(), c_float()
        self.lib.b2mi_Batch_GetTiming(batch, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def work_items(self, world: WorldId | None = None, batch: int | None = None) -> np.ndarray:
        """The last step's work items, one row each: slots, bodies, store, mode (0 one
        workgroup, 1 one grid launch per phase)."""
        f, h = (self.lib.b2mi_World_GetWorkItems, world) if world is not None else (self.lib.b2mi_Batch_GetWorkItems, batch)
        n = f(h, None, 0)
        out = np.zeros((n, 4), np.int32)
        if n:
            f(h, out.ctypes.data, n)
        return out

    def graph_stats(self, world: WorldId | None = None, batch: int | None = None):
        """(captures, replays) of the solver's launch graphs so far."""
        a, b = c_int(), c_int()
        if world is not None:
            self.lib.b2mi_World_GetLaunchGraphStats(world, ctypes.byref(a), ctypes.byref(b))
        else:
            self.lib.b2mi_Batch_GetLaunchGraphStats(batch, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def device_pair_stats(self, world: WorldId) -> dict:
        """The device pair queries' accumulated phase clocks (ms) and counts."""
        v = (c_float * 9)()
        self.lib.b2mi_World_GetDevicePairStats(world, v)
        keys = ("table_ms", "uploads_ms", "queries_ms", "readback_ms", "contacts_ms", "pairs", "moves", "updates", "spilled")
        return dict(zip(keys, list(v)))

    def world_timing(self, world: WorldId):
        a, b, c = c_float(), c_float(), c_float()
        self.lib.b2mi_World_GetTiming(world, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value
