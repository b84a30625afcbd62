"""Gradients of batched rollouts: a torch.autograd.Function over the C ABI's taped runs.

A rollout builds a batch of drop-box worlds (scenes.c: scene_drop_boxes_ex), records every
step on the GPU (b2mi_Batch_BeginTape), and returns the final poses. Its backward pass calls
b2mi_Batch_Backward, the hand-written adjoint of the recorded solver steps (adjoint.hip), and
returns gradients with respect to each world's friction and restitution and each box's initial
linear and angular velocity. The narrowphase manifold of every step is held fixed (it is a
constant of the step, not a function of the parameters).

Tensors cross the boundary on the host (float32); the simulation itself runs on the GPU.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from .capi import B2Api, BodyId, WorldId

POSE = 7  # center x, y, rotation cos, sin, velocity x, y, angular velocity


def bind_scene(api: B2Api):
    """The drop-box scene builder and the pose getters (any library with the public API)."""
    from .capi import Vec2
    L = api.lib
    L.b2Body_GetWorldCenterOfMass.argtypes = [BodyId]
    L.b2Body_GetWorldCenterOfMass.restype = Vec2
    S = api.scenes
    S.scene_drop_boxes_ex.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(WorldId), ctypes.c_void_p, ctypes.c_int,
                                      ctypes.c_void_p]
    S.scene_drop_boxes_ex.restype = ctypes.c_int


def bind(api: B2Api):
    bind_scene(api)
    L = api.lib
    L.b2mi_Batch_BeginTape.argtypes = [ctypes.c_int, ctypes.c_int]
    L.b2mi_Batch_EndTape.argtypes = [ctypes.c_int]
    L.b2mi_Batch_GetTapeStepCount.argtypes = [ctypes.c_int]
    L.b2mi_Batch_GetTapeStepCount.restype = ctypes.c_int
    L.b2mi_Batch_GetTapeBodies.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    L.b2mi_Batch_GetTapeBodies.restype = ctypes.c_int
    L.b2mi_Batch_GetTapePoses.argtypes = [ctypes.c_int, ctypes.c_void_p]
    L.b2mi_Batch_ReplayTape.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.b2mi_Batch_ReplayTapeNoAlloc.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.b2mi_Batch_ReplayTapeNoAlloc.restype = ctypes.c_int
    L.b2mi_Batch_Backward.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]


def _ptr(a: np.ndarray):
    return a.ctypes.data if a is not None else None


class DropBoxesRollout:
    """`worlds` drop-box worlds of `boxes` boxes each, `steps` steps of `substeps` substeps, a
    checkpoint every `checkpoint` steps. run() and backward() may be called repeatedly; each
    run() builds the worlds afresh from the seed and the given parameters."""

    def __init__(self, api: B2Api, worlds: int, boxes: int = 16, steps: int = 120, substeps: int = 4,
                 checkpoint: int = 10, seed: int = 11, dt: float = 1.0 / 60.0, pre_steps: int = 0):
        bind(api)
        self.api, self.worlds, self.boxes, self.steps = api, worlds, boxes, steps
        self.pre_steps = pre_steps  # untaped steps before the tape begins
        self.substeps, self.checkpoint, self.seed, self.dt = substeps, checkpoint, seed, dt
        self.batch = None
        self.start = None
        self.lds_budget = None  # b2mi_Batch_SetLdsBudget of the rollout's batch (None = default)
        self.forward_seconds = self.backward_seconds = 0.0

    def world_def(self):
        # continuous collision is off: falling boxes reach time-of-impact speeds, and the time of
        # impact solve is not part of the device pipeline
        return self.api.default_world_def(enableSleep=False, enableContinous=False)

    def default_params(self):
        """The seeded per-world friction, restitution and per-box initial velocities."""
        api, d = self.api, self.world_def()
        fr = np.zeros((self.worlds, 2), np.float32)
        v0 = np.zeros((self.worlds, self.boxes, 2), np.float32)
        w0 = np.zeros((self.worlds, self.boxes), np.float32)
        for i in range(self.worlds):
            w = WorldId()
            ids = (BodyId * self.boxes)()
            out = (ctypes.c_float * 2)()
            api.scenes.scene_drop_boxes_ex(ctypes.byref(d), self.seed, i, self.boxes, -1.0, -1.0, None, None, ctypes.byref(w),
                                           ids, self.boxes, out)
            fr[i] = (out[0], out[1])
            for b in range(self.boxes):
                v = api.lib.b2Body_GetLinearVelocity(ids[b])
                v0[i, b] = (v.x, v.y)
                w0[i, b] = api.lib.b2Body_GetAngularVelocity(ids[b])
            api.lib.b2DestroyWorld(w)
        return fr[:, 0].copy(), fr[:, 1].copy(), v0, w0

    def close(self):
        if self.batch is not None:
            self.api.lib.b2mi_Batch_EndTape(self.batch)
            self.api.lib.b2mi_DestroyBatch(self.batch)
            self.batch = None

    def run(self, friction, restitution, v0, w0, taped: bool = True) -> np.ndarray | None:
        """Final poses, shape (worlds, boxes, 7). With taped=False the same steps run without a
        tape (one work item per world all the same) and only forward_seconds is set: the
        untaped cost of the rollout, for comparison."""
        self.close()
        api = self.api
        friction = np.ascontiguousarray(friction, np.float32)
        restitution = np.ascontiguousarray(restitution, np.float32)
        v0 = np.ascontiguousarray(v0, np.float32).reshape(self.worlds, self.boxes, 2)
        w0 = np.ascontiguousarray(w0, np.float32).reshape(self.worlds, self.boxes)
        d = self.world_def()
        self.batch = api.create_batch()
        if self.lds_budget is not None:
            api.lib.b2mi_Batch_SetLdsBudget(self.batch, int(self.lds_budget))
        api.use_batch(self.batch)
        try:
            for i in range(self.worlds):
                w = WorldId()
                ids = (BodyId * self.boxes)()
                n = api.scenes.scene_drop_boxes_ex(ctypes.byref(d), self.seed, i, self.boxes, float(friction[i]),
                                                   float(restitution[i]), v0[i].ctypes.data, w0[i].ctypes.data, ctypes.byref(w),
                                                   ids, self.boxes, None)
                assert n == self.boxes
        finally:
            api.use_batch(None)
        for _ in range(self.pre_steps):
            api.lib.b2mi_Batch_Step(self.batch, ctypes.c_float(self.dt), self.substeps)
        if not taped:
            # the work items a tape asks for: one per world
            api.lib.b2mi_Batch_Configure(self.batch, 0, 1 << 24)
            t0 = time.perf_counter()
            for _ in range(self.steps):
                api.lib.b2mi_Batch_Step(self.batch, ctypes.c_float(self.dt), self.substeps)
            self.forward_seconds = time.perf_counter() - t0  # a step returns with its device work done
            return None
        api.lib.b2mi_Batch_BeginTape(self.batch, self.checkpoint)
        n = api.lib.b2mi_Batch_GetTapeBodies(self.batch, None, 0)
        assert n == self.worlds * self.boxes, "tape order must be world by world, box by box"
        ids = (BodyId * n)()
        api.lib.b2mi_Batch_GetTapeBodies(self.batch, ids, n)
        self.start = np.zeros((n, POSE), np.float32)
        for k in range(n):
            c = api.lib.b2Body_GetWorldCenterOfMass(ids[k])
            q = api.lib.b2Body_GetRotation(ids[k])
            v = api.lib.b2Body_GetLinearVelocity(ids[k])
            self.start[k] = (c.x, c.y, q.c, q.s, v.x, v.y, api.lib.b2Body_GetAngularVelocity(ids[k]))
        self.start = self.start.reshape(self.worlds, self.boxes, POSE)
        t0 = time.perf_counter()
        for _ in range(self.steps):
            api.lib.b2mi_Batch_Step(self.batch, ctypes.c_float(self.dt), self.substeps)
        out = np.zeros((n, POSE), np.float32)
        api.lib.b2mi_Batch_GetTapePoses(self.batch, out.ctypes.data)  # synchronizes
        self.forward_seconds = time.perf_counter() - t0
        api.lib.b2mi_Batch_GetTapePoses(self.batch, out.ctypes.data)
        return out.reshape(self.worlds, self.boxes, POSE)

    def backward(self, grad_end: np.ndarray):
        """d(loss)/d(friction per world), d(loss)/d(restitution per world),
        d(loss)/d(v0 per box), d(loss)/d(w0 per box), d(loss)/d(start poses)."""
        g = np.ascontiguousarray(grad_end, np.float32).reshape(-1, POSE)
        gs = np.zeros_like(g)
        gp = np.zeros((self.worlds, 2), np.float32)
        t0 = time.perf_counter()
        self.api.lib.b2mi_Batch_Backward(self.batch, g.ctypes.data, gs.ctypes.data, gp.ctypes.data)
        self.backward_seconds = time.perf_counter() - t0
        gs = gs.reshape(self.worlds, self.boxes, POSE)
        return gp[:, 0].copy(), gp[:, 1].copy(), gs[:, :, 4:6].copy(), gs[:, :, 6].copy(), gs

    def replay(self, start: np.ndarray, params: np.ndarray | None = None) -> np.ndarray:
        """The taped solver steps from `start` poses (worlds, boxes, 7), with per-world
        (friction, restitution) when given: the function backward() differentiates."""
        s = np.ascontiguousarray(start, np.float32).reshape(-1, POSE)
        out = np.zeros_like(s)
        p = None if params is None else np.ascontiguousarray(params, np.float32).reshape(-1, 2)
        self.api.lib.b2mi_Batch_ReplayTape(self.batch, s.ctypes.data, _ptr(p), out.ctypes.data)
        return out.reshape(self.worlds, self.boxes, POSE)


def autograd_function():
    """torch.autograd.Function mapping (friction, restitution, v0, w0) to final box centers."""
    import torch

    class Rollout(torch.autograd.Function):
        @staticmethod
        def forward(ctx, friction, restitution, v0, w0, rollout: DropBoxesRollout):
            poses = rollout.run(friction.detach().cpu().numpy(), restitution.detach().cpu().numpy(),
                                v0.detach().cpu().numpy(), w0.detach().cpu().numpy())
            ctx.rollout = rollout
            return torch.from_numpy(poses[:, :, 0:2].copy())

        @staticmethod
        def backward(ctx, grad_centers):
            r = ctx.rollout
            g = np.zeros((r.worlds, r.boxes, POSE), np.float32)
            g[:, :, 0:2] = grad_centers.detach().cpu().numpy()
            gf, gr, gv, gw, _ = r.backward(g)
            return torch.from_numpy(gf), torch.from_numpy(gr), torch.from_numpy(gv), torch.from_numpy(gw), None

    return Rollout
