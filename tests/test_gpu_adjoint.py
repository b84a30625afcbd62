"""The taped run and its hand-written adjoint (adjoint.hip) on the GPU.

1. Taping does not change the simulation: the taped batch's final poses equal the oracle's
   stepping of the same worlds, bit for bit.
2. The replay of the taped solver steps from the recorded start poses reproduces the run.
3. The adjoint agrees with central finite differences (h = 1e-3, as the benchmark's
   configuration states) of the replayed solver steps, the function it differentiates: the
   manifolds and contact sets of each step are constants. Bar: relative error of the gradient
   vector (all velocity coordinates, all friction and restitution coordinates) below 2e-2 for
   velocities and 5e-2 for the parameters, with the float32 forward's noise floor
   (~1e-7 |L| / h) stated in the assertion message.
4. torch.autograd.Function: gradients of a loss on final centers flow to friction,
   restitution and initial velocities.
"""
import ctypes

import numpy as np
import pytest

from box2c_amd import diff
from box2c_amd.capi import BodyId, WorldId
from tests import helpers

pytestmark = pytest.mark.gpu


def rollout(worlds=4, boxes=4, steps=24, checkpoint=5, seed=3, pre_steps=0):
    r = diff.DropBoxesRollout(helpers.product(), worlds, boxes, steps, 4, checkpoint, seed=seed, pre_steps=pre_steps)
    return r, r.default_params()


def test_taped_run_matches_oracle():
    r, (f, e, v0, w0) = rollout()
    try:
        poses = r.run(f, e, v0, w0)
        cpu = helpers.oracle()
        diff.bind_scene(cpu)
        d = cpu.default_world_def(enableSleep=False, enableContinous=False)
        for i in range(r.worlds):
            w = WorldId()
            ids = (BodyId * r.boxes)()
            cpu.scenes.scene_drop_boxes_ex(ctypes.byref(d), r.seed, i, r.boxes, float(f[i]), float(e[i]), v0[i].ctypes.data,
                                           w0[i].ctypes.data, ctypes.byref(w), ids, r.boxes, None)
            cpu.step(w, n=r.steps)
            for b in range(r.boxes):
                c = cpu.lib.b2Body_GetWorldCenterOfMass(ids[b])
                v = cpu.lib.b2Body_GetLinearVelocity(ids[b])
                got = poses[i, b]
                assert np.float32(c.x) == got[0] and np.float32(c.y) == got[1], (i, b)
                assert np.float32(v.x) == got[4] and np.float32(v.y) == got[5], (i, b)
            cpu.destroy(w)
    finally:
        r.close()


def test_replay_reproduces_run():
    r, (f, e, v0, w0) = rollout()
    try:
        poses = r.run(f, e, v0, w0)
        again = r.replay(r.start)
        assert np.array_equal(poses.view(np.uint32), again.view(np.uint32))
        params = np.stack([f, e], axis=1)
        again = r.replay(r.start, params)
        assert np.array_equal(poses.view(np.uint32), again.view(np.uint32))
    finally:
        r.close()


def _fd_check(worlds, boxes, steps, seed, h=1e-3, pre_steps=0):
    r, (f, e, v0, w0) = rollout(worlds, boxes, steps, 5, seed, pre_steps)
    try:
        r.run(f, e, v0, w0)
        rng = np.random.default_rng(seed)
        weights = rng.normal(size=(worlds, boxes, diff.POSE)).astype(np.float32)
        weights[:, :, 2:4] = 0.0
        gf, ge, gv, gw, gs = r.backward(weights)
        params = np.stack([f, e], axis=1).astype(np.float64)

        def loss(start, p):
            return float(np.sum(weights.astype(np.float64) * r.replay(start, p.astype(np.float32))))

        base = loss(r.start, params)
        fd_v = np.zeros_like(gv, np.float64)
        fd_w = np.zeros_like(gw, np.float64)
        for i in range(worlds):
            for b in range(boxes):
                for k, (arr, col) in enumerate(((fd_v, 4), (fd_v, 5), (fd_w, 6))):
                    sp, sm = r.start.copy(), r.start.copy()
                    sp[i, b, col] += h
                    sm[i, b, col] -= h
                    val = (loss(sp, params) - loss(sm, params)) / (2 * h)
                    if col == 6:
                        fd_w[i, b] = val
                    else:
                        fd_v[i, b, col - 4] = val
        fd_p = np.zeros((worlds, 2))
        for i in range(worlds):
            for k in range(2):
                pp, pm = params.copy(), params.copy()
                pp[i, k] += h
                pm[i, k] -= h
                fd_p[i, k] = (loss(r.start, pp) - loss(r.start, pm)) / (2 * h)
        return base, (gv, gw, gf, ge), (fd_v, fd_w, fd_p[:, 0], fd_p[:, 1])
    finally:
        r.close()


def _rel(a, b):
    a, b = np.ravel(a).astype(np.float64), np.ravel(b).astype(np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12))


@pytest.mark.parametrize("pre,steps", [(0, 12), (0, 40), (0, 90), (60, 1), (60, 3), (60, 10)])
def test_adjoint_matches_finite_differences(pre, steps):
    base, adj, fd = _fd_check(3, 4, steps, seed=5, pre_steps=pre)
    noise = 1e-7 * abs(base) / 1e-3
    ev = _rel(np.concatenate([np.ravel(adj[0]), np.ravel(adj[1])]), np.concatenate([np.ravel(fd[0]), np.ravel(fd[1])]))
    ep = _rel(np.concatenate([adj[2], adj[3]]), np.concatenate([fd[2], fd[3]]))
    msg = (f"velocity rel err {ev:.2e}, parameter rel err {ep:.2e}, loss {base:.3f}, fd noise ~{noise:.1e}\n"
           f"adj friction {adj[2]} fd {fd[2]}\nadj restitution {adj[3]} fd {fd[3]}")
    print(msg)
    assert ev < 2e-2, msg
    assert ep < 5e-2, msg


def test_autograd_function():
    import torch
    r, (f, e, v0, w0) = rollout(2, 4, 20, 5, 9)
    try:
        F = diff.autograd_function()
        tf = torch.tensor(f, requires_grad=True)
        te = torch.tensor(e, requires_grad=True)
        tv = torch.tensor(v0, requires_grad=True)
        tw = torch.tensor(w0, requires_grad=True)
        centers = F.apply(tf, te, tv, tw, r)
        target = torch.zeros_like(centers)
        loss = ((centers - target) ** 2).sum()
        loss.backward()
        for t in (tf, te, tv, tw):
            assert t.grad is not None and torch.isfinite(t.grad).all()
        assert float(tv.grad.abs().sum()) > 0.0
    finally:
        r.close()


@pytest.mark.parametrize("pre,steps", [(0, 40), (60, 10), (0, 90)])
def test_adjoint_matches_fp64_autograd(pre, steps):
    """The float32 GPU adjoint against torch autograd through an FP64 restatement of the same
    taped steps (tests/fp64_replay.py). Bars: final poses within 1e-3 of the float32 run; relative
    gradient error below 1e-2 (velocities and start centers) and 2e-2 (friction, restitution)."""
    import torch
    from tests import fp64_replay

    r, (f, e, v0, w0) = rollout(3, 4, steps, 5, 5, pre)
    try:
        final = r.run(f, e, v0, w0)
        rng = np.random.default_rng(1)
        weights = rng.normal(size=final.shape).astype(np.float32)
        weights[:, :, 2:4] = 0.0
        gf, ge, gv, gw, gs = r.backward(weights)
        steps_, wp, mass, inertia, world_of = fp64_replay.export(r)
        start = torch.tensor(r.start.reshape(-1, 7), dtype=torch.float64, requires_grad=True)
        params = torch.tensor(np.stack([f, e], 1), dtype=torch.float64, requires_grad=True)
        out = fp64_replay.replay(steps_, wp, mass, inertia, world_of, start, params)
        err_pose = float(np.abs(out.detach().numpy() - final.reshape(-1, 7)).max())
        loss = (torch.tensor(weights.reshape(-1, 7), dtype=torch.float64) * out).sum()
        g_start, g_params = torch.autograd.grad(loss, [start, params])
        g_start = g_start.numpy().reshape(final.shape)
        g_params = g_params.numpy()
        ev = _rel(np.concatenate([np.ravel(gs[:, :, 4:7]), np.ravel(gs[:, :, 0:2])]),
                  np.concatenate([np.ravel(g_start[:, :, 4:7]), np.ravel(g_start[:, :, 0:2])]))
        ep = _rel(np.concatenate([gf, ge]), np.concatenate([g_params[:, 0], g_params[:, 1]]))
        msg = (f"pose err {err_pose:.2e}, velocity/center grad rel err {ev:.2e}, parameter grad rel err {ep:.2e}\n"
               f"gpu friction {gf} fp64 {g_params[:, 0]}\ngpu restitution {ge} fp64 {g_params[:, 1]}")
        print(msg)
        assert err_pose < 1e-3, msg
        assert ev < 1e-2, msg
        assert ep < 2e-2, msg
    finally:
        r.close()
