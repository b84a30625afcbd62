"""The taped run and its hand-written adjoint (adjoint.hip) on the GPU.

1. Taping does not change the simulation: the taped batch's final poses equal the oracle's
   stepping of the same worlds, bit for bit.
2. The replay of the taped solver steps from the recorded start poses reproduces the run.
3. The adjoint agrees with central finite differences (h = 1e-3, as the benchmark's
   configuration states) of the replayed solver steps, the function it differentiates (the
   manifolds and contact sets of each step are constants), on every coordinate where the
   difference quotient is stable.
4. torch.autograd.Function: gradients of a loss on final centers flow to friction,
   restitution and initial velocities.
5. The adjoint against torch autograd through an FP64 restatement of the taped steps.
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
    """Adjoint and central differences at h and h/2 of the replayed taped steps, for every
    velocity coordinate of the start poses and every world's friction and restitution."""
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
        adj, fd1, fd2 = [], [], []
        for i in range(worlds):
            for b in range(boxes):
                for col, g in ((4, gv[i, b, 0]), (5, gv[i, b, 1]), (6, gw[i, b])):
                    est = []
                    for step in (h, h / 2):
                        sp, sm = r.start.copy(), r.start.copy()
                        sp[i, b, col] += step
                        sm[i, b, col] -= step
                        est.append((loss(sp, params) - loss(sm, params)) / (2 * step))
                    adj.append(g)
                    fd1.append(est[0])
                    fd2.append(est[1])
        for i in range(worlds):
            for k, g in ((0, gf[i]), (1, ge[i])):
                est = []
                for step in (h, h / 2):
                    pp, pm = params.copy(), params.copy()
                    pp[i, k] += step
                    pm[i, k] -= step
                    est.append((loss(r.start, pp) - loss(r.start, pm)) / (2 * step))
                adj.append(g)
                fd1.append(est[0])
                fd2.append(est[1])
        return base, np.array(adj, np.float64), np.array(fd1), np.array(fd2)
    finally:
        r.close()


def _rel(a, b):
    a, b = np.ravel(a).astype(np.float64), np.ravel(b).astype(np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12))


@pytest.mark.parametrize("pre,steps", [(0, 12), (0, 40), (60, 1), (60, 10)])
def test_adjoint_matches_finite_differences(pre, steps):
    """The replay is piecewise smooth (friction-cone, restitution and speculative-contact
    switches), so a difference quotient is a valid reference only where it is stable: a
    coordinate counts when its central differences at h = 1e-3 and h/2 agree to within the
    float32 noise floor (4e-7 |L| / h) plus 1e-3 of their size. On those coordinates (at least
    60% of them) the adjoint must agree with the h/2 quotient to 1e-3 relative (vector norm)."""
    h = 1e-3
    base, adj, fd1, fd2 = _fd_check(3, 4, steps, seed=5, h=h, pre_steps=pre)
    noise = 4e-7 * abs(base) / h
    smooth = np.abs(fd1 - fd2) <= noise + 1e-3 * np.abs(fd2)
    err = _rel(adj[smooth], fd2[smooth]) if smooth.any() else 0.0
    scale = max(np.linalg.norm(fd2[smooth]), 1e-12)
    msg = (f"{smooth.sum()}/{len(adj)} stable coordinates, adjoint vs fd rel err {err:.2e}, loss {base:.3f}, "
           f"noise floor {noise:.1e} (relative {noise / scale * np.sqrt(smooth.sum()):.1e})")
    print(msg)
    assert smooth.mean() >= 0.6, msg
    assert err <= max(1e-3, 2 * noise * np.sqrt(smooth.sum()) / scale), msg


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


def test_adjoint_state_in_lds_or_hbm_bit_identical():
    """The adjoint kernel keeps an item's body and slot adjoints in LDS when they fit and in HBM
    under an LDS budget of 0: the same arithmetic, so the same gradients bit for bit."""
    results = []
    for budget in (None, 0):
        r, (f, e, v0, w0) = rollout(3, 6, 30, 5, 13)
        r.lds_budget = budget
        try:
            final = r.run(f, e, v0, w0)
            rng = np.random.default_rng(2)
            g = rng.normal(size=final.shape).astype(np.float32)
            results.append((final,) + tuple(r.backward(g)))
        finally:
            r.close()
    for a, b in zip(results[0], results[1]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert float(np.abs(results[0][3]).sum()) > 0.0


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
