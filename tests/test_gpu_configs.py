"""The HIP path at the benchmark configurations' full sizes.

Each workload runs at the size it is benchmarked at, through the same C ABI calls as bench.py,
against the CPU oracle on the same seeded inputs:

  large pyramid     base 100 (5,050 boxes): bit-identical every step for 60 steps and every
                    50 steps to step 500, then the stability measures of the settled stack
                    (top height, horizontal drift, deepest penetration, kinetic energy);
  pyramid grid      40 x 40 pyramids of base 10 (88,000 boxes): bit-identical every step for the
                    first 40 (settling) steps;
  tumbler           40,000 shapes in a kinematic container: bit-identical every 10 steps for
                    120 steps, by which time the pile is one island large enough to run phase by
                    phase over the grid;
  batch             4,096 pyramid worlds of base U{4..40}: 20 steps, the autotuner's 180 trial
                    steps and 100 more, bit-identical on a seeded sample of 64 worlds at five
                    points, finiteness and penetration bounds on every world;
  adjoint           1,024 drop-box worlds x 16 boxes, 120 taped steps, checkpoint every 10:
                    the taped run against the oracle on sampled worlds and the adjoint against
                    torch autograd through the FP64 restatement on sampled worlds; on the last 10
                    of those steps, the adjoint against central finite differences (h = 1e-3) on
                    sampled coordinates.
"""
import ctypes

import numpy as np
import pytest

from box2c_amd import diff
from box2c_amd.capi import BodyId, WorldId
from tests import helpers

pytestmark = pytest.mark.gpu


def _compare(gpu, cpu, wg, wc, what):
    m = helpers.first_contact_mismatch(gpu.dump_contacts(wg), cpu.dump_contacts(wc))
    assert m is None, f"{what}: contacts differ: {m}"
    m = helpers.first_body_mismatch(gpu.dump_bodies(wg), cpu.dump_bodies(wc))
    assert m is None, f"{what}: bodies differ: {m}"


def test_large_pyramid_base100_parity_and_stability():
    gpu, cpu = helpers.product(), helpers.oracle()
    wg, _ = helpers.make_world(gpu, "large_pyramid", {"base": 100}, {})
    wc, _ = helpers.make_world(cpu, "large_pyramid", {"base": 100}, {})
    try:
        start = cpu.dump_bodies(wc)
        for s in range(1, 501):
            gpu.step(wg)
            cpu.step(wc)
            if s <= 60 or s % 50 == 0:
                _compare(gpu, cpu, wg, wc, f"large pyramid step {s}")
        assert gpu.counters(wg).contactCount == cpu.counters(wc).contactCount == 14950
        g = helpers.stability(gpu.dump_bodies(wg), gpu.dump_contacts(wg), start)
        c = helpers.stability(cpu.dump_bodies(wc), cpu.dump_contacts(wc), start)
        print("large pyramid after 500 steps:", g)
        assert g == c
        # The stack still stands: it has compressed (soft contacts, ~7 mm per row) but no box has
        # slid or toppled, the deepest penetration stays near the speculative band, and it has
        # come to rest.
        assert 98.5 < g["top"] < 99.5, g
        assert g["drift"] < 0.25, g
        assert g["penetration"] < 0.03, g
        assert g["kinetic"] < 1e-2, g
    finally:
        gpu.destroy(wg)
        cpu.destroy(wc)


def test_pyramid_grid_40x40_parity():
    gpu, cpu = helpers.product(), helpers.oracle()
    kw = {"rows": 40, "cols": 40, "base": 10, "seed": 7}
    wg, ids = helpers.make_world(gpu, "pyramid_grid", kw, {})
    wc, _ = helpers.make_world(cpu, "pyramid_grid", kw, {})
    assert len(ids) == 88000
    try:
        for s in range(1, 41):
            gpu.step(wg)
            cpu.step(wc)
            _compare(gpu, cpu, wg, wc, f"pyramid grid step {s}")
        cg, cc = gpu.counters(wg), cpu.counters(wc)
        assert list(cg.colorCounts) == list(cc.colorCounts)
        assert sum(cg.colorCounts) > 200000
    finally:
        gpu.destroy(wg)
        cpu.destroy(wc)


def test_tumbler_40000_parity():
    gpu, cpu = helpers.product(), helpers.oracle()
    kw = {"count": 40000, "seed": 5}
    wg, ids = helpers.make_world(gpu, "tumbler", kw, {})
    wc, _ = helpers.make_world(cpu, "tumbler", kw, {})
    assert len(ids) == 40000
    try:
        largest = 0
        for s in range(1, 121):
            gpu.step(wg)
            cpu.step(wc)
            items = gpu.work_items(world=wg)
            if len(items):
                largest = max(largest, int(items[:, 0].max()))
            if s % 10 == 0:
                _compare(gpu, cpu, wg, wc, f"tumbler step {s}")
        cg, cc = gpu.counters(wg), cpu.counters(wc)
        assert list(cg.colorCounts) == list(cc.colorCounts)
        assert cg.colorCounts[11] > 0, "overflow constraints expected at this size"
        items = gpu.work_items(world=wg)
        print(f"tumbler: {len(items)} work items, largest {int(items[:, 0].max())} slots, "
              f"modes {sorted(set(items[:, 3].tolist()))}, colours {list(cg.colorCounts)}")
        assert largest >= 16384, "the pile should form one island beyond the single-workgroup size"
        assert (items[:, 3] != 0).any(), "the largest island should run over the grid"
    finally:
        gpu.destroy(wg)
        cpu.destroy(wc)


def test_batch_4096_parity_sample_and_bounds():
    gpu, cpu = helpers.product(), helpers.oracle()
    n_worlds, seed = 4096, 7
    rng = np.random.default_rng(2024)
    sample = sorted(rng.choice(n_worlds, size=64, replace=False).tolist())
    batch = gpu.create_batch()
    gpu.use_batch(batch)
    worlds_g, bodies = [], 0
    try:
        for i in range(n_worlds):
            w, ids = gpu.scene("batch_pyramid", seed=seed, index=i)
            worlds_g.append(w)
            bodies += len(ids)
    finally:
        gpu.use_batch(None)
    worlds_c = {i: cpu.scene("batch_pyramid", seed=seed, index=i)[0] for i in sample}
    done = 0

    def advance_cpu(to):
        nonlocal done
        for wc in worlds_c.values():
            cpu.step(wc, n=to - done)
        done = to

    def check(step):
        advance_cpu(step)
        for i in sample:
            _compare(gpu, cpu, worlds_g[i], worlds_c[i], f"world {i} step {step}")

    try:
        assert bodies > 1_000_000
        for s in range(1, 21):
            gpu.batch_step(batch)
            if s in (1, 5, 20):
                check(s)
        gpu.lib.b2mi_Batch_Autotune(batch, 1.0 / 60.0, 4, 2)  # ordinary steps, 3 per candidate shape
        tuned = 20 + helpers.autotune_steps(gpu, batch, 2)
        check(tuned)
        gpu.batch_step(batch, n=100)
        check(tuned + 100)
        items = gpu.work_items(batch=batch)
        assert len(items) >= n_worlds and set(items[:, 2].tolist()) >= {2}
        # every world: finite state, bounded penetration, nothing through its ground box
        worst_pen, lowest = 0.0, np.inf
        for i, w in enumerate(worlds_g):
            b = gpu.dump_bodies(w)
            c = gpu.dump_contacts(w)
            dyn = b["type"] == 2
            assert np.isfinite(b["center"]).all() and np.isfinite(b["v"]).all() and np.isfinite(b["w"]).all(), i
            st = helpers.stability(b, c, None)
            worst_pen = max(worst_pen, st["penetration"])
            lowest = min(lowest, float(b["center"][dyn, 1].min()))
        print(f"batch: {bodies} bodies, {len(items)} items, worst penetration {worst_pen:.4f}, lowest center {lowest:.3f}")
        assert worst_pen < 0.03
        assert lowest > 0.4  # box centers stay above the ground's top face (y = 0) less a penetration
    finally:
        gpu.lib.b2mi_DestroyBatch(batch)
        for wc in worlds_c.values():
            cpu.destroy(wc)


def _oracle_drop_boxes(r, i, f, e, v0, w0):
    cpu = helpers.oracle()
    diff.bind_scene(cpu)
    d = cpu.default_world_def(enableSleep=False, enableContinous=False)
    w = WorldId()
    ids = (BodyId * r.boxes)()
    cpu.scenes.scene_drop_boxes_ex(ctypes.byref(d), r.seed, i, r.boxes, float(f[i]), float(e[i]), v0[i].ctypes.data,
                                   w0[i].ctypes.data, ctypes.byref(w), ids, r.boxes, None)
    cpu.step(w, n=r.steps)
    out = []
    for b in range(r.boxes):
        c = cpu.lib.b2Body_GetWorldCenterOfMass(ids[b])
        v = cpu.lib.b2Body_GetLinearVelocity(ids[b])
        out.append((c.x, c.y, v.x, v.y))
    cpu.destroy(w)
    return np.array(out, np.float32)


def _world_loss(out, target, i):
    d = out[i, :, 0:2].astype(np.float64) - target[i]
    return float(np.sum(d * d))


def _rel(a, b):
    a, b = np.ravel(a).astype(np.float64), np.ravel(b).astype(np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12))


def test_adjoint_1024x16_full_size():
    """The benchmark's differentiable configuration at its size: 1,024 worlds x 16 boxes, 120 taped
    steps, a checkpoint every 10. The taped run is the oracle's run bit for bit on sampled worlds;
    the float32 adjoint of the loss (squared distance of the final centers to seeded targets)
    agrees with torch autograd through the FP64 restatement of the same taped steps on sampled
    worlds. Over 120 steps the replay is only piecewise smooth (impulse clamps, friction-cone and
    restitution switches; the restitution skip on a zero max impulse is a jump), and the FP64
    trajectory, 1e-5 to 3e-4 away, sits on a neighbouring piece in many worlds (the gradients
    then differ by up to a few per cent); so the bar is a median relative error of 1e-2 over the
    sampled worlds and 5e-2 in every one."""
    import torch
    from tests import fp64_replay

    worlds, boxes, steps, checkpoint, seed = 1024, 16, 120, 10, 11
    r = diff.DropBoxesRollout(helpers.product(), worlds, boxes, steps, 4, checkpoint, seed=seed)
    f, e, v0, w0 = r.default_params()
    rng = np.random.default_rng(5)
    try:
        final = r.run(f, e, v0, w0)
        for i in sorted(rng.choice(worlds, size=16, replace=False).tolist()):
            want = _oracle_drop_boxes(r, i, f, e, v0, w0)
            got = final[i][:, [0, 1, 4, 5]]
            assert np.array_equal(want.view(np.uint32), got.view(np.uint32)), f"world {i}"

        target = np.random.default_rng(seed).uniform(-5, 5, size=final[:, :, 0:2].shape).astype(np.float32)
        g = np.zeros_like(final)
        g[:, :, 0:2] = 2.0 * (final[:, :, 0:2] - target)
        gf, ge, gv, gw, gs = r.backward(g)
        for a in (gf, ge, gv, gw):
            assert np.isfinite(a).all()
        params = np.stack([f, e], axis=1).astype(np.float32)

        errs = []
        # six sampled worlds: the FP64 autograd replay runs on the host, about 8 s per world
        for i in sorted(rng.choice(worlds, size=6, replace=False).tolist()):
            steps_, wp, mass, inertia, world_of = fp64_replay.export_worlds(r, [i])
            start = torch.tensor(r.start[[i]].reshape(-1, 7), dtype=torch.float64, requires_grad=True)
            p64 = torch.tensor(params[[i]], dtype=torch.float64, requires_grad=True)
            out = fp64_replay.replay(steps_, wp, mass, inertia, world_of, start, p64)
            err_pose = float(np.abs(out.detach().numpy() - final[[i]].reshape(-1, 7)).max())
            loss = (torch.tensor(g[[i]].reshape(-1, 7), dtype=torch.float64) * out).sum()
            g_start, g_params = torch.autograd.grad(loss, [start, p64])
            ours = np.concatenate([np.ravel(gs[i][:, 4:7]), [gf[i], ge[i]]])
            ref = np.concatenate([np.ravel(g_start.numpy().reshape(boxes, 7)[:, 4:7]), np.ravel(g_params.numpy())])
            errs.append(_rel(ours, ref))
            print(f"world {i}: pose err {err_pose:.1e}, gradient rel err {errs[-1]:.1e} (friction {gf[i]:.5g} vs "
                  f"{g_params[0, 0].item():.5g}, restitution {ge[i]:.5g} vs {g_params[0, 1].item():.5g})")
            assert err_pose < 1e-3
        print(f"fp64 autograd: median gradient rel err {np.median(errs):.1e}, max {max(errs):.1e}")
        assert np.median(errs) < 1e-2
        assert max(errs) < 5e-2
    finally:
        r.close()


def test_adjoint_1024x16_finite_differences():
    """Central finite differences (h = 1e-3) of the replayed taped steps on sampled coordinates of
    the same 1,024-world scene, over the last 10 steps of its 120 (110 untaped steps first, then a
    10-step tape with one checkpoint): on that window the replay is smooth at most coordinates. A
    coordinate counts when its central differences at h and h/2 agree to within the float32
    noise floor of the world's loss (4e-7 |L| / h) plus 1e-3 of their size; on those (at least half)
    the adjoint agrees with the h/2 quotient to 1e-3 relative plus twice the noise floor."""
    worlds, boxes, seed = 1024, 16, 11
    r = diff.DropBoxesRollout(helpers.product(), worlds, boxes, 10, 4, 10, seed=seed, pre_steps=110)
    f, e, v0, w0 = r.default_params()
    rng = np.random.default_rng(9)
    try:
        final = r.run(f, e, v0, w0)
        target = np.random.default_rng(seed).uniform(-5, 5, size=final[:, :, 0:2].shape).astype(np.float32)
        g = np.zeros_like(final)
        g[:, :, 0:2] = 2.0 * (final[:, :, 0:2] - target)
        gf, ge, gv, gw, gs = r.backward(g)
        params = np.stack([f, e], axis=1).astype(np.float32)
        h = 1e-3
        coords = []
        for i in sorted(rng.choice(worlds, size=10, replace=False).tolist()):
            b = int(rng.integers(boxes))
            coords += [(i, "friction", gf[i]), (i, "restitution", ge[i]), (i, ("vx", b), gv[i, b, 0]),
                       (i, ("w", b), gw[i, b])]
        rows = []
        for i, what, adj in coords:
            est = []
            for step in (h, h / 2):
                vals = []
                for sign in (1.0, -1.0):
                    s, p = r.start.copy(), params.copy()
                    if what == "friction":
                        p[i, 0] += sign * step
                    elif what == "restitution":
                        p[i, 1] += sign * step
                    elif what[0] == "vx":
                        s[i, what[1], 4] += sign * step
                    else:
                        s[i, what[1], 6] += sign * step
                    vals.append(_world_loss(r.replay(s, p), target, i))
                est.append((vals[0] - vals[1]) / (2 * step))
            rows.append((i, what, float(adj), est[0], est[1], 4e-7 * abs(_world_loss(final, target, i)) / h))
        stable = [row for row in rows if abs(row[3] - row[4]) <= row[5] + 1e-3 * abs(row[4])]
        bad = [row for row in stable if abs(row[2] - row[4]) > 1e-3 * abs(row[4]) + 2 * row[5]]
        print(f"adjoint vs central differences: {len(stable)}/{len(rows)} stable coordinates, {len(bad)} outside the bar")
        for row in rows:
            print("  ", row, "stable" if row in stable else "")
        assert len(stable) >= len(rows) // 2, rows
        assert not bad, bad
    finally:
        r.close()
