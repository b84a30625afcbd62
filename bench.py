#!/usr/bin/env python3
"""Throughput of the HIP step pipeline in bodies x substeps / s on one MI355X.

Workload (default): one world holding a 40 x 40 grid of base-10 box pyramids (88,000
dynamic boxes, half-extent 0.5, friction 0.6), dt = 1/60 s, 4 substeps, synthetic input
from box2c_amd/csrc/scenes.c. W warmup steps (the first creates every contact) are excluded;
K steps are timed. Every step of the timed region runs the full pipeline: broadphase pair
update, narrowphase, colouring updates, the solver (prepare, 4 x (integrate velocities, warm
start, solve, integrate positions, relax), restitution, store impulses) and finalize.

Printed JSON line: metric/value/unit per the benchmark contract, plus
  roofline      the solver kernel (k_solve, the dominant launch): algorithmic bytes per
                launch / its mean duration from HIP events on its stream;
  cpu_baseline  the reference (oracle/_ref, built from its own sources) on the host cores
                with a pthread task pool, stepping the same scene for a bounded sample.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "bodies×substeps/s (1 GPU): plain vs batched vs fwd+adjoint at ≤1e-5 step error"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(n_bodies: int, n_contacts: int, substeps: int) -> dict:
    """Bytes one step must move, counting each field once per pass (per-body and per-contact
    figures derived from the reference's data layout: integrate velocities 60 B + integrate
    positions 44 B per body per substep; warm start 128 B + solve 232 B + relax 232 B per
    contact per substep; prepare 264 + restitution 232 + store 56 B per contact per step;
    finalize 172 B per body per step; narrowphase ~300 B per contact per step)."""
    solve = substeps * (104 * n_bodies + 592 * n_contacts) + (264 + 232 + 56) * n_contacts + 172 * n_bodies
    collide = 300 * n_contacts
    return {"solve": solve, "collide": collide, "step": solve + collide}


def run_gpu(api, scene: str, steps: int, warmup: int, rows: int, cols: int):
    d = api.default_world_def(enableSleep=False)
    w, ids = api.scene(scene, d, rows=rows, cols=cols, base=10, seed=7)
    n_bodies = len(ids)
    api.step(w, n=warmup)
    counters = api.counters(w)
    t_solve = t_collide = 0.0
    host = {"pairs": 0.0, "broadphase": 0.0, "buildIslands": 0.0, "finalizeBodies": 0.0, "step": 0.0}
    t0 = time.perf_counter()
    for _ in range(steps):
        api.lib.b2World_Step(w, 1.0 / 60.0, 4)
        c, s, _h = api.world_timing(w)
        t_collide += c
        t_solve += s
        prof = api.lib.b2World_GetProfile(w)
        for k in host:
            host[k] += getattr(prof, k)
    elapsed = time.perf_counter() - t0
    counters = api.counters(w)
    n_contacts = sum(counters.colorCounts)
    api.destroy(w)
    return {
        "bodies": n_bodies, "contacts": n_contacts, "elapsed": elapsed,
        "ms_solve": t_solve / steps, "ms_collide": t_collide / steps,
        "colors": list(counters.colorCounts),
        "host_ms": {k: v / steps for k, v in host.items()},
    }


def run_gpu_batch(api, n_worlds: int, steps: int, warmup: int, seed: int = 7, autotune: bool = True):
    """4096 (default) independent worlds, each one pyramid of base U{4..40} on its own ground box,
    advanced together: two kernel launches per step for all worlds."""
    d = api.default_world_def(enableSleep=False)
    batch = api.create_batch()
    api.use_batch(batch)
    n_bodies = 0
    try:
        for i in range(n_worlds):
            _, ids = api.scene("batch_pyramid", d, seed=seed, index=i)
            n_bodies += len(ids)
    finally:
        api.use_batch(None)
    api.batch_step(batch, n=warmup)
    shape = None
    if autotune:
        best = api.lib.b2mi_Batch_Autotune(batch, 1.0 / 60.0, 4, 2)
        b, t, l = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        api.lib.b2mi_Batch_GetLaunchShape(batch, ctypes.byref(b), ctypes.byref(t), ctypes.byref(l))
        shape = {"block": b.value, "target_slots": t.value, "lds_budget": l.value, "tuned_solve_ms": best}
        api.batch_step(batch)  # rebuilds the work items under the chosen shape
    t_solve = t_collide = t_host = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        api.lib.b2mi_Batch_Step(batch, 1.0 / 60.0, 4)
        c, s, h = api.batch_timing(batch)
        t_collide += c
        t_solve += s
        t_host += h
    elapsed = time.perf_counter() - t0
    items = api.lib.b2mi_Batch_GetWorkItemCount(batch)
    api.lib.b2mi_DestroyBatch(batch)
    return {"bodies": n_bodies, "elapsed": elapsed, "ms_solve": t_solve / steps, "ms_collide": t_collide / steps,
            "ms_host": t_host / steps, "work_items": items, "launch_shape": shape}


def run_tumbler(api, shapes: int, steps: int, warmup: int, seed: int = 5):
    """A kinematic square container spinning at 0.25 rad/s, filled with `shapes` half boxes and
    half circles of size 0.125 on a jittered 0.4 m grid (the container grows with the count)."""
    d = api.default_world_def(enableSleep=False)
    w, ids = api.scene("tumbler", d, count=shapes, seed=seed)
    api.step(w, n=warmup)
    t_solve = t_collide = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        api.lib.b2World_Step(w, 1.0 / 60.0, 4)
        c, s, _h = api.world_timing(w)
        t_collide += c
        t_solve += s
    elapsed = time.perf_counter() - t0
    counters = api.counters(w)
    api.destroy(w)
    return {"bodies": len(ids), "contacts": sum(counters.colorCounts), "colors": list(counters.colorCounts),
            "elapsed": elapsed, "ms_solve": t_solve / steps, "ms_collide": t_collide / steps}


def cpu_baseline_tumbler(shapes: int, threads: int, sample_steps: int, warmup: int, seed: int = 5):
    """The reference on the same tumbler with `threads` workers, over a bounded number of steps."""
    from box2c_amd.capi import B2Api
    ref_dir = os.path.join(ROOT, "oracle", "_ref")
    lib, scenes = os.path.join(ref_dir, "libbox2d_ref.so"), os.path.join(ref_dir, "libref_scenes.so")
    if not (os.path.exists(lib) and os.path.exists(scenes)):
        return None
    ref = B2Api(lib, scenes)
    ref.lib.ref_pool_start(threads)
    d = ref.default_world_def(enableSleep=False)
    ref.lib.ref_pool_apply(__import__("ctypes").byref(d))
    w, ids = ref.scene("tumbler", d, count=shapes, seed=seed)
    ref.step(w, n=warmup)
    t0 = time.perf_counter()
    ref.step(w, n=sample_steps)
    elapsed = time.perf_counter() - t0
    ref.destroy(w)
    ref.lib.ref_pool_stop()
    return {"value": len(ids) * 4 * sample_steps / elapsed, "unit": "bodies×substeps/s", "cores": threads,
            "kind": "reference",
            "sample": f"tumbler of {len(ids)} shapes, {sample_steps} steps after {warmup} warmup, {threads} worker threads, "
                      f"{elapsed:.1f} s"}


def run_adjoint(api, worlds: int, boxes: int, steps: int, checkpoint: int, seed: int = 11):
    """Forward (taped) rollout of `worlds` drop-box worlds and the adjoint of the loss
    sum |p_final - target|^2 with respect to friction, restitution and initial velocities."""
    import numpy as np
    from box2c_amd import diff
    r = diff.DropBoxesRollout(api, worlds, boxes, steps, 4, checkpoint, seed=seed)
    f, e, v0, w0 = r.default_params()
    try:
        r.run(f, e, v0, w0)  # warm-up: kernels, allocations
        final = r.run(f, e, v0, w0)
        target = np.random.default_rng(seed).uniform(-5, 5, size=final[:, :, 0:2].shape).astype(np.float32)
        g = np.zeros_like(final)
        g[:, :, 0:2] = 2.0 * (final[:, :, 0:2] - target)
        gf, ge, gv, gw, _ = r.backward(g)
        fwd, bwd = r.forward_seconds, r.backward_seconds
        finite = bool(np.isfinite(gf).all() and np.isfinite(ge).all() and np.isfinite(gv).all() and np.isfinite(gw).all())
    finally:
        r.close()
    n = worlds * boxes
    return {
        "workload": f"drop_boxes_{worlds}x{boxes}_{steps}steps_adjoint_f32", "bodies": n, "steps": steps,
        "checkpoint_interval": checkpoint, "forward_s": fwd, "backward_s": bwd, "adjoint_over_forward": bwd / fwd,
        "value_forward": n * 4 * steps / fwd, "value_fwd_plus_adjoint": n * 4 * steps / (fwd + bwd),
        "unit": "bodies×substeps/s", "gradients_finite": finite,
        "note": "forward = the taped GPU step loop (narrowphase, solver, recording); backward = replay of each "
                "checkpoint segment plus the adjoint kernels; manifolds held fixed per step; continuous collision off",
    }


def cpu_baseline_batch(threads: int, n_worlds: int, sample_steps: int, warmup: int, seed: int = 7):
    """The reference on the same kind of worlds with world-level parallelism (each thread steps
    its own worlds, the reference's best schedule for many small worlds). The reference holds at
    most 128 worlds at once, so the sample is one wave: the first `n_worlds` (<= 128) worlds of
    the batch."""
    import ctypes
    from box2c_amd.capi import B2Api, WorldId
    ref_dir = os.path.join(ROOT, "oracle", "_ref")
    lib, scenes = os.path.join(ref_dir, "libbox2d_ref.so"), os.path.join(ref_dir, "libref_scenes.so")
    if not (os.path.exists(lib) and os.path.exists(scenes)):
        return None
    ref = B2Api(lib, scenes)
    f = ref.lib.ref_step_worlds_parallel
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int]
    f.restype = ctypes.c_double
    d = ref.default_world_def(enableSleep=False)
    made = []
    for i in range(n_worlds):
        w, ids = ref.scene("batch_pyramid", d, seed=seed, index=i)
        made.append((len(ids), w))
    made.sort(key=lambda t: -t[0])  # largest first: threads take worlds from a shared counter
    worlds = (WorldId * n_worlds)(*[w for _, w in made])
    n_bodies = sum(n for n, _ in made)
    f(worlds, n_worlds, threads, warmup, 1.0 / 60.0, 4)
    elapsed = f(worlds, n_worlds, threads, sample_steps, 1.0 / 60.0, 4)
    for i in range(n_worlds):
        ref.destroy(worlds[i])
    return {
        "value": n_bodies * 4 * sample_steps / elapsed, "unit": "bodies×substeps/s", "cores": threads,
        "kind": "reference",
        "sample": f"{n_worlds} batch-pyramid worlds ({n_bodies} boxes, the first wave of the batch), {sample_steps} steps "
                  f"after {warmup} warmup, worlds taken largest first by {threads} threads, {elapsed:.1f} s",
    }


def cpu_baseline(rows: int, cols: int, threads: int, sample_steps: int, warmup: int = 20):
    """The reference stepping the same scene with `threads` workers, timed over a bounded
    number of steps after the same warmup as the GPU leg (bench.py's CPU leg; the oracle directory holds the
    reference build and is used here only as the baseline)."""
    from box2c_amd.capi import B2Api
    ref_dir = os.path.join(ROOT, "oracle", "_ref")
    lib, scenes = os.path.join(ref_dir, "libbox2d_ref.so"), os.path.join(ref_dir, "libref_scenes.so")
    if not (os.path.exists(lib) and os.path.exists(scenes)):
        return None
    ref = B2Api(lib, scenes)
    ref.lib.ref_pool_start(threads)
    d = ref.default_world_def(enableSleep=False)
    ref.lib.ref_pool_apply(__import__("ctypes").byref(d))
    w, ids = ref.scene("pyramid_grid", d, rows=rows, cols=cols, base=10, seed=7)
    ref.step(w, n=warmup)
    t0 = time.perf_counter()
    ref.step(w, n=sample_steps)
    elapsed = time.perf_counter() - t0
    ref.destroy(w)
    ref.lib.ref_pool_stop()
    return {
        "value": len(ids) * 4 * sample_steps / elapsed, "unit": "bodies×substeps/s", "cores": threads,
        "kind": "reference",
        "sample": f"pyramid_grid {rows}x{cols} ({len(ids)} boxes), {sample_steps} steps after {warmup} warmup, "
                  f"{threads} worker threads, {elapsed:.1f} s",
    }


def calibrate_copy(mib: int = 1024, reps: int = 5):
    """A plain device copy of known size (torch elementwise kernel, `mib` MiB read and `mib`
    MiB written per launch), run under the same HBM counters as the solver so that the
    counters' bytes per dispatch can be scaled to true bytes."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, device="cuda:0", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(reps):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kernel_bytes_read": n * 4, "copy_kernel_bytes_written": n * 4, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--rows", type=int, default=40)
    ap.add_argument("--cols", type=int, default=40)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=600)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-worlds", type=int, default=4096, help="worlds of the batched leg (0 = skip it)")
    ap.add_argument("--batch-steps", type=int, default=300)
    ap.add_argument("--batch-cpu-steps", type=int, default=1000)
    ap.add_argument("--adjoint-worlds", type=int, default=1024, help="worlds of the adjoint leg (0 = skip it)")
    ap.add_argument("--tumbler-shapes", type=int, default=40000, help="shapes of the tumbler leg (0 = skip it)")
    ap.add_argument("--tumbler-steps", type=int, default=600)
    ap.add_argument("--tumbler-cpu-steps", type=int, default=60)
    ap.add_argument("--legs", default="plain,batched,adjoint,tumbler",
                    help="comma-separated legs to run (a profiling run may keep one; the JSON line then lacks the others)")
    ap.add_argument("--batch-no-autotune", action="store_true", help="batched leg at the default launch shape")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain copy kernel of known size (for HBM counter scaling)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: worlds are independent and the multi-GPU "
                                   "layout (worlds sharded across GPUs) is design only; run with --gpus 1"}))
        sys.exit(2)

    import box2c_amd
    api = box2c_amd.load()
    if api.lib.b2mi_DeviceCount() < 1:
        raise SystemExit("no HIP device visible")

    legs = set(args.legs.split(","))
    substeps = 4
    if "plain" not in legs:
        out = {"metric": METRIC, "legs": sorted(legs)}
        if "batched" in legs and args.batch_worlds > 0:
            rb = run_gpu_batch(api, args.batch_worlds, args.batch_steps, args.warmup, autotune=not args.batch_no_autotune)
            out["batched"] = {"value": rb["bodies"] * substeps * args.batch_steps / rb["elapsed"],
                              "ms_per_step": 1e3 * rb["elapsed"] / args.batch_steps, "launch_shape": rb["launch_shape"]}
        if "adjoint" in legs and args.adjoint_worlds > 0:
            out["adjoint"] = run_adjoint(api, args.adjoint_worlds, 16, 120, 10)
        if "tumbler" in legs and args.tumbler_shapes > 0:
            rt = run_tumbler(api, args.tumbler_shapes, args.tumbler_steps, args.warmup)
            out["tumbler"] = {"value": rt["bodies"] * substeps * args.tumbler_steps / rt["elapsed"],
                              "ms_per_step": 1e3 * rt["elapsed"] / args.tumbler_steps}
        print(json.dumps(out), flush=True)
        return
    r = run_gpu(api, "pyramid_grid", args.steps, args.warmup, args.rows, args.cols)
    value = r["bodies"] * substeps * args.steps / r["elapsed"]
    ab = algorithmic_bytes(r["bodies"], r["contacts"], substeps)
    achieved = ab["solve"] / (r["ms_solve"] * 1e-3) / 1e9

    # HBM bytes per k_solve launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE passes of the same
    # workload, scaled by the counters' reading of a copy kernel of known size
    workload = f"pyramid_grid_{args.rows}x{args.cols}_base10_plain_f32"
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"hbm_pyramid_grid_{args.rows}x{args.cols}_base10.json")
    if os.path.exists(tfile):
        t = json.load(open(tfile))
        if t.get("workload") == workload:
            traffic = t.get("k_solve_bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "bodies×substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "bodies": r["bodies"], "touching_contacts": r["contacts"], "substeps": substeps,
            "dt": 1.0 / 60.0, "graph_colors": r["colors"],
            "ms_solve_kernel": r["ms_solve"], "ms_collide_kernel": r["ms_collide"],
            "ms_step_phases": r["host_ms"],
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": "k_solve", "algorithmic_bytes_per_launch": ab["solve"],
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.rows, args.cols, args.cpu_threads, args.cpu_steps, args.warmup)
    if args.batch_worlds > 0:
        rb = run_gpu_batch(api, args.batch_worlds, args.batch_steps, args.warmup)
        bval = rb["bodies"] * substeps * args.batch_steps / rb["elapsed"]
        batched = {
            "workload": f"batch_pyramids_{args.batch_worlds}_base4-40_plain_f32", "value": bval,
            "unit": "bodies×substeps/s", "bodies": rb["bodies"], "worlds": args.batch_worlds,
            "steps": args.batch_steps, "warmup": args.warmup, "ms_per_step": 1e3 * rb["elapsed"] / args.batch_steps,
            "ms_solve_kernel": rb["ms_solve"], "ms_collide_kernel": rb["ms_collide"], "ms_host": rb["ms_host"],
            "work_items": rb["work_items"], "launch_shape": rb["launch_shape"], "cpu_baseline": None,
        }
        if not args.no_cpu:
            cb = cpu_baseline_batch(args.cpu_threads, min(128, args.batch_worlds), args.batch_cpu_steps, args.warmup)
            batched["cpu_baseline"] = cb
            if cb:
                batched["vs_cpu"] = bval / cb["value"]
        out["batched"] = batched
    if args.adjoint_worlds > 0:
        out["adjoint"] = run_adjoint(api, args.adjoint_worlds, 16, 120, 10)
    if args.tumbler_shapes > 0:
        rt = run_tumbler(api, args.tumbler_shapes, args.tumbler_steps, args.warmup)
        tval = rt["bodies"] * substeps * args.tumbler_steps / rt["elapsed"]
        tumbler = {
            "workload": f"tumbler_{args.tumbler_shapes}_kinematic_container_plain_f32", "value": tval,
            "unit": "bodies×substeps/s", "bodies": rt["bodies"], "touching_contacts": rt["contacts"],
            "graph_colors": rt["colors"], "steps": args.tumbler_steps, "warmup": args.warmup,
            "ms_per_step": 1e3 * rt["elapsed"] / args.tumbler_steps, "ms_solve_kernel": rt["ms_solve"],
            "ms_collide_kernel": rt["ms_collide"], "cpu_baseline": None,
        }
        if not args.no_cpu:
            ct = cpu_baseline_tumbler(args.tumbler_shapes, args.cpu_threads, args.tumbler_cpu_steps, args.warmup)
            tumbler["cpu_baseline"] = ct
            if ct:
                tumbler["vs_cpu"] = tval / ct["value"]
        out["tumbler"] = tumbler
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
