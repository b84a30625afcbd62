#!/usr/bin/env python3
"""Throughput of the HIP step pipeline in bodies x substeps / s on one MI355X.

Workload of the headline (`value`): one world holding a 40 x 40 grid of base-10 box pyramids
(88,000 dynamic boxes, half-extent 0.5, friction 0.6), dt = 1/60 s, 4 substeps, synthetic input
from box2c_amd/csrc/scenes.c. W warmup steps (the first creates every contact) are excluded;
steps W+1 .. W+K are timed. Every timed step runs the full pipeline: broadphase pair update,
narrowphase, colouring updates, the solver (prepare, 4 x (integrate velocities, warm start,
solve, integrate positions, relax), restitution, store impulses) and finalize.

Further legs, each an object in the same JSON line: the reference's own CPU benchmark scene
(large pyramid of base 100), the batch of 4,096 uneven pyramid worlds, the differentiable
1,024-world drop-box rollout with its adjoint, the 40,000-shape tumbler, and the reference's own
joint benchmarks (its motorized tumbler with 2,025 boxes, and the 100 x 100 joint grid).

Each leg carries
  roofline      its solver kernel (k_solve, the dominant launch): algorithmic bytes per launch
                (counted fields, each once per pass) / the kernel's mean duration from HIP events
                on the stream it runs on, against the 8 TB/s HBM peak; beside it the HBM bytes per
                launch measured by rocprofv3 counters (profiles/hbm_*.json) and the rate they give;
  cpu_baseline  the reference (oracle/_ref, built from its own sources) on this host's cores with a
                pthread task pool, over the same step window as the GPU leg (or a stated bounded
                part of it), at the thread count of the host's CPU share, plus a 1..N thread curve.
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
SUBSTEPS = 4
DT = 1.0 / 60.0


def algorithmic_bytes(n_bodies: int, n_contacts: int, substeps: int) -> dict:
    """Bytes one step must move, counting each field once per pass (per-body and per-contact
    figures derived from the reference's data layout: integrate velocities 60 B + integrate
    positions 44 B per body per substep; warm start 128 B + solve 232 B + relax 232 B per
    contact per substep; prepare 264 + restitution 232 + store 56 B per contact per step;
    finalize 172 B per body per step; narrowphase ~300 B per contact per step)."""
    solve = substeps * (104 * n_bodies + 592 * n_contacts) + (264 + 232 + 56) * n_contacts + 172 * n_bodies
    collide = 300 * n_contacts
    return {"solve": solve, "collide": collide, "step": solve + collide}


def roofline(workload: str, n_bodies: int, n_contacts: int, ms_solve: float, ms_collide: float = 0.0) -> dict:
    """The solver's algorithmic-byte rate against the HBM peak (achieved / peak = frac), with the
    measured HBM bytes of the same workload's solver per step (`traffic`, from the committed
    rocprofv3 counter summary profiles/hbm_<workload>.json when there is one) and the rate they
    give; the narrowphase kernel's figures beside them."""
    ab = algorithmic_bytes(n_bodies, n_contacts, SUBSTEPS)
    achieved = ab["solve"] / (ms_solve * 1e-3) / 1e9 if ms_solve > 0 else 0.0
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
           "traffic": None, "kernel": "k_solve", "algorithmic_bytes_per_launch": ab["solve"],
           "note": "achieved = algorithmic bytes (each field once per pass) / the step's solver time (HIP events around its "
                   "launches); the kernel keeps work items in LDS across passes, so this can exceed what crosses HBM: "
                   "traffic and measured_* are the HBM bytes rocprofv3 counts"}
    collide = {"kernel": "k_collide", "algorithmic_bytes_per_launch": ab["collide"],
               "achieved": ab["collide"] / (ms_collide * 1e-3) / 1e9 if ms_collide > 0 else 0.0}
    tfile = os.path.join(ROOT, "profiles", f"hbm_{workload}.json")
    if os.path.exists(tfile):
        t = json.load(open(tfile))
        if t.get("workload") == workload and t.get("solve_bytes_per_step"):
            traffic = float(t["solve_bytes_per_step"])
            launches = int(t.get("solve_launches_per_step", 1))
            out["traffic"] = traffic
            if launches > 1:
                out["kernel"] = f"solver ({launches} launches per step)"
                out["traffic_note"] = "per step: the sum over the step's solver launches"
            rate = traffic / (ms_solve * 1e-3) / 1e9 if ms_solve > 0 else 0.0
            out["measured_hbm_GBs"] = rate
            out["measured_hbm_frac"] = rate / HBM_PEAK_GBS
            if t.get("collide_bytes_per_launch") and ms_collide > 0:
                collide["traffic"] = float(t["collide_bytes_per_launch"])
                collide["measured_hbm_GBs"] = collide["traffic"] / (ms_collide * 1e-3) / 1e9
                collide["measured_hbm_frac"] = collide["measured_hbm_GBs"] / HBM_PEAK_GBS
                collide["traffic_over_algorithmic"] = collide["traffic"] / ab["collide"] if ab["collide"] else None
    out["collide"] = collide
    return out


# ---- host CPU share ------------------------------------------------------------------------

def host_cores() -> dict:
    """The CPU share this process may use: physical cores among the CPUs of its affinity mask,
    capped by a cgroup CPU quota and by the OMP_NUM_THREADS share when those are set."""
    try:
        cpus = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cpus = list(range(os.cpu_count() or 1))
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology"
        try:
            cores.add((open(f"{base}/physical_package_id").read().strip(), open(f"{base}/core_id").read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    physical = len(cores)
    quota = None
    try:
        q, p = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = max(1, int(int(q) / int(p)))
    except (OSError, ValueError):
        pass
    share = None
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        share = int(os.environ["OMP_NUM_THREADS"])
    used = physical
    for cap in (quota, share):
        if cap:
            used = min(used, cap)
    return {"affinity_cpus": len(cpus), "physical_cores": physical, "cgroup_quota": quota, "omp_share": share,
            "threads": max(1, used)}


def progress(msg: str):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def thread_curve(n: int) -> list:
    out, t = [], 1
    while t < n:
        out.append(t)
        t *= 2
    return out + [n]


def _ref():
    from box2c_amd.capi import B2Api
    ref_dir = os.path.join(ROOT, "oracle", "_ref")
    lib, scenes = os.path.join(ref_dir, "libbox2d_ref.so"), os.path.join(ref_dir, "libref_scenes.so")
    if not (os.path.exists(lib) and os.path.exists(scenes)):
        return None
    return B2Api(lib, scenes)


def cpu_world_rate(ref, scene: str, kw: dict, threads: int, warmup: int, steps: int):
    """The reference stepping one world with `threads` workers: W warmup steps, then `steps`
    timed (benchmark/main.c:196-209: wall clock over the steps). Returns (bodies x substeps / s,
    seconds)."""
    ref.lib.ref_pool_start(threads)
    try:
        d = ref.default_world_def(enableSleep=False)
        ref.lib.ref_pool_apply(ctypes.byref(d))
        w, ids = ref.scene(scene, d, **kw)
        ref.step(w, n=warmup)
        t0 = time.perf_counter()
        ref.step(w, n=steps)
        elapsed = time.perf_counter() - t0
        ref.destroy(w)
    finally:
        ref.lib.ref_pool_stop()
    return len(ids) * SUBSTEPS * steps / elapsed, elapsed


def cpu_world_baseline(scene: str, kw: dict, label: str, host: dict, warmup: int, steps: int, curve_steps: int):
    ref = _ref()
    if ref is None:
        return None
    n = host["threads"]
    value, elapsed = cpu_world_rate(ref, scene, kw, n, warmup, steps)
    curve = {}
    for t in thread_curve(n):
        if t == n and curve_steps == steps:
            curve[str(t)] = value
            continue
        curve[str(t)] = cpu_world_rate(ref, scene, kw, t, warmup, curve_steps)[0]
    return {"value": value, "unit": "bodies×substeps/s", "cores": n, "kind": "reference",
            "sample": f"{label}: steps {warmup + 1}..{warmup + steps} (the GPU leg's window), {n} worker threads "
                      f"(the host's CPU share: {host}), {elapsed:.1f} s",
            "thread_curve": curve,
            "thread_curve_sample": f"steps {warmup + 1}..{warmup + curve_steps} at each thread count"}


# ---- GPU legs ------------------------------------------------------------------------------

def run_world(api, scene: str, kw: dict, steps: int, warmup: int):
    d = api.default_world_def(enableSleep=False)
    w, ids = api.scene(scene, d, **kw)
    api.step(w, n=warmup)
    t_solve = t_collide = 0.0
    host = {"pairs": 0.0, "broadphase": 0.0, "buildIslands": 0.0, "finalizeBodies": 0.0, "step": 0.0}
    per_step = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        api.lib.b2World_Step(w, DT, SUBSTEPS)
        per_step.append(time.perf_counter() - ts)
        c, s, _h = api.world_timing(w)
        t_collide += c
        t_solve += s
        prof = api.lib.b2World_GetProfile(w)
        for k in host:
            host[k] += getattr(prof, k)
    elapsed = time.perf_counter() - t0
    counters = api.counters(w)
    items = api.work_items(world=w)
    api.destroy(w)
    return {
        "bodies": len(ids), "contacts": sum(counters.colorCounts), "elapsed": elapsed, "per_step": per_step,
        "ms_solve": t_solve / steps, "ms_collide": t_collide / steps, "colors": list(counters.colorCounts),
        "host_ms": {k: v / steps for k, v in host.items()},
        "work_items": {"count": int(len(items)), "largest_slots": int(items[:, 0].max()) if len(items) else 0,
                       "modes": sorted(set(items[:, 3].tolist())) if len(items) else []},
    }


def world_leg(api, name: str, scene: str, kw: dict, steps: int, warmup: int, host: dict, cpu: bool, cpu_steps: int,
              curve_steps: int):
    progress(f"{name}: GPU, {warmup} warmup + {steps} timed steps")
    r = run_world(api, scene, kw, steps, warmup)
    value = r["bodies"] * SUBSTEPS * steps / r["elapsed"]
    leg = {
        "workload": name, "value": value, "unit": "bodies×substeps/s", "bodies": r["bodies"],
        "touching_contacts": r["contacts"], "graph_colors": r["colors"], "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * r["elapsed"] / steps, "ms_solve_kernel": r["ms_solve"], "ms_collide_kernel": r["ms_collide"],
        "ms_step_phases": r["host_ms"], "work_items": r["work_items"],
        "roofline": roofline(name, r["bodies"], r["contacts"], r["ms_solve"], r["ms_collide"]), "cpu_baseline": None,
    }
    if cpu:
        progress(f"{name}: CPU reference, {cpu_steps} steps at {host['threads']} threads and the thread curve")
        cb = cpu_world_baseline(scene, kw, name, host, warmup, cpu_steps, curve_steps)
        leg["cpu_baseline"] = cb
        if cb:
            # GPU and CPU over the same steps (the CPU sample's window when it is shorter)
            k = min(cpu_steps, steps)
            window = r["bodies"] * SUBSTEPS * k / sum(r["per_step"][:k])
            leg["vs_cpu"] = window / cb["value"]
            if k < steps:
                leg["value_cpu_window"] = window
            leg["vs_cpu_note"] = f"GPU and CPU over the same steps {warmup + 1}..{warmup + k}"
    return leg


def build_batch(api, n_worlds: int, seed: int):
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
    return batch, n_bodies


def run_gpu_batch(api, n_worlds: int, steps: int, warmup: int, seed: int = 7, autotune: bool = True, shape=None):
    """4096 (default) independent worlds, each one pyramid of base U{4..40} on its own ground box,
    advanced together. The autotuner runs on a batch of its own (its 180 trial steps are ordinary
    steps), so the timed batch starts from step 0 like the CPU's: W warmup steps, K timed."""
    tuned = None
    if shape is None and autotune:
        tb, _ = build_batch(api, n_worlds, seed)
        api.batch_step(tb, n=warmup)
        best = api.lib.b2mi_Batch_Autotune(tb, DT, SUBSTEPS, 2)
        b, t, l, f = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        api.lib.b2mi_Batch_GetLaunchShape(tb, ctypes.byref(b), ctypes.byref(t), ctypes.byref(l), ctypes.byref(f))
        shape = (b.value, t.value, l.value, f.value)
        tuned = {"block": b.value, "target_slots": t.value, "lds_budget": l.value, "lds_full_budget": f.value,
                 "tuned_solve_ms": best}
        api.lib.b2mi_DestroyBatch(tb)
    batch, n_bodies = build_batch(api, n_worlds, seed)
    if shape is not None:
        api.lib.b2mi_Batch_Configure(batch, shape[0], shape[1])
        api.lib.b2mi_Batch_SetLdsBudget(batch, shape[2])
        api.lib.b2mi_Batch_SetLdsFullBudget(batch, shape[3] if len(shape) > 3 else -1)
    api.batch_step(batch, n=warmup)
    t_solve = t_collide = t_host = 0.0
    per_step = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        api.lib.b2mi_Batch_Step(batch, DT, SUBSTEPS)
        per_step.append(time.perf_counter() - ts)
        c, s, h = api.batch_timing(batch)
        t_collide += c
        t_solve += s
        t_host += h
    elapsed = time.perf_counter() - t0
    items = api.work_items(batch=batch)
    contacts = 0
    api.lib.b2mi_DestroyBatch(batch)
    return {"bodies": n_bodies, "elapsed": elapsed, "per_step": per_step, "ms_solve": t_solve / steps,
            "ms_collide": t_collide / steps, "ms_host": t_host / steps, "work_items": int(len(items)),
            "slots": int(items[:, 0].sum()) if len(items) else contacts,
            "launch_shape": tuned or ({"block": shape[0], "target_slots": shape[1], "lds_budget": shape[2],
                                       "lds_full_budget": shape[3] if len(shape) > 3 else -1} if shape else None)}


def cpu_baseline_batch(host: dict, n_worlds: int, warmup: int, steps: int, curve_steps: int, seed: int = 7):
    """The reference on every world of the batch with world-level parallelism (each thread steps
    whole worlds taken from a shared counter, largest first: the reference's best schedule for
    many small worlds). It holds at most 128 worlds at once (src/core.h:129), so the batch runs in
    waves of 128: each wave's worlds are built, stepped through the warmup, and timed over the
    same steps. The thread curve is taken on the first wave."""
    from box2c_amd.capi import WorldId
    ref = _ref()
    if ref is None:
        return None
    f = ref.lib.ref_step_worlds_parallel
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int]
    f.restype = ctypes.c_double
    d = ref.default_world_def(enableSleep=False)
    n = host["threads"]

    def wave(first, count, threads, timed):
        made = []
        for i in range(first, first + count):
            w, ids = ref.scene("batch_pyramid", d, seed=seed, index=i)
            made.append((len(ids), w))
        made.sort(key=lambda t: -t[0])
        arr = (WorldId * count)(*[w for _, w in made])
        f(arr, count, threads, warmup, DT, SUBSTEPS)
        elapsed = f(arr, count, threads, timed, DT, SUBSTEPS)
        for i in range(count):
            ref.destroy(arr[i])
        return sum(b for b, _ in made), elapsed

    bodies = elapsed = 0.0
    for first in range(0, n_worlds, 128):
        b, e = wave(first, min(128, n_worlds - first), n, steps)
        bodies += b
        elapsed += e
    value = bodies * SUBSTEPS * steps / elapsed
    curve = {}
    for t in thread_curve(n):
        b, e = wave(0, min(128, n_worlds), t, curve_steps)
        curve[str(t)] = b * SUBSTEPS * curve_steps / e
    return {"value": value, "unit": "bodies×substeps/s", "cores": n, "kind": "reference",
            "sample": f"all {n_worlds} batch-pyramid worlds ({int(bodies)} boxes) in waves of 128, steps {warmup + 1}.."
                      f"{warmup + steps} of each, worlds taken largest first by {n} threads (the host's CPU share: {host}), "
                      f"{elapsed:.1f} s of timed stepping",
            "thread_curve": curve,
            "thread_curve_sample": f"the first wave (128 worlds), steps {warmup + 1}..{warmup + curve_steps}"}


def batched_leg(api, args, host: dict, cpu: bool):
    shape = tuple(int(x) for x in args.batch_shape.split(",")) if args.batch_shape else None
    progress(f"batched: GPU, {args.batch_worlds} worlds")
    rb = run_gpu_batch(api, args.batch_worlds, args.batch_steps, args.batch_warmup, autotune=not args.batch_no_autotune,
                       shape=shape)
    bval = rb["bodies"] * SUBSTEPS * args.batch_steps / rb["elapsed"]
    name = f"batch_pyramids_{args.batch_worlds}_base4-40"
    leg = {
        "workload": name, "value": bval, "unit": "bodies×substeps/s", "bodies": rb["bodies"],
        "worlds": args.batch_worlds, "steps": args.batch_steps, "warmup": args.batch_warmup,
        "ms_per_step": 1e3 * rb["elapsed"] / args.batch_steps, "ms_solve_kernel": rb["ms_solve"],
        "ms_collide_kernel": rb["ms_collide"], "ms_host": rb["ms_host"], "work_items": rb["work_items"],
        "launch_shape": rb["launch_shape"], "cpu_baseline": None,
        "roofline": roofline(name, rb["bodies"], rb["slots"], rb["ms_solve"], rb["ms_collide"]),
    }
    k = min(args.batch_steps, args.batch_cpu_steps)
    leg["value_cpu_window"] = rb["bodies"] * SUBSTEPS * k / sum(rb["per_step"][:k])
    if cpu:
        progress(f"batched: CPU reference, {args.batch_worlds} worlds in waves of 128")
        cb = cpu_baseline_batch(host, args.batch_worlds, args.batch_warmup, k, min(k, 20))
        leg["cpu_baseline"] = cb
        if cb:
            leg["vs_cpu"] = leg["value_cpu_window"] / cb["value"]
            leg["vs_cpu_note"] = f"GPU and CPU over the same steps {args.batch_warmup + 1}..{args.batch_warmup + k}"
    return leg


def run_adjoint(api, worlds: int, boxes: int, steps: int, checkpoint: int, seed: int = 11):
    """Forward (taped) rollout of `worlds` drop-box worlds and the adjoint of the loss
    sum |p_final - target|^2 with respect to friction, restitution and initial velocities."""
    import numpy as np
    from box2c_amd import diff
    r = diff.DropBoxesRollout(api, worlds, boxes, steps, SUBSTEPS, checkpoint, seed=seed)
    f, e, v0, w0 = r.default_params()
    try:
        r.run(f, e, v0, w0)  # warm-up: kernels, allocations
        final = r.run(f, e, v0, w0)
        target = np.random.default_rng(seed).uniform(-5, 5, size=final[:, :, 0:2].shape).astype(np.float32)
        g = np.zeros_like(final)
        g[:, :, 0:2] = 2.0 * (final[:, :, 0:2] - target)
        gf, ge, gv, gw, _ = r.backward(g)
        fwd, bwd = r.forward_seconds, r.backward_seconds
        r.run(f, e, v0, w0, taped=False)  # the same rollout without a tape
        plain = r.forward_seconds
        finite = bool(np.isfinite(gf).all() and np.isfinite(ge).all() and np.isfinite(gv).all() and np.isfinite(gw).all())
    finally:
        r.close()
    n = worlds * boxes
    return {
        "workload": f"drop_boxes_{worlds}x{boxes}_{steps}steps_adjoint", "bodies": n, "steps": steps,
        "checkpoint_interval": checkpoint, "forward_s": fwd, "backward_s": bwd, "adjoint_over_forward": bwd / fwd,
        "value_forward": n * SUBSTEPS * steps / fwd, "value_fwd_plus_adjoint": n * SUBSTEPS * steps / (fwd + bwd),
        "untaped_forward_s": plain, "value_untaped": n * SUBSTEPS * steps / plain,
        "fwd_plus_adjoint_over_untaped": (fwd + bwd) / plain,
        "unit": "bodies×substeps/s", "gradients_finite": finite,
        "note": "forward = the taped GPU step loop (narrowphase, solver, recording); backward = replay of each "
                "checkpoint segment plus the adjoint kernels; manifolds held fixed per step; continuous collision off",
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
    ap.add_argument("--steps", type=int, default=300, help="timed steps of the headline (pyramid grid) leg")
    ap.add_argument("--warmup", type=int, default=20, help="warmup steps of the headline leg")
    ap.add_argument("--rows", type=int, default=40)
    ap.add_argument("--cols", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU worker threads (0 = the host's CPU share)")
    ap.add_argument("--cpu-curve-steps", type=int, default=40, help="steps of each point of the CPU thread curve")
    ap.add_argument("--pyramid-steps", type=int, default=500, help="timed steps of the large-pyramid leg")
    ap.add_argument("--batch-worlds", type=int, default=4096, help="worlds of the batched leg (0 = skip it)")
    ap.add_argument("--batch-steps", type=int, default=300)
    ap.add_argument("--batch-warmup", type=int, default=5)
    ap.add_argument("--batch-cpu-steps", type=int, default=30, help="steps of the batched leg's CPU sample")
    ap.add_argument("--batch-no-autotune", action="store_true", help="batched leg at the default launch shape")
    ap.add_argument("--batch-shape", default="", help="block,target,ldsBudget[,ldsFullBudget]: skip the autotuner, use this shape")
    ap.add_argument("--adjoint-worlds", type=int, default=1024, help="worlds of the adjoint leg (0 = skip it)")
    ap.add_argument("--tumbler-shapes", type=int, default=40000, help="shapes of the tumbler leg (0 = skip it)")
    ap.add_argument("--tumbler-steps", type=int, default=600)
    ap.add_argument("--tumbler-cpu-steps", type=int, default=600, help="steps of the tumbler leg's CPU sample")
    ap.add_argument("--joint-tumbler-steps", type=int, default=750, help="timed steps of the joint tumbler leg")
    ap.add_argument("--joint-grid-steps", type=int, default=500, help="timed steps of the joint grid leg")
    ap.add_argument("--legs", default="plain,large_pyramid,batched,adjoint,tumbler,joint_tumbler,joint_grid",
                    help="comma-separated legs to run (a profiling run may keep one; the JSON line then lacks the others)")
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
    host = host_cores()
    if args.cpu_threads > 0:
        host["threads"] = args.cpu_threads
    cpu = not args.no_cpu
    out = {"metric": METRIC}
    grid = f"pyramid_grid_{args.rows}x{args.cols}_base10"
    if "plain" in legs:
        leg = world_leg(api, grid, "pyramid_grid", {"rows": args.rows, "cols": args.cols, "base": 10, "seed": 7},
                        args.steps, args.warmup, host, cpu, args.steps, min(args.steps, args.cpu_curve_steps))
        out.update({
            "value": leg["value"], "unit": "bodies×substeps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": leg["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": grid, "bodies": leg["bodies"], "touching_contacts": leg["touching_contacts"],
                       "substeps": SUBSTEPS, "dt": DT, "graph_colors": leg["graph_colors"],
                       "ms_solve_kernel": leg["ms_solve_kernel"], "ms_collide_kernel": leg["ms_collide_kernel"],
                       "ms_step_phases": leg["ms_step_phases"], "work_items": leg["work_items"]},
            "roofline": leg["roofline"], "cpu_baseline": leg["cpu_baseline"],
        })
        if leg.get("vs_cpu"):
            out["vs_cpu"] = leg["vs_cpu"]
    if "large_pyramid" in legs:
        # the reference's own benchmark scene (benchmark/large_pyramid.c): one warm-up step, then timed steps
        out["large_pyramid"] = world_leg(api, "large_pyramid_base100", "large_pyramid", {"base": 100}, args.pyramid_steps, 1,
                                         host, cpu, args.pyramid_steps, min(args.pyramid_steps, 100))
    if "batched" in legs and args.batch_worlds > 0:
        out["batched"] = batched_leg(api, args, host, cpu)
    if "adjoint" in legs and args.adjoint_worlds > 0:
        progress("adjoint: taped forward and backward")
        out["adjoint"] = run_adjoint(api, args.adjoint_worlds, 16, 120, 10)
    if "tumbler" in legs and args.tumbler_shapes > 0:
        out["tumbler"] = world_leg(api, f"tumbler_{args.tumbler_shapes}", "tumbler", {"count": args.tumbler_shapes, "seed": 5},
                                   args.tumbler_steps, args.warmup, host, cpu, min(args.tumbler_steps, args.tumbler_cpu_steps),
                                   min(args.tumbler_steps, args.tumbler_cpu_steps, 30))
    # the reference's own joint benchmarks (benchmark/tumbler.c: a motorized revolute container with
    # 45 x 45 boxes, 750 steps; benchmark/joint_grid.c: 100 x 100 circles on 19,800 revolute joints
    # in zero gravity, 500 steps), each after one warm-up step as benchmark/main.c:194-205 runs them
    if "joint_tumbler" in legs and args.joint_tumbler_steps > 0:
        out["joint_tumbler"] = world_leg(api, "joint_tumbler_45x45", "joint_tumbler", {"grid": 45}, args.joint_tumbler_steps, 1, host,
                                         cpu, args.joint_tumbler_steps, min(args.joint_tumbler_steps, 100))
    if "joint_grid" in legs and args.joint_grid_steps > 0:
        out["joint_grid"] = world_leg(api, "joint_grid_100x100", "joint_grid", {"N": 100, "gravity": 0.0}, args.joint_grid_steps, 1,
                                      host, cpu, args.joint_grid_steps, min(args.joint_grid_steps, 100))
    out["host"] = host
    progress("done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
