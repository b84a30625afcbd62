"""Parity of the HIP step pipeline with the CPU oracle.

Bar: bit-identical. After every step, every awake contact record (contact id, shape pair,
graph colour, index in its colour array, flags, friction, restitution, the manifold with
feature ids, anchors, separations and accumulated impulses) and every body record (pose,
center, velocity) must match the oracle byte for byte; struct padding is excluded.
"""
import ctypes

import numpy as np
import pytest

from box2c_amd.capi import BodyId, Vec2
from tests import helpers

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,kw,overrides,steps,substeps", helpers.PARITY_SCENES,
                         ids=[s[0] for s in helpers.PARITY_SCENES])
def test_step_parity_with_oracle(name, kw, overrides, steps, substeps):
    gpu, cpu = helpers.product(), helpers.oracle()
    wg, _ = helpers.make_world(gpu, name, kw, overrides)
    wc, _ = helpers.make_world(cpu, name, kw, overrides)
    try:
        for s in range(steps):
            gpu.step(wg, substeps=substeps)
            cpu.step(wc, substeps=substeps)
            m = helpers.first_contact_mismatch(gpu.dump_contacts(wg), cpu.dump_contacts(wc))
            assert m is None, f"{name} step {s}: contacts differ: {m}"
            m = helpers.first_body_mismatch(gpu.dump_bodies(wg), cpu.dump_bodies(wc))
            assert m is None, f"{name} step {s}: bodies differ: {m}"
        cg, cc = gpu.counters(wg), cpu.counters(wc)
        assert list(cg.colorCounts) == list(cc.colorCounts)
        assert cg.contactCount == cc.contactCount
    finally:
        gpu.destroy(wg)
        cpu.destroy(wc)


def test_batch_parity_with_oracle():
    """16 uneven pyramid worlds in one batch launch == the same worlds stepped one by one."""
    gpu, cpu = helpers.product(), helpers.oracle()
    batch = gpu.create_batch()
    gpu.use_batch(batch)
    worlds_g, worlds_c = [], []
    try:
        for i in range(16):
            wg, _ = gpu.scene("batch_pyramid", seed=11, index=i)
            wc, _ = cpu.scene("batch_pyramid", seed=11, index=i)
            worlds_g.append(wg)
            worlds_c.append(wc)
    finally:
        gpu.use_batch(None)
    try:
        for s in range(40):
            gpu.batch_step(batch)
            for wc in worlds_c:
                cpu.step(wc)
            if s % 8 == 7:
                for i, (wg, wc) in enumerate(zip(worlds_g, worlds_c)):
                    m = helpers.first_contact_mismatch(gpu.dump_contacts(wg), cpu.dump_contacts(wc))
                    assert m is None, f"world {i} step {s}: {m}"
                    m = helpers.first_body_mismatch(gpu.dump_bodies(wg), cpu.dump_bodies(wc))
                    assert m is None, f"world {i} step {s}: {m}"
    finally:
        gpu.lib.b2mi_DestroyBatch(batch)
        for wc in worlds_c:
            cpu.destroy(wc)


def test_gpu_run_to_run_bit_reproducible():
    gpu = helpers.product()
    dumps = []
    for _ in range(2):
        w, _ = helpers.make_world(gpu, "tumbler", {"count": 300, "seed": 4}, {})
        gpu.step(w, n=60)
        dumps.append((helpers.masked(gpu.dump_contacts(w)), gpu.dump_bodies(w)))
        gpu.destroy(w)
    assert np.array_equal(dumps[0][0].view(np.uint8), dumps[1][0].view(np.uint8))
    assert np.array_equal(dumps[0][1].view(np.uint8), dumps[1][1].view(np.uint8))


def test_hello_world_rests_on_ground():
    """The reference's HelloWorld test (test/test_world.c:17-97): a box dropped on the
    ground for 90 steps ends within 0.01 of (0, 1) with angle below 0.01. Sleep is off here: with
    the reference's default (on) the box's island reaches the sleep time at about step 80, which
    this pipeline reports as unsupported (test_island_sleep_fails_loudly)."""
    gpu = helpers.product()
    d = gpu.default_world_def(enableSleep=False)
    w, ids = gpu.scene("hello", d)
    gpu.step(w, n=90)
    p = gpu.lib.b2Body_GetPosition(ids[0])
    q = gpu.lib.b2Body_GetRotation(ids[0])
    gpu.destroy(w)
    assert abs(p.x) < 0.01 and abs(p.y - 1.0) < 0.01
    assert abs(np.arctan2(q.s, q.c)) < 0.01


@pytest.mark.parametrize("name,kw,overrides,steps,substeps", helpers.PARITY_SCENES[:4] + helpers.PARITY_SCENES[6:],
                         ids=[s[0] for s in helpers.PARITY_SCENES[:4] + helpers.PARITY_SCENES[6:]])
@pytest.mark.parametrize("budget", [0, 8 * 1024, 24 * 1024], ids=["hbm", "lds8k", "lds24k"])
def test_step_parity_solver_storage(name, kw, overrides, steps, substeps, budget):
    """The solver keeps a work item's state in LDS when it fits the budget, only its bodies in LDS
    when those fit, and everything in HBM otherwise; all give the oracle's bits (budget 0 forces
    HBM for every item, 8 KiB makes the larger items keep only their bodies in LDS)."""
    gpu, cpu = helpers.product(), helpers.oracle()
    wg, _ = helpers.make_world(gpu, name, kw, overrides)
    wc, _ = helpers.make_world(cpu, name, kw, overrides)
    gpu.lib.b2mi_World_SetLdsBudget(wg, budget)
    try:
        for s in range(min(steps, 40)):
            gpu.step(wg, substeps=substeps)
            cpu.step(wc, substeps=substeps)
        m = helpers.first_contact_mismatch(gpu.dump_contacts(wg), cpu.dump_contacts(wc))
        assert m is None, f"{name}: contacts differ: {m}"
        m = helpers.first_body_mismatch(gpu.dump_bodies(wg), cpu.dump_bodies(wc))
        assert m is None, f"{name}: bodies differ: {m}"
    finally:
        gpu.destroy(wg)
        cpu.destroy(wc)


@pytest.mark.parametrize("name", ["hello", "large_pyramid", "tumbler", "shape_zoo", "drop_boxes", "overflow",
                                  "overflow_negzero"])
@pytest.mark.parametrize("grid", [-1, 0, 3], ids=["per_phase", "grid_auto", "grid3"])
def test_step_parity_split_items(name, grid):
    """Every work item kept in HBM and run over the whole grid (split threshold of one slot): as
    one cooperative launch with grid-wide barriers (its automatic workgroup count, or three), or
    as one launch per phase; each gives the oracle's bits, overflow levels and restitution
    included."""
    scene = {s[0]: s for s in helpers.PARITY_SCENES}[name]
    _, kw, overrides, steps, substeps = scene
    gpu, cpu = helpers.product(), helpers.oracle()
    wg, _ = helpers.make_world(gpu, name, kw, overrides)
    wc, _ = helpers.make_world(cpu, name, kw, overrides)
    gpu.lib.b2mi_World_SetLdsBudget(wg, 0)
    gpu.lib.b2mi_World_SetSplitSlots(wg, 1)
    gpu.lib.b2mi_World_SetGridBlocks(wg, grid)
    try:
        for s in range(min(steps, 40)):
            gpu.step(wg, substeps=substeps)
            cpu.step(wc, substeps=substeps)
        m = helpers.first_contact_mismatch(gpu.dump_contacts(wg), cpu.dump_contacts(wc))
        assert m is None, f"{name}: contacts differ: {m}"
        m = helpers.first_body_mismatch(gpu.dump_bodies(wg), cpu.dump_bodies(wc))
        assert m is None, f"{name}: bodies differ: {m}"
        items = gpu.work_items(world=wg)
        modes = set(items[items[:, 0] > 0, 3].tolist())  # every item with a slot runs over the grid
        assert not modes or modes == ({1} if grid < 0 else {2}), modes
    finally:
        gpu.destroy(wg)
        cpu.destroy(wc)


@pytest.mark.parametrize("name", ["overflow", "overflow_negzero", "tumbler"])
@pytest.mark.parametrize("grid", [-1, 0], ids=["per_phase", "grid_auto"])
@pytest.mark.parametrize("run", [0, 1, 4], ids=["levels_over_grid", "run1", "run4"])
def test_step_parity_overflow_level_runs(name, grid, run):
    """The overflow levels of an item run over the grid: each level over the whole grid (a run
    limit of 0), or the levels of at most one or four slots in one workgroup between grid barriers
    (or in one launch) and the larger ones over the grid; every split gives the oracle's bits."""
    scene = {s[0]: s for s in helpers.PARITY_SCENES}[name]
    _, kw, overrides, steps, substeps = scene
    gpu, cpu = helpers.product(), helpers.oracle()
    wg, _ = helpers.make_world(gpu, name, kw, overrides)
    wc, _ = helpers.make_world(cpu, name, kw, overrides)
    gpu.lib.b2mi_World_SetLdsBudget(wg, 0)
    gpu.lib.b2mi_World_SetSplitSlots(wg, 1)
    gpu.lib.b2mi_World_SetGridBlocks(wg, grid)
    gpu.lib.b2mi_World_SetLevelRunSlots(wg, run)
    try:
        for s in range(min(steps, 30)):
            gpu.step(wg, substeps=substeps)
            cpu.step(wc, substeps=substeps)
            m = helpers.first_body_mismatch(gpu.dump_bodies(wg), cpu.dump_bodies(wc))
            assert m is None, f"{name} step {s}: bodies differ: {m}"
        m = helpers.first_contact_mismatch(gpu.dump_contacts(wg), cpu.dump_contacts(wc))
        assert m is None, f"{name}: contacts differ: {m}"
    finally:
        gpu.destroy(wg)
        cpu.destroy(wc)


def _batch_worlds(gpu, cpu, n, seed):
    batch = gpu.create_batch()
    gpu.use_batch(batch)
    worlds_g, worlds_c = [], []
    try:
        for i in range(n):
            worlds_g.append(gpu.scene("batch_pyramid", seed=seed, index=i)[0])
            worlds_c.append(cpu.scene("batch_pyramid", seed=seed, index=i)[0])
    finally:
        gpu.use_batch(None)
    return batch, worlds_g, worlds_c


def _compare_worlds(gpu, cpu, worlds_g, worlds_c, label):
    for i, (wg, wc) in enumerate(zip(worlds_g, worlds_c)):
        m = helpers.first_contact_mismatch(gpu.dump_contacts(wg), cpu.dump_contacts(wc))
        assert m is None, f"{label} world {i}: {m}"
        m = helpers.first_body_mismatch(gpu.dump_bodies(wg), cpu.dump_bodies(wc))
        assert m is None, f"{label} world {i}: {m}"


def test_batch_large_lds_budget_parity():
    """Items up to 160 KiB of LDS (the kernel's dynamic-LDS attribute above 64 KiB)."""
    gpu, cpu = helpers.product(), helpers.oracle()
    batch, worlds_g, worlds_c = _batch_worlds(gpu, cpu, 12, 21)
    try:
        gpu.lib.b2mi_Batch_SetLdsBudget(batch, 160 * 1024)
        for s in range(30):
            gpu.batch_step(batch)
            for wc in worlds_c:
                cpu.step(wc)
        _compare_worlds(gpu, cpu, worlds_g, worlds_c, "lds160k")
    finally:
        gpu.lib.b2mi_DestroyBatch(batch)
        for wc in worlds_c:
            cpu.destroy(wc)


def test_batch_mixed_stores_parity():
    """A 96 KiB budget puts the largest worlds of this batch in the HBM store and the next ones in
    the bodies-in-LDS store, launched side by side every step: each item's constraint records
    must stay inside its own region of the shared record array."""
    gpu, cpu = helpers.product(), helpers.oracle()
    batch, worlds_g, worlds_c = _batch_worlds(gpu, cpu, 10, 5)
    try:
        gpu.lib.b2mi_Batch_SetLdsBudget(batch, 96 * 1024)
        gpu.batch_step(batch)
        n = gpu.lib.b2mi_Batch_GetSolverLaunches(batch, None, 0)
        launches = np.zeros((n, 4), np.int32)
        gpu.lib.b2mi_Batch_GetSolverLaunches(batch, launches.ctypes.data, n)
        assert 0 in launches[:, 3] and (launches[:, 3] > 0).any(), launches
        for s in range(30):
            gpu.batch_step(batch)
        for wc in worlds_c:
            cpu.step(wc, n=31)
        _compare_worlds(gpu, cpu, worlds_g, worlds_c, "mixed stores")
    finally:
        gpu.lib.b2mi_DestroyBatch(batch)
        for wc in worlds_c:
            cpu.destroy(wc)


def test_batch_autotune_keeps_parity():
    """The autotuner's trial steps are ordinary steps: after tuning (60 candidate shapes and up to
    four footprint limits, one rebuild step and one measured step each) and more steps under the
    chosen shape, every world still matches the oracle bit for bit."""
    gpu, cpu = helpers.product(), helpers.oracle()
    batch, worlds_g, worlds_c = _batch_worlds(gpu, cpu, 10, 5)
    try:
        gpu.batch_step(batch, n=3)
        best = gpu.lib.b2mi_Batch_Autotune(batch, 1.0 / 60.0, 4, 1)
        assert best > 0.0
        gpu.batch_step(batch, n=5)
        tuned = helpers.autotune_steps(gpu, batch, 1)
        for wc in worlds_c:
            cpu.step(wc, n=3 + tuned + 5)
        _compare_worlds(gpu, cpu, worlds_g, worlds_c, "autotune")
    finally:
        gpu.lib.b2mi_DestroyBatch(batch)
        for wc in worlds_c:
            cpu.destroy(wc)


def test_small_batch_parity():
    """Three uneven worlds in one batch (regions with slack between them in the launch arrays)."""
    gpu, cpu = helpers.product(), helpers.oracle()
    batch, worlds_g, worlds_c = _batch_worlds(gpu, cpu, 3, 17)
    try:
        for s in range(20):
            gpu.batch_step(batch)
            for wc in worlds_c:
                cpu.step(wc)
        _compare_worlds(gpu, cpu, worlds_g, worlds_c, "small batch")
    finally:
        gpu.lib.b2mi_DestroyBatch(batch)
        for wc in worlds_c:
            cpu.destroy(wc)


@pytest.mark.parametrize("chunks", [1, 3])
def test_chunked_batch_parity(chunks):
    """72 worlds, in one chunk or in three whose solver launches follow each chunk's narrowphase:
    a chunk whose contacts change state runs its solver again after the host work while the
    others keep their speculative results. Falling boxes change state at some steps and not at
    others."""
    gpu, cpu = helpers.product(), helpers.oracle()
    d = gpu.default_world_def(enableSleep=False, enableContinous=False)
    dc = cpu.default_world_def(enableSleep=False, enableContinous=False)
    batch = gpu.create_batch()
    gpu.use_batch(batch)
    worlds_g, worlds_c = [], []
    try:
        for i in range(72):
            worlds_g.append(gpu.scene("drop_boxes", d, count=4, seed=9, index=i)[0])
            worlds_c.append(cpu.scene("drop_boxes", dc, count=4, seed=9, index=i)[0])
    finally:
        gpu.use_batch(None)
    gpu.lib.b2mi_Batch_SetWorldChunks(batch, chunks)
    try:
        for s in range(45):
            gpu.batch_step(batch)
            for wc in worlds_c:
                cpu.step(wc)
        _compare_worlds(gpu, cpu, worlds_g, worlds_c, "chunked batch")
    finally:
        gpu.lib.b2mi_DestroyBatch(batch)
        for wc in worlds_c:
            cpu.destroy(wc)


@pytest.mark.parametrize("target", [128, 1024])
def test_batch_per_item_block_sizes_parity(target):
    """Workgroup size per work item (blockSize 0): items of each size go to their own launch,
    the launches run side by side on separate streams, and every world keeps the oracle's bits."""
    gpu, cpu = helpers.product(), helpers.oracle()
    batch, worlds_g, worlds_c = _batch_worlds(gpu, cpu, 12, 3)
    try:
        gpu.lib.b2mi_Batch_Configure(batch, 0, target)
        for s in range(30):
            gpu.batch_step(batch)
            for wc in worlds_c:
                cpu.step(wc)
        _compare_worlds(gpu, cpu, worlds_g, worlds_c, f"per-item target {target}")
    finally:
        gpu.lib.b2mi_DestroyBatch(batch)
        for wc in worlds_c:
            cpu.destroy(wc)


@pytest.mark.parametrize("block", [64, 512, 1024])
def test_step_parity_block_sizes(block):
    """Any solver workgroup size gives the oracle's bits (strided loops over slots and bodies)."""
    gpu, cpu = helpers.product(), helpers.oracle()
    wg, _ = helpers.make_world(gpu, "large_pyramid", {"base": 20}, {})
    wc, _ = helpers.make_world(cpu, "large_pyramid", {"base": 20}, {})
    gpu.lib.b2mi_World_Configure(wg, block, 0)
    try:
        for s in range(30):
            gpu.step(wg)
            cpu.step(wc)
        assert helpers.first_contact_mismatch(gpu.dump_contacts(wg), cpu.dump_contacts(wc)) is None
        assert helpers.first_body_mismatch(gpu.dump_bodies(wg), cpu.dump_bodies(wc)) is None
    finally:
        gpu.destroy(wg)
        cpu.destroy(wc)


def test_worlds_stepped_from_two_threads():
    """Two worlds stepped at the same time from two host threads (the API allows it for different
    worlds): both use the library's host thread pool, and each must still match the oracle bit
    for bit."""
    import threading

    gpu, cpu = helpers.product(), helpers.oracle()
    cases = [("pyramid_grid", {"rows": 2, "cols": 3, "base": 10, "seed": 3}, {}),
             ("tumbler", {"count": 400, "seed": 5}, {})]
    worlds = [(helpers.make_world(gpu, n, kw, o)[0], helpers.make_world(cpu, n, kw, o)[0]) for n, kw, o in cases]
    errors = []

    def run(wg):
        try:
            for _ in range(30):
                gpu.lib.b2World_Step(wg, ctypes.c_float(1.0 / 60.0), 4)
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    try:
        threads = [threading.Thread(target=run, args=(wg,)) for wg, _ in worlds]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        assert not errors, errors
        for (name, _, _), (wg, wc) in zip(cases, worlds):
            cpu.step(wc, n=30)
            m = helpers.first_contact_mismatch(gpu.dump_contacts(wg), cpu.dump_contacts(wc))
            assert m is None, f"{name}: contacts differ: {m}"
            m = helpers.first_body_mismatch(gpu.dump_bodies(wg), cpu.dump_bodies(wc))
            assert m is None, f"{name}: bodies differ: {m}"
    finally:
        for wg, wc in worlds:
            gpu.destroy(wg)
            cpu.destroy(wc)


def test_overflow_kinematic_negzero_set_after_first_step():
    """The rake (kinematic, touching 30 overflow constraints) is given a -0 velocity component
    after the first step: the overflow levels are rebuilt with it counted as changed, so the
    serial order's signed zeros are kept."""
    gpu, cpu = helpers.product(), helpers.oracle()
    count = 30
    worlds = [helpers.make_world(api, "overflow", {"count": count}, {}) for api in (gpu, cpu)]
    try:
        for api, (w, ids) in zip((gpu, cpu), worlds):
            api.step(w)
            plank = ids[0]
            rake = BodyId(plank.index1 + count + 1, plank.world0, plank.revision)
            assert api.lib.b2Body_GetType(rake) == 1  # kinematic
            api.lib.b2Body_SetLinearVelocity(rake, Vec2(-0.0, 0.0))
        for s in range(30):
            gpu.step(worlds[0][0])
            cpu.step(worlds[1][0])
            m = helpers.first_contact_mismatch(gpu.dump_contacts(worlds[0][0]), cpu.dump_contacts(worlds[1][0]))
            assert m is None, f"step {s}: contacts differ: {m}"
            m = helpers.first_body_mismatch(gpu.dump_bodies(worlds[0][0]), cpu.dump_bodies(worlds[1][0]))
            assert m is None, f"step {s}: bodies differ: {m}"
    finally:
        gpu.destroy(worlds[0][0])
        cpu.destroy(worlds[1][0])


def test_island_sleep_fails_loudly():
    """Island sleep is not part of the device pipeline. With the reference's default world
    definition (sleep on) a resting box's island reaches the sleep time, after which the
    reference would put it to sleep: the step reports that through the assert callback and the
    last error instead of simulating on with results the reference would not give. Before that
    step the run is the oracle's, bit for bit."""
    gpu, cpu = helpers.product(), helpers.oracle()
    seen = []
    ASSERT = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int)

    def record(cond, file, line):
        seen.append(cond.decode())
        return 0  # do not abort

    cb = ASSERT(record)
    gpu.lib.b2SetAssertFcn.argtypes = [ctypes.c_void_p]
    gpu.lib.b2SetAssertFcn(ctypes.cast(cb, ctypes.c_void_p))
    wg, _ = gpu.scene("hello", gpu.default_world_def())
    wc, _ = cpu.scene("hello", cpu.default_world_def(enableSleep=False))
    try:
        first = None
        for s in range(150):
            gpu.step(wg)
            cpu.step(wc)
            if seen:
                first = s
                break
            m = helpers.first_body_mismatch(gpu.dump_bodies(wg), cpu.dump_bodies(wc))
            assert m is None, f"step {s}: {m}"
        assert first is not None and 60 <= first < 150, first
        assert "sleep" in seen[0], seen
        assert b"sleep" in gpu.lib.b2mi_GetLastError()
    finally:
        gpu.lib.b2SetAssertFcn(None)
        gpu.destroy(wg)
        cpu.destroy(wc)


@pytest.mark.parametrize("case", ["pyramid", "per_phase", "grid", "batch"])
def test_launch_graph_replays_keep_parity(case):
    """Once a step enqueues the same solver launches as the step before, they are captured as a
    HIP graph and replayed, with the step's own launches, copies and host work between replays:
    every step still matches the oracle bit for bit, and the same run with graphs off gives the
    same bits. The per-phase launches of a very large item are replayed from a graph too; a step
    whose solver is one plain launch (the pyramid here) is not captured."""
    gpu, cpu = helpers.product(), helpers.oracle()
    if case == "batch":
        batch, worlds_g, worlds_c = _batch_worlds(gpu, cpu, 10, 5)
        # two stores (as in the mixed-store test), so two launches side by side every step
        gpu.lib.b2mi_Batch_SetLdsBudget(batch, 96 * 1024)
        try:
            for s in range(40):
                gpu.batch_step(batch)
                for wc in worlds_c:
                    cpu.step(wc)
                if s % 10 == 9:
                    _compare_worlds(gpu, cpu, worlds_g, worlds_c, f"graph batch step {s}")
            captures, replays = gpu.graph_stats(batch=batch)
            assert replays > 0, (captures, replays)
        finally:
            gpu.lib.b2mi_DestroyBatch(batch)
            for wc in worlds_c:
                cpu.destroy(wc)
        return
    kw = {"base": 20}
    runs = []
    for graphs in (1, 0):
        wg, _ = helpers.make_world(gpu, "large_pyramid", kw, {})
        gpu.lib.b2mi_World_SetLaunchGraphs(wg, graphs)
        if case != "pyramid":
            gpu.lib.b2mi_World_SetLdsBudget(wg, 0)
            gpu.lib.b2mi_World_SetSplitSlots(wg, 1)
            gpu.lib.b2mi_World_SetGridBlocks(wg, -1 if case == "per_phase" else 2)
        wc = helpers.make_world(cpu, "large_pyramid", kw, {})[0] if graphs else None
        try:
            for s in range(60):
                gpu.step(wg)
                if wc is not None:
                    cpu.step(wc)
                    m = helpers.first_body_mismatch(gpu.dump_bodies(wg), cpu.dump_bodies(wc))
                    assert m is None, f"{case} step {s}: {m}"
            runs.append((helpers.masked(gpu.dump_contacts(wg)), gpu.dump_bodies(wg), gpu.graph_stats(world=wg)))
            if wc is not None:
                m = helpers.first_contact_mismatch(gpu.dump_contacts(wg), cpu.dump_contacts(wc))
                assert m is None, f"{case}: {m}"
        finally:
            gpu.destroy(wg)
            if wc is not None:
                cpu.destroy(wc)
    assert np.array_equal(runs[0][0].view(np.uint8), runs[1][0].view(np.uint8))
    assert np.array_equal(runs[0][1].view(np.uint8), runs[1][1].view(np.uint8))
    print(case, "graph captures/replays with graphs on:", runs[0][2], "off:", runs[1][2])
    assert runs[1][2] == (0, 0)
    if case == "pyramid":
        assert runs[0][2] == (0, 0), runs[0][2]
    elif case == "per_phase":
        assert runs[0][2][1] > 0, runs[0][2]


@pytest.mark.parametrize("name,kw,overrides,steps,substeps", helpers.PARITY_SCENES,
                         ids=[s[0] for s in helpers.PARITY_SCENES])
def test_step_parity_device_pairs(name, kw, overrides, steps, substeps):
    """Broadphase pair queries on the device for every step (threshold of zero moved proxies):
    the new pairs, and so the contact ids, colours and everything after them, are the oracle's."""
    gpu, cpu = helpers.product(), helpers.oracle()
    wg, _ = helpers.make_world(gpu, name, kw, overrides)
    wc, _ = helpers.make_world(cpu, name, kw, overrides)
    gpu.lib.b2mi_World_SetDevicePairs(wg, 0)
    try:
        for s in range(steps):
            gpu.step(wg, substeps=substeps)
            cpu.step(wc, substeps=substeps)
            if s % 10 == 9 or s < 3:
                m = helpers.first_contact_mismatch(gpu.dump_contacts(wg), cpu.dump_contacts(wc))
                assert m is None, f"{name} step {s}: contacts differ: {m}"
        m = helpers.first_body_mismatch(gpu.dump_bodies(wg), cpu.dump_bodies(wc))
        assert m is None, f"{name}: bodies differ: {m}"
        stats = gpu.device_pair_stats(wg)
        assert stats["updates"] > 0, stats
        if name == "tumbler":
            # the container walls' queries run past one thread's visit budget
            assert stats["spilled"] > 0, stats
    finally:
        gpu.destroy(wg)
        cpu.destroy(wc)
