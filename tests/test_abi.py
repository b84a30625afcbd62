"""The C-ABI boundary: libb2mi.so loads, exports every function include/*.h declares, builds
worlds on the host exactly as the oracle does, and refuses to step without a GPU."""
from __future__ import annotations

import os
import re
import subprocess
import sys

import numpy as np
import pytest

from tests import helpers

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "b2mi.h")] + [
    os.path.join(ROOT, "include", "box2d", f) for f in sorted(os.listdir(os.path.join(ROOT, "include", "box2d")))]


def declared_functions() -> list[str]:
    names = []
    for h in HEADERS:
        text = open(h).read()
        for m in re.finditer(r"^B2_API\s+[^;{]*?\b(\w+)\s*\(", text, re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_headers_declare_the_api():
    names = declared_functions()
    for must in ("b2CreateWorld", "b2World_Step", "b2CreateBody", "b2CreatePolygonShape", "b2CollidePolygons",
                 "b2mi_Batch_Step", "b2mi_World_DumpContacts"):
        assert must in names
    assert len(names) > 100


def test_library_exports_every_declared_symbol():
    import box2c_amd
    lib = box2c_amd.load().lib
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_exports_nothing_undeclared():
    """Only the declared C API is visible (the library builds with -fvisibility=hidden)."""
    import box2c_amd
    out = subprocess.run(["nm", "-D", "--defined-only", box2c_amd.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.split()[-2] in ("T", "W")}
    extra = sorted(n for n in exported - set(declared_functions()) if not n.startswith("_"))
    assert not extra, f"undeclared exports: {extra}"


@pytest.mark.parametrize("name,kw,overrides,steps,substeps", helpers.PARITY_SCENES,
                         ids=[s[0] for s in helpers.PARITY_SCENES])
def test_host_world_setup_matches_oracle(name, kw, overrides, steps, substeps):
    """Before the first step (no GPU needed), body records, masses and inertias agree bit for
    bit with the oracle."""
    gpu, cpu = helpers.product(), helpers.oracle()
    wg, ig = helpers.make_world(gpu, name, kw, overrides)
    wc, ic = helpers.make_world(cpu, name, kw, overrides)
    try:
        assert helpers.first_body_mismatch(gpu.dump_bodies(wg), cpu.dump_bodies(wc)) is None
        mg = np.array([[gpu.lib.b2Body_GetMass(b), gpu.lib.b2Body_GetInertiaTensor(b)] for b in ig], np.float32)
        mc = np.array([[cpu.lib.b2Body_GetMass(b), cpu.lib.b2Body_GetInertiaTensor(b)] for b in ic], np.float32)
        assert np.array_equal(mg.view(np.uint32), mc.view(np.uint32))
        cg, cc = gpu.counters(wg), cpu.counters(wc)
        assert (cg.bodyCount, cg.shapeCount, cg.staticBodyCount) == (cc.bodyCount, cc.shapeCount, cc.staticBodyCount)
    finally:
        gpu.destroy(wg)
        cpu.destroy(wc)


def test_step_without_gpu_fails_loudly():
    """The product has no CPU fallback: stepping with no HIP device aborts with a message."""
    import box2c_amd
    if box2c_amd.load().lib.b2mi_DeviceCount() > 0:
        pytest.skip("a GPU is visible")
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import box2c_amd\n"
            "api = box2c_amd.load()\n"
            "w, _ = api.scene('hello', api.default_world_def())\n"
            "api.lib.b2World_Step(w, 1.0 / 60.0, 4)\n"
            "print('stepped')\n") % ROOT
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "stepped" not in r.stdout
    assert "no HIP device" in (r.stdout + r.stderr)
