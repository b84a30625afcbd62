"""b2mi_Batch_ReplayTapeNoAlloc, the replay of a taped run that allocates nothing: called through
the C ABI, it must give the bits of b2mi_Batch_ReplayTape on the same tape, with and without
per-world parameter overrides, on two calls in a row (the second on a smaller taped run, so the
scratch array holds the larger run's contents when the smaller one starts)."""
import numpy as np
import pytest

from box2c_amd import diff
from tests import helpers

pytestmark = pytest.mark.gpu


def _no_alloc(r, start, params=None):
    s = np.ascontiguousarray(start, np.float32).reshape(-1, diff.POSE)
    out = np.full_like(s, np.nan)
    p = None if params is None else np.ascontiguousarray(params, np.float32).reshape(-1, 2)
    ok = r.api.lib.b2mi_Batch_ReplayTapeNoAlloc(r.batch, s.ctypes.data, None if p is None else p.ctypes.data, out.ctypes.data)
    assert ok == 1
    return out.reshape(r.worlds, r.boxes, diff.POSE)


def test_replay_no_alloc_matches_replay():
    gpu = helpers.product()
    for worlds, boxes, steps, seed in ((12, 6, 30, 3), (3, 4, 12, 8)):
        r = diff.DropBoxesRollout(gpu, worlds, boxes, steps, 4, 5, seed=seed)
        try:
            f, e, v0, w0 = r.default_params()
            final = r.run(f, e, v0, w0)
            want = r.replay(r.start)
            got = _no_alloc(r, r.start)
            assert np.array_equal(want.view(np.uint32), got.view(np.uint32)), (worlds, "plain")
            assert np.array_equal(final.view(np.uint32), got.view(np.uint32)), (worlds, "run")
            params = np.stack([f * 0.9, np.minimum(e + 0.1, 1.0)], axis=1).astype(np.float32)
            want = r.replay(r.start, params)
            got = _no_alloc(r, r.start, params)
            assert np.array_equal(want.view(np.uint32), got.view(np.uint32)), (worlds, "params")
            # perturbed start poses: a different input through the same scratch
            start = r.start.copy()
            start[:, :, 4] += 0.25
            want = r.replay(start, params)
            got = _no_alloc(r, start, params)
            assert np.array_equal(want.view(np.uint32), got.view(np.uint32)), (worlds, "start")
        finally:
            r.close()
