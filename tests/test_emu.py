"""The step pipeline, kernels included, on the CPU emulation of the HIP runtime
(tests/emu/): the same host driver and kernel sources compiled as host C++ with
AddressSanitizer, so that an out-of-bounds access of any kernel or host path is reported on a
machine without a GPU. A subset of the GPU tests runs against it in a child process (the
sanitizer runtime must be loaded before the interpreter's libraries) and must match the oracle
bit for bit, as on the GPU."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EMU = os.path.join(HERE, "emu")
LIB = os.path.join(EMU, "_build", "libb2mi.so")

# small cases that cover the taped run and its adjoint, the batch, the storage variants, the
# per-item launch classes, the steps whose solver launch goes right behind the narrowphase, the
# grid-wide item modes with their overflow level runs, the joints and continuous collision
CASES = [
    "tests/test_gpu_adjoint.py::test_taped_run_matches_oracle",
    "tests/test_gpu_adjoint.py::test_replay_reproduces_run",
    "tests/test_gpu_adjoint.py::test_autograd_function",
    "tests/test_gpu_adjoint.py::test_adjoint_state_in_lds_or_hbm_bit_identical",
    "tests/test_gpu_parity.py::test_step_parity_with_oracle[hello]",
    "tests/test_gpu_parity.py::test_step_parity_with_oracle[shape_zoo]",
    "tests/test_gpu_parity.py::test_step_parity_with_oracle[polygon_zoo]",
    "tests/test_gpu_parity.py::test_step_parity_with_oracle[batch_pyramid]",
    "tests/test_gpu_parity.py::test_small_batch_parity",
    "tests/test_gpu_parity.py::test_chunked_batch_parity[3]",
    "tests/test_gpu_parity.py::test_gpu_run_to_run_bit_reproducible",
    "tests/test_gpu_parity.py::test_step_parity_solver_storage[lds8k-tilted_stacks]",
    "tests/test_gpu_parity.py::test_step_parity_solver_storage[hbm-tumbler]",
    "tests/test_gpu_parity.py::test_step_parity_with_oracle[overflow]",
    "tests/test_gpu_parity.py::test_step_parity_with_oracle[overflow_negzero]",
    "tests/test_gpu_parity.py::test_step_parity_split_items[per_phase-tumbler]",
    "tests/test_gpu_parity.py::test_step_parity_split_items[grid3-tumbler]",
    "tests/test_gpu_parity.py::test_step_parity_split_items[grid_auto-overflow]",
    "tests/test_gpu_parity.py::test_step_parity_split_items[grid3-overflow_negzero]",
    "tests/test_gpu_parity.py::test_worlds_stepped_from_two_threads",
    "tests/test_gpu_parity.py::test_overflow_kinematic_negzero_set_after_first_step",
    "tests/test_gpu_parity.py::test_step_parity_device_pairs[tumbler]",
    "tests/test_gpu_parity.py::test_step_parity_device_pairs[shape_zoo]",
    "tests/test_events.py::test_events_match_oracle[drop_boxes]",
    "tests/test_events.py::test_events_taken_before_an_edit",
    "tests/test_gpu_parity.py::test_step_parity_overflow_level_runs[levels_over_grid-grid_auto-overflow]",
    "tests/test_joints.py::test_joint_parity[zoo16]",
    "tests/test_joints.py::test_joint_parity[grid12g]",
    "tests/test_joints.py::test_joint_parity_stores[bodies_lds]",
    "tests/test_joints.py::test_joint_zoo_island_modes[per_phase]",
    "tests/test_joints.py::test_joint_edits_between_steps",
    "tests/test_continuous.py::test_continuous_parity[bullets10]",
    "tests/test_continuous.py::test_continuous_batch_parity",
]


def _sources_newer_than(path):
    if not os.path.exists(path):
        return True
    t = os.path.getmtime(path)
    dirs = [os.path.join(ROOT, "box2c_amd", "csrc"), EMU, os.path.join(EMU, "hip"), os.path.join(ROOT, "include"),
            os.path.join(ROOT, "include", "box2d")]
    for d in dirs:
        for f in os.listdir(d):
            if f.endswith((".cpp", ".hip", ".hpp", ".h", ".c")) and os.path.getmtime(os.path.join(d, f)) > t:
                return True
    return False


def _asan_runtime():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_pipeline_on_cpu_emulation():
    asan = _asan_runtime()
    if asan is None:
        pytest.skip("no AddressSanitizer runtime")
    if _sources_newer_than(LIB):
        subprocess.run([sys.executable, os.path.join(EMU, "build_emu.py")], check=True, capture_output=True)
    env = dict(os.environ)
    env["B2MI_TEST_LIBRARY"] = "emu"
    env["ASAN_OPTIONS"] = "detect_leaks=0:verify_asan_link_order=0"
    env["LD_PRELOAD"] = asan + (":" + env["LD_PRELOAD"] if env.get("LD_PRELOAD") else "")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *CASES], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=1200)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
