"""The batch path (j3dg_render_batch: many frames in one persistent k_trace
launch + k_shade_batch, captured in a graph) run on the CPU before the GPU.

j3d_amd/csrc/j3dg.hip is compiled for x86 against a stand-in HIP runtime
(tests/emu: fibers in lockstep per 64-lane wave, every device allocation
fenced by unmapped memory, a fault handler naming the allocation an access
fell outside of) and driven through the same C ABI by tests/emu/batch_case.py:
13 frames of 5 meshes, five sizes, sub-rectangles and every shading mode,
rendered alone and as a batch (capture, replay, updated frame tables), each
batch frame equal to its single render and to the oracle.

tests/emu/tail_case.py runs frames of one and of three posed objects and ray
queries with each tail mode (the coop_tail tuning) against the oracle.

The two mistakes that faulted the GPU with an illegal memory access are
planted in copies of the source and must fault here too, while the current
source runs clean:
  wrong_table       the batch launch passed the single-frame object table;
  no_frame_handoff  the cooperative tail moved a ray's state but not its frame
                    (index, object table, object count), so a lane completed
                    another frame's hit with the wrong object's records.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu")
sys.path.insert(0, EMU)
import build_emu  # noqa: E402

pytestmark = pytest.mark.skipif(build_emu.compiler() is None, reason="needs clang++")

VARIANTS = [None, "wrong_table", "no_frame_handoff"]


@pytest.fixture(scope="module")
def libs(tmp_path_factory):
    base = tmp_path_factory.mktemp("emu")

    def one(v):
        return v, build_emu.build(str(base / (v or "current")), v)
    with ThreadPoolExecutor(3) as ex:
        return dict(ex.map(one, VARIANTS))


def _run(lib, case="batch_case.py", **env_extra):
    env = dict(os.environ, J3DG_LIBRARY=lib, **env_extra)
    return subprocess.run([sys.executable, os.path.join(EMU, case)], env=env, stdout=subprocess.PIPE,
                          stderr=subprocess.STDOUT, text=True, timeout=600)


@pytest.mark.parametrize("tail", ["1", "4"])
def test_batch_on_the_cpu_equals_single_frames_and_oracle(libs, tail):
    # tail 4 (the default) on a grid of many waves (EMU_CUS): every wave ends in the split tail
    r = _run(libs[None], EMU_TAIL=tail, **({"EMU_CUS": "16"} if tail == "4" else {}))
    print(r.stdout[-3000:])
    assert r.returncode == 0 and "BATCH OK" in r.stdout, r.stdout[-3000:]
    assert "shuffles that read a lane outside the exec mask: 0" in r.stdout


@pytest.mark.parametrize("tail,depth", [("4", "8"), ("4", "32"), ("3", "8"), ("2", "16"), ("1", "8")])
def test_tail_modes_on_the_cpu_match_the_oracle(libs, tail, depth):
    """Frames of one and of three posed objects and closest-/any-hit queries
    with each tail mode: the split tail from 32, 16 and 8 rays left (groups of
    2, 4 and 8 lanes traversing parts of a ray's tree on stacks of their own,
    in columns of finished lanes, regrouped into wider groups as rays finish;
    depth 8 spills to global memory) and the cooperative one, against the
    oracle.  Handing a regrouped ray's new lanes columns that other lanes
    still own (a mistake planted in a copy of the source) faults here in the
    (4, 32) case."""
    r = _run(libs[None], "tail_case.py", EMU_TAIL=tail, EMU_STACK=depth, EMU_CUS="16")
    print(r.stdout[-3000:])
    assert r.returncode == 0 and "TAIL OK" in r.stdout, r.stdout[-3000:]


@pytest.mark.parametrize("variant", ["wrong_table", "no_frame_handoff"])
def test_planted_batch_mistakes_fault_on_the_cpu(libs, variant):
    r = _run(libs[variant])
    print(r.stdout[-3000:])
    assert r.returncode == 86 and "EMU FAULT" in r.stdout and "in kernel k_trace" in r.stdout, r.stdout[-3000:]
