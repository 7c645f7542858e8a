"""The batched work-queue kernel's loop is wave-uniform in the gfx950 code.

batch_kernel takes work items from a device counter, one item per wave.  Its
first form read the counter with "if (lane == 0) it = atomicAdd(...); it =
__shfl(it, 0)"; the compiler turned that loop into an inner loop that lanes
1-63 re-entered without lane 0, reading a stale item index of 0 forever, and
the launch never finished.  The item index is now taken with every lane in
the atomic and read back with readfirstlane into a scalar register.  This test
compiles the kernel and checks the properties that make the loop terminate:
the counter's atomic result read back with readfirstlane, no cross-lane
shuffle between the atomic and the item record, and the item record addressed
from that scalar index (the index scaled by the 24-byte record size in scalar
arithmetic).  The persistent closest-hit kernel takes its ray chunks from a
counter the same way and is checked likewise.
"""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "rayengine_amd" / "csrc" / "kernels" / "batch.hip"
HIPCC = Path("/opt/rocm/bin/hipcc")


def _kernel_asm(tmp_path, name_part, src=SRC):
    hipcc = str(HIPCC) if HIPCC.exists() else shutil.which("hipcc")
    if hipcc is None:
        pytest.skip("hipcc not found")
    from rayengine_amd.build import COMMON, HIP_FLAGS
    out = tmp_path / (src.stem + ".s")
    cmd = [hipcc, *HIP_FLAGS, *COMMON, "--cuda-device-only", "-S", str(src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    text = out.read_text()
    bodies = {}
    for m in re.finditer(r"^(_Z\w*" + name_part + r"\w*):", text, re.M):
        end = text.index("s_endpgm", m.end())
        bodies[m.group(1)] = text[m.end():end]
    assert bodies, f"{name_part} not found in the device assembly"
    return bodies


def test_batch_kernel_item_loop_is_wave_uniform(tmp_path):
    for name, body in _kernel_asm(tmp_path, "batch_kernel").items():
        # the item record address is formed from a scalar index (BatchItem is 24 bytes)
        m = re.search(r"s_mul_i32\s+s\d+, s\d+, 24", body)
        assert m, f"{name}: item index is not in a scalar register"
        # the counter value is broadcast with readfirstlane right after the atomic
        atom = body.index("global_atomic_add")
        assert atom < m.start(), f"{name}: item record addressed before the counter is read"
        assert "v_readfirstlane_b32" in body[atom:atom + 400], f"{name}: atomic result not read back uniformly"
        # and no cross-lane shuffle stands between the atomic and the item
        # record (where the first version read the index with a shuffle)
        between = body[atom:m.start()]
        assert "ds_bpermute" not in between and "ds_permute" not in between, f"{name}: shuffle of the item index"


@pytest.mark.parametrize("kernel", ["trace_persistent_kernel", "trace_coop_refill_kernel"])
def test_persistent_trace_kernel_chunk_fetch_is_wave_uniform(tmp_path, kernel):
    """The persistent closest-hit kernels (variants 1 and 5) take chunks of
    rays from a device counter with the same every-lane atomic and
    readfirstlane read-back."""
    src = ROOT / "rayengine_amd" / "csrc" / "kernels" / "trace.hip"
    for name, body in _kernel_asm(tmp_path, kernel, src).items():
        atom = body.index("global_atomic_add_x2")
        window = body[atom:atom + 1500]   # the part bounds' scalar arithmetic may be scheduled in between
        assert window.count("v_readfirstlane_b32") >= 2, f"{name}: chunk start not read back uniformly"
        assert "ds_bpermute" not in window, f"{name}: shuffle of the chunk start"
