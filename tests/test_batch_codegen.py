"""The batched work-queue kernel's loop is wave-uniform in the gfx950 code.

batch_kernel takes work items from a device counter, one item per wave.  Its
first form read the counter with "if (lane == 0) it = atomicAdd(...); it =
__shfl(it, 0)"; the compiler turned that loop into an inner loop that lanes
1-63 re-entered without lane 0, reading a stale item index of 0 forever, and
the launch never finished.  The item index is now taken with every lane in
the atomic and read back with readfirstlane into a scalar register.  This test
compiles the kernel and checks the properties that make the loop terminate:
no cross-lane shuffle in the kernel, the item index scaled by the item record
size (24 bytes) in scalar arithmetic, and the loop exit taken on a comparison
of that scalar index with the item count.
"""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "rayengine_amd" / "csrc" / "kernels" / "batch.hip"
HIPCC = Path("/opt/rocm/bin/hipcc")


def _kernel_asm(tmp_path, name_part):
    hipcc = str(HIPCC) if HIPCC.exists() else shutil.which("hipcc")
    if hipcc is None:
        pytest.skip("hipcc not found")
    from rayengine_amd.build import COMMON, HIP_FLAGS
    out = tmp_path / "batch.s"
    cmd = [hipcc, *HIP_FLAGS, *COMMON, "--cuda-device-only", "-S", str(SRC), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    text = out.read_text()
    bodies = {}
    for m in re.finditer(r"^(_Z\w*batch_kernel\w*):", text, re.M):
        end = text.index("s_endpgm", m.end())
        bodies[m.group(1)] = text[m.end():end]
    assert bodies, "batch_kernel not found in the device assembly"
    return bodies


def test_batch_kernel_item_loop_is_wave_uniform(tmp_path):
    for name, body in _kernel_asm(tmp_path, "batch_kernel").items():
        assert "ds_bpermute" not in body and "ds_permute" not in body, f"{name}: cross-lane shuffle in the item loop"
        # the item record address is formed from a scalar index (BatchItem is 24 bytes)
        assert re.search(r"s_mul_i32\s+s\d+, s\d+, 24", body), f"{name}: item index is not in a scalar register"
        # the counter value is broadcast with readfirstlane right after the atomic
        atom = body.index("global_atomic_add")
        assert "v_readfirstlane_b32" in body[atom:atom + 400], f"{name}: atomic result not read back uniformly"
