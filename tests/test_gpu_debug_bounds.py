"""The bounds-checked debug library (librayengine_debug.so, built with
RE_DEBUG_BOUNDS): every trace kernel carrying the checks gives the oracle's
results on the small, median-split, multi-object and cluster scenes with no
bounds-check bit set.  It runs in a child process, since a process holds one
of the two libraries; the release library answers RE_ERR_UNSUPPORTED."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

import rayengine_amd as re

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent


def test_release_library_has_no_bounds_word(gpu_ctx):
    with pytest.raises(re.ReError, match="RE_DEBUG_BOUNDS"):
        gpu_ctx.debug_bounds_errors()


@pytest.mark.timeout(300, method="thread")
def test_debug_library_traces_clean():
    assert (ROOT / "rayengine_amd" / "librayengine_debug.so").exists(), \
        "build with python -m rayengine_amd.build --debug-bounds"
    env = dict(os.environ, RE_LIBRARY="debug")
    print("debug-bounds child: start", flush=True)
    r = subprocess.run([sys.executable, "-u", str(ROOT / "tests" / "debug_bounds_child.py")], env=env,
                       capture_output=True, text=True, timeout=280)
    print(r.stdout[-4000:], r.stderr[-4000:], flush=True)
    assert r.returncode == 0, f"child exit {r.returncode}"
    assert r.stdout.rstrip().endswith("debug-bounds ok")
