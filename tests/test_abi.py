"""The C-ABI library loads and exports every symbol include/rayengine.h declares."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "rayengine.h"
LIB = ROOT / "rayengine_amd" / "librayengine.so"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"RE_API\s+[\w\s\*]+?\b(re_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ["re_create", "re_destroy", "re_last_error", "re_kd_build", "re_scene_upload", "re_trace_closest"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    assert LIB.exists(), "build with python -m rayengine_amd.build"
    lib = ctypes.CDLL(str(LIB))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"declared but not exported: {missing}"


def test_debug_library_exports_every_declared_symbol():
    dbg = LIB.with_name("librayengine_debug.so")
    assert dbg.exists(), "build with python -m rayengine_amd.build --debug-bounds"
    lib = ctypes.CDLL(str(dbg))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"declared but not exported: {missing}"


def test_python_binding_covers_header():
    from rayengine_amd import _lib
    missing = [s for s in declared_symbols() if s not in _lib.SIGNATURES]
    assert not missing, f"no ctypes signature for: {missing}"


def test_no_device_call_without_gpu():
    """Without a GPU, re_create fails with a status instead of exiting."""
    import rayengine_amd as rea
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present")
    except Exception:
        pass
    with pytest.raises(rea.ReError):
        rea.Context(0)
