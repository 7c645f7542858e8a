"""Builds librayengine.so in-tree for gfx950.

Host C++ (scene generation, kd-tree build) is compiled with g++, the C-ABI
layer and the kernels with hipcc --offload-arch=gfx950.  Every object is
compiled with -ffp-contract=off and without fast-math so that the host and
the device evaluate the shared ray-query routines with identical IEEE
rounding.  The result is linked by hipcc into rayengine_amd/librayengine.so.

--debug-bounds builds librayengine_debug.so instead: the HIP sources are
compiled with -DRE_DEBUG_BOUNDS (bounds checks on every node entry, leaf
record and triangle run index the trace kernels read), the host objects are
shared with the normal build.  RE_LIBRARY=debug makes rayengine_amd load it.

Usage: python -m rayengine_amd.build [-j N] [--verbose] [--debug-bounds]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "librayengine.so"
OBJ = PKG / "_build"
OUT_DEBUG = PKG / "librayengine_debug.so"
OBJ_DEBUG = PKG / "_build_debug"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))

HOST_SOURCES = sorted((CSRC / "host").glob("*.cpp"))
HIP_SOURCES = sorted(CSRC.glob("capi*.cpp")) + sorted((CSRC / "kernels").glob("*.hip"))

COMMON = ["-O3", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-std=c++17", "-Wall",
          "-Wno-unused-function", f"-I{PKG.parent / 'include'}"]
HIP_FLAGS = ["--offload-arch=gfx950", "-x", "hip", "-fno-gpu-rdc",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-fhip-fp32-correctly-rounded-divide-sqrt", "-munsafe-fp-atomics"]


def _hipcc() -> str:
    p = ROCM / "bin" / "hipcc"
    return str(p) if p.exists() else (shutil.which("hipcc") or "hipcc")


def _compile(src: Path, verbose: bool, debug: bool = False) -> Path:
    rel = src.relative_to(CSRC)
    hip = src.suffix == ".hip" or src in HIP_SOURCES
    obj = (OBJ_DEBUG if debug and hip else OBJ) / (str(rel).replace(os.sep, "_") + ".o")
    if hip:
        cmd = [_hipcc(), *HIP_FLAGS, *COMMON, *(["-DRE_DEBUG_BOUNDS"] if debug else []), "-c", str(src), "-o", str(obj)]
    else:
        cmd = ["g++", *COMMON, "-fopenmp", f"-I{ROCM / 'include'}", "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    for line in r.stderr.splitlines():
        if "warning" in line and "packed-fp32-ops" not in line:
            print(line, file=sys.stderr)
    return obj


def build(jobs: int | None = None, verbose: bool = False, debug: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    OBJ_DEBUG.mkdir(exist_ok=True)
    out = OUT_DEBUG if debug else OUT
    srcs = HOST_SOURCES + HIP_SOURCES
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose, debug), srcs))
    tmp = out.with_suffix(".so.tmp")
    cmd = [_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
           "-fopenmp", "-Wl,--no-undefined", "-lamdhip64"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--debug-bounds", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.verbose, a.debug_bounds))


if __name__ == "__main__":
    main()
