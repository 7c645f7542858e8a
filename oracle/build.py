"""Builds the CPU oracle (test infrastructure) into oracle/liboracle.so.

g++ -O2 -ffp-contract=off -fno-fast-math -fopenmp: the FP32 paths evaluate
exactly the reference's operation order with IEEE rounding (x86-64 SSE, no
FMA contraction).  Only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline load the result.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "liboracle.so"


def build(verbose: bool = False) -> Path:
    srcs = sorted(HERE.glob("*.cpp"))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
           "-shared", "-Wall", "-o", str(tmp), *map(str, srcs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    tmp.replace(OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
