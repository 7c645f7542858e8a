"""Writes tests/golden/small_scene.npz: the small icosphere workload's
triangle soup and rays (SHA-256 of their bytes) and the oracle's closest hits at the
reference cap of 50 leaf visits.

Regenerate only when the workload definition changes:
    python tests/golden/make_golden.py
"""
import hashlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from rayengine_amd import workloads as W  # noqa: E402


def main():
    tri_v, tree, o, d = W.small_scene()
    nodes, leaves, tris = tree.export()
    p, t = oracle.trace(nodes, leaves, tris, tree.identity_object(), o, d, cap=50)
    out = Path(__file__).resolve().parent / "small_scene.npz"
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    np.savez_compressed(out, tri_v_sha=sha(tri_v), ray_o_sha=sha(o), ray_d_sha=sha(d), prim=p,
                        t_bits=t.view(np.uint32))
    print(out, out.stat().st_size, "bytes;", int((p >= 0).sum()), "hits")


if __name__ == "__main__":
    main()
