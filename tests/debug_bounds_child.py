"""Child process of test_gpu_debug_bounds.py: loads the bounds-checked debug
library (RE_LIBRARY=debug is set by the parent before anything is imported),
runs every trace kernel that carries the checks on the small, median-split,
multi-object and 2^20-triangle cluster scenes, and checks each result against
the CPU oracle and the error word against 0; then makes the leaf check fire
with RE_DEBUG_BOUNDS_LEAF_LIMIT (every read stays inside the arrays).  Prints one line per step and
"debug-bounds ok" at the end."""
import os
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402
import torch  # noqa: E402,F401  (one HIP runtime: torch's, loaded first)

import oracle  # noqa: E402
import rayengine_amd as re  # noqa: E402
from rayengine_amd import _lib  # noqa: E402
from rayengine_amd import workloads as W  # noqa: E402


def say(*a):
    print(*a, flush=True)


def same(p, t, po, to, what):
    assert np.array_equal(p, po), f"{what}: {int((p != po).sum())} prim ids differ"
    assert np.array_equal(t.view(np.uint32), to.view(np.uint32)), f"{what}: t differs"


def clean(ctx, what):
    m = ctx.debug_bounds_errors()
    assert m == 0, f"{what}: bounds-check bits {m:#x}"


def scene_cases(ctx, tri_v, tree, o, d, name, variants=(0, 1, 2, 3, 4, 5)):
    scene = ctx.upload_scene([re.SceneObject(tree)])
    nodes, leaves, tris = tree.export()
    obj = tree.identity_object()
    po, to, co = oracle.trace(nodes, leaves, tris, obj, o, d, counts=True, cut=True)
    for v in variants:
        say(f"{name}: closest hit, variant {v}")
        p, t = scene.trace_closest(o, d, variant=v)
        same(p, t, po, to, f"{name} variant {v}")
        clean(ctx, f"{name} variant {v}")
    say(f"{name}: any hit")
    so, sd = W.shadow_rays(tri_v, o, d, po, to, (3.0, 4.0, 5.0))
    got = scene.occluded(so, sd)
    assert np.array_equal(got, oracle.occluded(nodes, leaves, tris, obj, so, sd)), f"{name}: any hit differs"
    clean(ctx, f"{name} any hit")
    say(f"{name}: status")
    p, t, st = scene.trace_status(o, d)
    same(p, t, po, to, f"{name} status")
    assert np.array_equal(st & 0x7F, co[:, 3].astype(np.uint8)), f"{name}: end reasons differ"
    clean(ctx, f"{name} status")
    scene.release()
    return int((po >= 0).sum())


def main():
    assert os.environ.get("RE_LIBRARY") == "debug"
    assert _lib.LIB_PATH.name == "librayengine_debug.so", _lib.LIB_PATH
    ctx = re.Context(0)
    assert ctx.debug_bounds_errors() == 0
    tri_v, tree, o, d = W.small_scene()
    assert scene_cases(ctx, tri_v, tree, o, d, "small") > 8000
    scene_cases(ctx, tri_v, re.KDTree(tri_v, split_rule=1, clip=False), o, d, "median", variants=(0, 4, 5))

    a = re.icosphere(3, 1.0)
    b = re.random_tris(5, 3000, -1.5, 1.5, 0.05, 0.3)
    ta, tb = re.KDTree(a), re.KDTree(b)
    scene = ctx.upload_scene([re.SceneObject(ta, location=(0.5, -0.2, 0.1), rotation_y=0.7),
                              re.SceneObject(tb, location=(-0.3, 0.4, 0.0), rotation_y=-1.3, prim_offset=len(a))])
    nodes, leaves, tris, obs = scene.export()
    o2, d2 = re.camera_rays(re.Camera(pos=(0.3, 0.5, 6.0), view=(0.0, -0.1, -1.0)), 160, 120)
    po, to = oracle.trace(nodes, leaves, tris, obs, o2, d2, secondary=True)
    for v in (0, 4, 5):
        say(f"multi-object: closest hit, variant {v}")
        p, t = scene.trace_closest(o2, d2, secondary=True, variant=v)
        same(p, t, po, to, f"multi-object variant {v}")
        clean(ctx, f"multi-object variant {v}")
    scene.release()

    say("cluster: building the 2^20-triangle soup and its tree")
    cv = W.cluster_soup_default()
    ct = re.KDTree(cv)
    scene = ctx.upload_scene([re.SceneObject(ct)])
    o3, d3 = re.camera_rays(W.CLUSTER_CAMERA, 480, 270)
    p, t = scene.trace_closest(o3, d3)
    o3, d3 = W.cluster_rays(cv, p, t, 480, 270)          # primaries and one bounce per hit
    po, to = oracle.trace(*ct.export(), ct.identity_object(), o3, d3)
    for v in (0, 4, 5):
        say(f"cluster: closest hit, variant {v}")
        p, t = scene.trace_closest(o3, d3, variant=v)
        same(p, t, po, to, f"cluster variant {v}")
        clean(ctx, f"cluster variant {v}")
    assert (po >= 0).sum() > 1000
    scene.release()
    # the check fires: leaf ids checked against half the leaf count are
    # flagged (bit 2) and read as leaf 0, inside the array
    scene = ctx.upload_scene([re.SceneObject(tree)])
    po, to = oracle.trace(*tree.export(), tree.identity_object(), o, d)
    os.environ["RE_DEBUG_BOUNDS_LEAF_LIMIT"] = str(max(1, len(tree.export()[1]) // 2))
    for v in (0, 4, 5):
        say(f"leaf limit: closest hit, variant {v}")
        scene.trace_closest(o, d, variant=v)
        m = ctx.debug_bounds_errors()
        assert m & 2 and not m & 5, f"variant {v}: bits {m:#x}"
    scene.trace_status(o, d)
    assert ctx.debug_bounds_errors() == 2
    del os.environ["RE_DEBUG_BOUNDS_LEAF_LIMIT"]
    p, t = scene.trace_closest(o, d)
    same(p, t, po, to, "after the leaf limit")
    clean(ctx, "after the leaf limit")
    scene.release()
    ctx.close()
    say("debug-bounds ok")


if __name__ == "__main__":
    main()
