"""The device kd-tree build gives the host build's tree, byte for byte:
node entries, leaf records (boxes, ropes, triangle runs) and triangle
records, over the small and cluster scenes, degenerate soups and every build
parameter."""
import sys
import time

import numpy as np
import pytest

import rayengine_amd as re
from rayengine_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _same_tree(host, dev, label):
    hi, di = host.info, dev.info
    for k in ("n_tris", "n_nodes", "n_leaves", "n_trirecs", "depth", "max_leaf"):
        assert hi[k] == di[k], f"{label}: info {k} {hi[k]} != {di[k]}"
    assert np.array_equal(hi["root_min"], di["root_min"]) and np.array_equal(hi["root_max"], di["root_max"])
    for name, a, b in zip(("nodes", "leaves", "tris"), host.export(), dev.export()):
        ab, bb = a.view(np.uint8).reshape(len(a), -1), b.view(np.uint8).reshape(len(b), -1)
        bad = np.flatnonzero((ab != bb).any(1))
        assert bad.size == 0, f"{label}: {name} differ at {bad.size} records, first {bad[:5]}: {a[bad[:2]]} vs {b[bad[:2]]}"


def test_small_scene_identical(gpu_ctx):
    tri_v = W.small_soup()
    host = re.KDTree(tri_v)
    dev = re.KDTree(tri_v, ctx=gpu_ctx)
    say(f"small scene: {host.info['n_nodes']} entries, device build {dev.build_ms:.3f} ms")
    _same_tree(host, dev, "small")


@pytest.mark.parametrize("params", [
    dict(clip=False), dict(max_leaf_tris=1), dict(max_leaf_tris=4), dict(max_depth=6), dict(max_depth=40),
    dict(n_bins=16), dict(n_bins=64), dict(n_bins=3), dict(empty_bonus=1.0), dict(cost_intersect=0.2),
])
def test_parameters_identical(gpu_ctx, params):
    tri_v = W.small_soup()
    say(f"params {params}")
    _same_tree(re.KDTree(tri_v, **params), re.KDTree(tri_v, ctx=gpu_ctx, **params), str(params))


def test_degenerate_and_tiny_soups(gpu_ctx):
    rng = np.random.default_rng(5)
    soups = {
        "one": re.random_tris(1, 1),
        "two": re.random_tris(2, 2),
        "nine": re.random_tris(3, 9),
        "seventy": re.random_tris(4, 70),
        "flat_z": np.concatenate([rng.uniform(-1, 1, (500, 2)), np.zeros((500, 1))] * 3, 1).astype(np.float32),
        "duplicates": np.repeat(re.random_tris(6, 40), 5, axis=0),
        "points": np.tile(rng.uniform(-1, 1, (300, 3)), (1, 3)).astype(np.float32),
        "axis_slivers": np.tile(np.float32([0, 0, 0, 1, 0, 0, 2, 0, 0]), (200, 1)) + rng.uniform(0, 1, (200, 1)).astype(np.float32),
        "signed_zeros": np.tile(np.float32([-0.0, 0.0, 1.0, 0.0, -0.0, 1.0, 1.0, 1.0, -0.0]), (100, 1)),
        "wide_range": (re.random_tris(7, 2000) * np.float32(1e4)),
    }
    for name, tri_v in soups.items():
        say(f"soup {name}: {len(tri_v)} tris")
        _same_tree(re.KDTree(tri_v), re.KDTree(tri_v, ctx=gpu_ctx), name)
    empty = re.KDTree(np.zeros((0, 9), np.float32), ctx=gpu_ctx)
    assert empty.info["n_leaves"] == 1 and empty.info["n_trirecs"] == 0


def test_cluster_scene_identical_and_timed(gpu_ctx, cluster):
    tri_v, host = cluster
    t0 = time.perf_counter()
    dev = re.KDTree(tri_v, ctx=gpu_ctx)
    wall = time.perf_counter() - t0
    say(f"cluster: {host.info['n_nodes']} entries, {host.info['n_trirecs']} tri records; device build "
        f"{dev.build_ms:.1f} ms (call {wall * 1e3:.0f} ms)")
    _same_tree(host, dev, "cluster")


def test_bad_parameters(gpu_ctx):
    with pytest.raises(re.ReError):
        re.KDTree(W.small_soup(), n_bins=100, ctx=gpu_ctx)
    with pytest.raises(re.ReError):
        re.KDTree(np.full((3, 9), np.nan, np.float32), ctx=gpu_ctx)
