"""A walk whose state (cur, t_entry, t_exit) repeats within two visits is
periodic and accepts nothing more; the library stops it there instead of
walking on to the visit cap.  These tests pin that shortcut on the CPU: the
oracle's walk with and without it gives identical results, every stopped ray
is one the full walk takes to the cap, and the shortened walk never does more
work."""
import numpy as np
import pytest

import oracle
import rayengine_amd as re
from rayengine_amd import workloads as W


def _check(nodes, leaves, tris, obj, o, d, cap, **kw):
    p0, t0, c0 = oracle.trace(nodes, leaves, tris, obj, o, d, cap=cap, counts=True, **kw)
    p1, t1, c1 = oracle.trace(nodes, leaves, tris, obj, o, d, cap=cap, counts=True, cut=True, **kw)
    assert np.array_equal(p0, p1)
    assert np.array_equal(t0.view(np.uint32), t1.view(np.uint32))
    assert (c1[:, :3] <= c0[:, :3]).all()
    if len(obj) > 1:  # the end reason is the last object's: per-walk checks need one object
        return int((c1[:, :3] < c0[:, :3]).any(1).sum())
    stopped = c1[:, 3] == 6
    assert (c0[stopped, 3] == 1).all(), "a stopped walk must reach the cap when not stopped"
    assert (c1[~stopped] == c0[~stopped]).all(), "walks that do not repeat are unchanged"
    assert (c1[stopped, :3] <= c0[stopped, :3]).all()
    return int(stopped.sum())


@pytest.mark.parametrize("cap", [50, 4096])
@pytest.mark.parametrize("two_sided", [False, True])
def test_small_scene_cut_equals_full_walk(small, cap, two_sided):
    tri_v, tree, o, d = small
    n_stopped = _check(*tree.export(), tree.identity_object(), o, d, cap, two_sided=two_sided)
    assert n_stopped > 0


def test_multi_object_cut_equals_full_walk():
    trees = [re.KDTree(re.icosphere(3, 1.0)), re.KDTree(re.random_tris(5, 2000, -2, 2, 0.05, 0.3))]
    n1, l1, t1 = trees[0].export()
    n2, l2, t2 = trees[1].export()
    # concatenate as the scene upload does (entry, leaf and triangle offsets)
    nb, lb, tb = len(n1), len(l1), len(t1)
    n2 = n2.copy()
    leaf = (n2["meta"] >> 30) == 3
    n2["meta"] = np.where(leaf, (3 << 30) | ((n2["meta"] & 0x3FFFFFFF) + lb),
                          (n2["meta"] & 0xC0000000) | ((n2["meta"] & 0x3FFFFFFF) + nb))
    l2 = l2.copy()
    rp = l2["rope"]
    l2["rope"] = np.where(rp == -1, -1, np.where(rp < -1, ((rp & 0x7FFFFFFF) + lb) | -0x80000000, rp + nb))
    l2["first"] += tb
    nodes, leaves, tris = np.concatenate([n1, n2]), np.concatenate([l1, l2]), np.concatenate([t1, t2])
    objs = np.concatenate([trees[0].identity_object(), trees[1].identity_object()])
    objs[1]["root"] = nb
    objs[0]["loc"] = (0.6, 0.0, 0.0)
    objs[1]["is_character"] = 1
    objs[1]["loc"] = (-0.3, 0.2, 0.1)
    o, d = re.camera_rays(W.SMALL_CAMERA, 96, 96)
    for secondary in (False, True):
        assert _check(nodes, leaves, tris, objs, o, d, 4096, secondary=secondary) > 0


@pytest.mark.slow
def test_cluster_workload_cut_equals_full_walk():
    tri_v = W.cluster_soup_default()
    tree = re.KDTree(tri_v)
    nodes, leaves, tris = tree.export()
    obj = tree.identity_object()
    o, d = re.camera_rays(W.CLUSTER_CAMERA, 1920, 1080)
    p, t = oracle.trace(nodes, leaves, tris, obj, o, d)
    ro, rd = W.cluster_rays(tri_v, p, t, order=W.tile_order(1920, 1080, 8))
    assert _check(nodes, leaves, tris, obj, ro, rd, 50) > 1000
