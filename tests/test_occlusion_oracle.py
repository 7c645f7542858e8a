"""Pins the oracle's any-hit (trace_shadow) restatement: per object it is
exactly "the closest-hit walk found a hit", over objects an OR, and it gives
the known answers of simple configurations."""
import numpy as np

import oracle
import rayengine_amd as re
from rayengine_amd import workloads as W


def test_single_object_equals_closest_hit_found():
    tri_v, tree, o, d = W.small_scene(96, 96)
    nodes, leaves, tris = tree.export()
    obj = tree.identity_object()
    for secondary in (True, False):
        for two_sided in (False, True):
            occ = oracle.occluded(nodes, leaves, tris, obj, o, d, two_sided=two_sided, secondary=secondary)
            p, _ = oracle.trace(nodes, leaves, tris, obj, o, d, two_sided=two_sided, secondary=secondary)
            assert np.array_equal(occ, (p >= 0).astype(np.uint8))
    p, t = oracle.trace(nodes, leaves, tris, obj, o, d)
    so, sd = W.shadow_rays(tri_v, o, d, p, t, (3.0, 4.0, 5.0))
    occ = oracle.occluded(nodes, leaves, tris, obj, so, sd)
    ps, _ = oracle.trace(nodes, leaves, tris, obj, so, sd, secondary=True)
    assert np.array_equal(occ, (ps >= 0).astype(np.uint8))
    assert 0 < occ.sum() < len(occ)


def test_multi_object_is_the_or_of_objects():
    ctx_free = [re.KDTree(re.icosphere(3, 1.0)), re.KDTree(re.random_tris(5, 2000, -2, 2, 0.05, 0.3))]
    n1, l1, t1 = ctx_free[0].export()
    n2, l2, t2 = ctx_free[1].export()
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
    objs = np.concatenate([ctx_free[0].identity_object(), ctx_free[1].identity_object()])
    objs[1]["root"] = nb
    objs[0]["loc"] = (0.6, 0.0, 0.0)
    objs[1]["is_character"] = 1
    objs[1]["loc"] = (-0.3, 0.2, 0.1)
    rot = np.float32(0.4)
    objs[1]["sin_y"], objs[1]["cos_y"] = np.float32(np.sin(np.float64(rot))), np.float32(np.cos(np.float64(rot)))
    o, d = re.camera_rays(W.SMALL_CAMERA, 64, 64)
    occ = oracle.occluded(nodes, leaves, tris, objs, o, d, secondary=True)
    per = [oracle.trace(nodes, leaves, tris, objs[k:k + 1], o, d, secondary=True)[0] >= 0 for k in range(2)]
    assert np.array_equal(occ, (per[0] | per[1]).astype(np.uint8))
    assert per[0].sum() > 0 and per[1].sum() > 0


def test_known_answers():
    tree = re.KDTree(re.icosphere(3, 1.0))
    nodes, leaves, tris = tree.export()
    obj = tree.identity_object()
    o = np.float32([[0, 0, 5], [0, 0, 5], [0.01, 0.02, 0.03], [0.05, -0.02, 0.01]])
    d = np.float32([[0, 0, -1], [0, 0, 1], [0.3, 0.5, 0.8], [0.1, -0.9, 0.2]])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # towards the sphere: hit; away from it: none; from inside (off the split
    # planes) the triangles are back faces: none one-sided, a hit two-sided
    assert list(oracle.occluded(nodes, leaves, tris, obj, o, d)) == [1, 0, 0, 0]
    assert list(oracle.occluded(nodes, leaves, tris, obj, o, d, two_sided=True)) == [1, 0, 1, 1]
