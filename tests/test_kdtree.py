"""Host-side scene generation and rope kd-tree: structural invariants, and the
oracle traversal over the built tree against brute force and closed forms."""
import hashlib
from pathlib import Path

import numpy as np
import pytest

import oracle
import rayengine_amd as re
from rayengine_amd import workloads as W

GOLDEN = Path(__file__).resolve().parent / "golden"


# ------------------------------------------------------------- generators

def test_icosphere_geometry():
    v = re.icosphere(4, 1.0).reshape(-1, 3, 3)
    assert v.shape[0] == 5120
    r = np.linalg.norm(v.reshape(-1, 3), axis=1)
    assert np.allclose(r, 1.0, atol=2e-7)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    assert (np.einsum("ij,ij->i", n, v.sum(1)) > 0).all()      # outward, counter-clockwise
    assert re.icosphere(0).shape == (20, 9)


def test_random_tris_ranges():
    v = re.random_tris(0, 1024, -2, 2, 0.02, 0.2).reshape(-1, 3, 3)
    c = v.mean(1)
    assert (np.abs(c) <= 2.0 + 1e-6).all()
    for e in (v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]):
        L = np.linalg.norm(e, axis=1)
        assert L.min() >= 0.02 * (1 - 1e-5) and L.max() <= 0.2 * (1 + 1e-5)
    assert np.array_equal(v, re.random_tris(0, 1024, -2, 2, 0.02, 0.2).reshape(-1, 3, 3))


def test_cluster_soup_statistics():
    v = re.cluster_soup(1, 1 << 16).reshape(-1, 3, 3)
    L = np.linalg.norm(v[:, 1] - v[:, 0], axis=1)
    assert abs(np.log(L).mean() - (-3.0)) < 0.02 and abs(np.log(L).std() - 0.5) < 0.02
    c = v.mean(1)
    assert (np.abs(c) < 12).all()


def test_camera_rays_reference_convention():
    o, d = re.camera_rays(W.SMALL_CAMERA, 4, 3)
    assert np.allclose(o, [0, 0, 5])
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)
    # basis h = view x camdown = (-1,0,0), v = h x view = (0,-1,0): pixel (0,0)
    # (sx = 0, sy = 1) looks along -h + v = (+x, -y); pixel (3,2) along h - v.
    assert d[0, 0] > 0 and d[0, 1] < 0
    assert d[2 * 4 + 3, 0] < 0 and d[2 * 4 + 3, 1] > 0


# ------------------------------------------------------------- tree structure

def _node_boxes(nodes, root_min, root_max):
    """Box of every node entry, from the root box and the split planes."""
    boxes = {}
    stack = [(0, np.array(root_min, np.float32), np.array(root_max, np.float32))]
    while stack:
        e, lo, hi = stack.pop()
        boxes[e] = (lo, hi)
        ax = int(nodes[e]["meta"] >> 30)
        if ax == 3:
            continue
        s = nodes[e]["split"]
        l_hi = hi.copy(); l_hi[ax] = s
        r_lo = lo.copy(); r_lo[ax] = s
        stack.append((e + 1, lo, l_hi))
        stack.append((int(nodes[e]["meta"] & 0x3FFFFFFF), r_lo, hi))
    return boxes


def test_tree_invariants(small):
    tri_v, tree, _, _ = small
    nodes, leaves, tris = tree.export()
    info = tree.info
    assert info["n_leaves"] == len(leaves) and info["depth"] <= 24
    boxes = _node_boxes(nodes, info["root_min"], info["root_max"])
    leaf_entry = {}
    for e, (lo, hi) in boxes.items():
        if nodes[e]["meta"] >> 30 == 3:
            lid = int(nodes[e]["meta"] & 0x3FFFFFFF)
            leaf_entry[lid] = e
            assert np.array_equal(leaves[lid]["bmin"], lo) and np.array_equal(leaves[lid]["bmax"], hi)
    assert len(leaf_entry) == len(leaves)
    # leaves tile the root box
    vol = lambda lo, hi: float(np.prod(hi.astype(np.float64) - lo.astype(np.float64)))
    total = sum(vol(l["bmin"], l["bmax"]) for l in leaves)
    assert abs(total - vol(info["root_min"], info["root_max"])) < 1e-6 * vol(info["root_min"], info["root_max"])
    # every triangle lands in a leaf, copies overlap their leaf, records hold
    # v0 and the FP32 edges
    seen = np.zeros(len(tri_v), bool)
    tv = tri_v.reshape(-1, 3, 3)
    for lid, l in enumerate(leaves):
        recs = tris[l["first"]:l["first"] + l["count"]]
        if l["count"] > 8:
            assert tree.info["depth"] == 24
        for r in recs:
            p = r["prim"]
            seen[p] = True
            assert np.array_equal(r["v0"], tv[p, 0]) and np.array_equal(r["e1"], tv[p, 1] - tv[p, 0])
            assert (tv[p].min(0) <= l["bmax"]).all() and (tv[p].max(0) >= l["bmin"]).all()
    assert seen.all()


def _check_descent_bounds(nodes, n_leaves):
    """The device descent (device_paths.hpp:descend) follows, from an inner
    entry e, e + 1 (below) or the entry in meta (above).  Both lie after e and
    inside the array, and a leaf entry's id inside the leaf array, so every
    descent ends at a leaf within (tree depth) steps."""
    meta = nodes["meta"].astype(np.int64)
    axis = meta >> 30
    idx = meta & 0x3FFFFFFF
    e = np.arange(len(nodes))
    inner = axis != 3
    n = len(nodes)
    assert (e[inner] + 1 < n).all()
    assert (idx[inner] > e[inner] + 1).all() and (idx[inner] < n).all()
    assert (idx[~inner] < n_leaves).all()
    # each leaf id appears at exactly one entry (the walk from the root reaches it once)
    assert np.array_equal(np.sort(idx[~inner]), np.arange(n_leaves))


def test_descent_indices_increase(small):
    tri_v, tree, _, _ = small
    nodes, leaves, _ = tree.export()
    _check_descent_bounds(nodes, len(leaves))
    soup = re.cluster_soup(3, 40000, 16, -2.0, 2.0, 0.3, -3.0, 0.5)
    for kw in (dict(), dict(max_leaf_tris=1, max_depth=40), dict(max_depth=6)):
        t = re.KDTree(soup, **kw)
        nodes, leaves, _ = t.export()
        _check_descent_bounds(nodes, len(leaves))
        assert t.info["depth"] <= kw.get("max_depth", 24)


FACE_AXIS = [0, 2, 0, 2, 1, 1]
FACE_HIGH = [False, True, True, False, True, False]


def test_ropes_point_across_their_face(small):
    _, tree, _, _ = small
    nodes, leaves, _ = tree.export()
    info = tree.info
    boxes = _node_boxes(nodes, info["root_min"], info["root_max"])
    leaf_box = {int(nodes[e]["meta"] & 0x3FFFFFFF): b for e, b in boxes.items() if nodes[e]["meta"] >> 30 == 3}
    for lid, l in enumerate(leaves):
        for f in range(6):
            a, high = FACE_AXIS[f], FACE_HIGH[f]
            plane = l["bmax"][a] if high else l["bmin"][a]
            r = int(l["rope"][f])
            if r == -1:
                root = info["root_max"][a] if high else info["root_min"][a]
                assert plane == root
                continue
            lo, hi = leaf_box[r & 0x7FFFFFFF] if r < -1 else boxes[r]
            assert (lo[a] if high else hi[a]) == plane          # shares the face plane
            others = [k for k in range(3) if k != a]
            for k in others:                                     # covers the whole face
                assert lo[k] <= l["bmin"][k] and hi[k] >= l["bmax"][k]


def test_empty_and_tiny_trees():
    t = re.KDTree(np.zeros((0, 9), np.float32))
    assert t.info["n_leaves"] == 1 and t.info["n_trirecs"] == 0
    o = np.array([[0.2, 0.2, 1]], np.float32)
    d = np.array([[0, 0, -1]], np.float32)
    # A flat scene seen head-on has a zero-thickness root box: t_entry == t_exit
    # and the traversal loop (ray_functions.cuh:284) never runs.
    flat = re.KDTree(np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32))
    p, tt = oracle.trace(*flat.export(), flat.identity_object(), o, d)
    assert p[0] == -1 and tt[0] == np.float32(1e20)
    tilted = re.KDTree(np.array([[0, 0, 0, 1, 0, 0.5, 0, 1, 0]], np.float32))
    p, tt = oracle.trace(*tilted.export(), tilted.identity_object(), o, d)
    assert p[0] == 0 and abs(tt[0] - 0.9) < 1e-6


def test_build_rejects_non_finite():
    bad = np.array([[0, 0, 0, 1, np.nan, 0, 0, 1, 0]], np.float32)
    with pytest.raises(re.ReError):
        re.KDTree(bad)


# ------------------------------------------------------------- oracle traversal

def test_traversal_equals_brute_force_below_cap(small):
    tri_v, tree, o, d = small
    nodes, leaves, tris = tree.export()
    p, t, cnt = oracle.trace(nodes, leaves, tris, tree.identity_object(), o, d, cap=1 << 12, counts=True)
    pb, tb = oracle.brute(tri_v, tree.info["root_min"], tree.info["root_max"], o, d)
    # Rays whose walk ended through one of the reference's silent failure
    # exits (cap, leaf slab miss, no face within K_EPSILON) may miss hits; all
    # others must equal brute force.
    ok = np.isin(cnt[:, 3], [0, 4, 5])
    assert ok.mean() > 0.99
    assert np.array_equal(t[ok].view(np.uint32), tb[ok].view(np.uint32))
    # ties at equal t (shared edges) may resolve to another triangle id; none at distinct t
    diff = ok & (p != pb)
    assert diff.sum() <= 0.001 * ok.sum()


def test_fp32_t_against_fp64_brute_force(small):
    tri_v, tree, o, d = small
    nodes, leaves, tris = tree.export()
    p, t = oracle.trace(nodes, leaves, tris, tree.identity_object(), o, d)
    sel = np.nonzero(p >= 0)[0][::16]
    p64, t64 = oracle.brute_f64(tri_v, o[sel], d[sel])
    same = p64 == p[sel]
    assert same.mean() > 0.99
    rel = np.abs(t[sel][same] - t64[same]) / t64[same]
    assert rel.max() < 1e-5


def test_icosphere_centre_ray_closed_form():
    tri_v = re.icosphere(4, 1.0)
    tree = re.KDTree(tri_v)
    nodes, leaves, tris = tree.export()
    o = np.array([[0, 0, 5]], np.float32)
    d = np.array([[0, 0, -1]], np.float32)
    p, t = oracle.trace(nodes, leaves, tris, tree.identity_object(), o, d)
    # hits the outer facet between the circumsphere (t = 4) and the inscribed
    # sphere of the 5120-facet polyhedron (cos of its largest facet angle > 0.9995).
    # (0,0,1) is a vertex of the subdivision, so t = 4 up to FP32 rounding.
    assert p[0] >= 0 and 4.0 - 1e-6 <= t[0] <= 5.0 - 0.9995


def test_cap_semantics(small):
    _, tree, o, d = small
    nodes, leaves, tris = tree.export()
    ob = tree.identity_object()
    p1, t1, c1 = oracle.trace(nodes, leaves, tris, ob, o, d, cap=1, counts=True)
    assert (c1[:, 1] <= 1).all()
    p50, t50, c50 = oracle.trace(nodes, leaves, tris, ob, o, d, cap=50, counts=True)
    assert (c50[:, 1] <= 50).all()
    # more leaf visits extend the same walk, and t_exit only shrinks
    assert (t50 <= t1).all()
    assert ((p1 >= 0) <= (p50 >= 0)).all()


def test_small_scene_golden(small):
    g = np.load(GOLDEN / "small_scene.npz")
    tri_v, tree, o, d = small
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    assert sha(tri_v) == str(g["tri_v_sha"]) and sha(o) == str(g["ray_o_sha"]) and sha(d) == str(g["ray_d_sha"])
    nodes, leaves, tris = tree.export()
    p, t = oracle.trace(nodes, leaves, tris, tree.identity_object(), o, d, cap=50)
    assert np.array_equal(p, g["prim"])
    assert np.array_equal(t.view(np.uint32), g["t_bits"])


# ------------------------------------------------------- reference median rule

def test_median_rule_splits_at_box_midpoints(small):
    """split_rule 1 is the reference's spatial median (KDTree.cpp:192-199,
    368-389): every inner node splits its box's longest side (x on a strict
    win, then y on a strict win over z, else z) at lo + (hi - lo) / 2 in
    FP32; triangles go below when their minimum is < the split and above when
    their maximum is >= it (KDTree.cpp:581-600, clip off)."""
    tri_v = small[0]
    t = re.KDTree(tri_v, split_rule=1, clip=False, max_leaf_tris=20, max_depth=12)
    nodes, leaves, tris = t.export()
    _check_descent_bounds(nodes, len(leaves))
    boxes = _node_boxes(nodes, t.info["root_min"], t.info["root_max"])
    n_inner = 0
    for e, (lo, hi) in boxes.items():
        ax = int(nodes[e]["meta"] >> 30)
        if ax == 3:
            continue
        n_inner += 1
        ext = hi - lo
        want = 0 if (ext[0] > ext[1] and ext[0] > ext[2]) else (1 if ext[1] > ext[2] else 2)
        assert ax == want
        assert nodes[e]["split"] == np.float32(lo[ax] + np.float32((hi[ax] - lo[ax]) / np.float32(2)))
    assert n_inner > 100
    # bounding-box assignment: every triangle whose box reaches into a leaf is in it
    tv = tri_v.reshape(-1, 3, 3)
    tmin, tmax = tv.min(1), tv.max(1)
    for l in leaves[:200]:
        inside = set(tris[l["first"]:l["first"] + l["count"]]["prim"].tolist())
        for p in inside:
            assert (tmin[p] <= l["bmax"]).all() and (tmax[p] >= l["bmin"]).all()
    # the SAH tree of the same soup is a different tree
    assert t.info["n_leaves"] != re.KDTree(tri_v).info["n_leaves"]


def test_median_tree_traversal_equals_brute_force(small):
    tri_v, _, o, d = small
    t = re.KDTree(tri_v, split_rule=1, clip=False)
    nodes, leaves, tris = t.export()
    p, tt, cnt = oracle.trace(nodes, leaves, tris, t.identity_object(), o, d, cap=1 << 12, counts=True)
    pb, tb = oracle.brute(tri_v, t.info["root_min"], t.info["root_max"], o, d)
    ok = np.isin(cnt[:, 3], [0, 4, 5])
    assert ok.mean() > 0.99
    assert np.array_equal(tt[ok].view(np.uint32), tb[ok].view(np.uint32))
    with pytest.raises(re.ReError):
        re.KDTree(tri_v, split_rule=2)
