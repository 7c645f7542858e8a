"""The premise of the walk-start hint check, on the CPU.

re_trace_closest_hinted lets a walk start at a hinted leaf when the walk's
first point p satisfies bmin <= p < bmax on every axis of that leaf's cell,
instead of descending from the root (below where p < split, above
otherwise, ray_functions.cuh:289-291).  That is sound when every leaf's
cell box is exactly the box its ancestors' splits cut out of the root box,
with every split inside its node's cell.  Checked here on the built trees
(SAH with and without clipping, the reference's median rule, several leaf
sizes), and the consequence checked directly: for random points and points
placed exactly on split planes, the leaves whose box passes the check are
only the leaf the descent reaches.  A planted mistake (one leaf's box
widened by one ulp) is caught."""
import numpy as np
import pytest

import rayengine_amd as re
from rayengine_amd import workloads as W


def _cells(nodes, root_min, root_max):
    """The box each leaf's ancestors cut out of the root box: (n_leaves, 2, 3)."""
    n_leaves = int(max((m & 0x3fffffff) for m in nodes["meta"][(nodes["meta"] >> 30) == 3]) + 1)
    out = np.full((n_leaves, 2, 3), np.nan, np.float32)
    stack = [(0, np.array(root_min, np.float32), np.array(root_max, np.float32))]
    while stack:
        cur, lo, hi = stack.pop()
        meta = int(nodes["meta"][cur])
        ax = meta >> 30
        if ax == 3:
            leaf = meta & 0x3fffffff
            assert np.isnan(out[leaf, 0, 0]), f"leaf {leaf} reached twice"
            out[leaf, 0], out[leaf, 1] = lo, hi
            continue
        s = np.float32(nodes["split"][cur])
        assert lo[ax] <= s <= hi[ax], f"split {s} of entry {cur} outside its cell"
        b_hi = hi.copy()
        b_hi[ax] = s
        a_lo = lo.copy()
        a_lo[ax] = s
        stack.append((cur + 1, lo, b_hi))
        stack.append((meta & 0x3fffffff, a_lo, hi))
    return out


def _descend(nodes, p):
    cur = 0
    while True:
        meta = int(nodes["meta"][cur])
        ax = meta >> 30
        if ax == 3:
            return meta & 0x3fffffff
        cur = cur + 1 if p[ax] < nodes["split"][cur] else meta & 0x3fffffff


def _trees():
    soup = W.cluster_soup_default(n=1 << 14, seed=2)
    ico = re.icosphere(3, 1.0)
    rnd = re.random_tris(5, 3000, -1.5, 1.5, 0.05, 0.3)
    return {"cluster leaf 8": re.KDTree(soup), "cluster leaf 2": re.KDTree(soup, max_leaf_tris=2),
            "cluster no clip": re.KDTree(soup, clip=False), "icosphere median": re.KDTree(ico, split_rule=1),
            "random tris": re.KDTree(rnd)}


@pytest.fixture(scope="module")
def trees():
    return _trees()


def _check_tree(tree, leaves_override=None):
    nodes, leaves, _ = tree.export()
    if leaves_override is not None:
        leaves = leaves_override
    cells = _cells(nodes, tree.info["root_min"], tree.info["root_max"])
    assert len(cells) == len(leaves) and not np.isnan(cells).any(), "a leaf no entry reaches"
    assert np.array_equal(cells[:, 0].view(np.uint32), leaves["bmin"].view(np.uint32)), "leaf bmin is not its cell's"
    assert np.array_equal(cells[:, 1].view(np.uint32), leaves["bmax"].view(np.uint32)), "leaf bmax is not its cell's"
    # the check's consequence on points: random, and on split planes
    rng = np.random.default_rng(7)
    lo, hi = tree.info["root_min"], tree.info["root_max"]
    pts = rng.uniform(lo, hi, (400, 3)).astype(np.float32)
    inner = nodes[(nodes["meta"] >> 30) != 3]
    pick = rng.integers(0, len(inner), 400)
    pts_on = pts.copy()
    pts_on[np.arange(400), inner["meta"][pick] >> 30] = inner["split"][pick]
    bmin, bmax = leaves["bmin"], leaves["bmax"]
    for p in np.concatenate([pts, pts_on]):
        inside = np.flatnonzero(np.all((p >= bmin) & (p < bmax), axis=1))
        want = _descend(nodes, p)
        assert set(inside.tolist()) <= {want}, f"point {p}: check passes for {inside}, descent ends at {want}"


@pytest.mark.parametrize("name", ["cluster leaf 8", "cluster leaf 2", "cluster no clip", "icosphere median",
                                  "random tris"])
def test_leaf_cells_are_the_descent_cells(trees, name):
    _check_tree(trees[name])


def test_widened_leaf_box_is_caught(trees):
    tree = trees["random tris"]
    _, leaves, _ = tree.export()
    bad = leaves.copy()
    k = int(np.flatnonzero(bad["bmax"][:, 0] < tree.info["root_max"][0])[0])
    bad["bmax"][k, 0] = np.nextafter(bad["bmax"][k, 0], np.float32(np.inf))
    with pytest.raises(AssertionError):
        _check_tree(tree, leaves_override=bad)
