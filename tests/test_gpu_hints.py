"""Walk-start hints and hit leaves (re_trace_closest_hinted) on the GPU.

Bar: prim ids and t bit-identical to the oracle's walk from the root for any
hints - the hit leaves of the primaries passed to their bounce rays, random
leaf ids (in range, out of range, -1) and leaves of other objects - on the
cluster workload and a multi-object scene with both transforms, variants 4
and 5.  The hit leaves are pinned by what defines them: -1 exactly for the
misses, and otherwise a leaf of the hit object whose cell holds the hit point
(the walk's last leaf)."""
import sys

import numpy as np
import pytest

import oracle
import rayengine_amd as re
from rayengine_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _hinted(scene, o, d, variant, start=None, want_leaf=False, **params):
    import torch
    n = len(o)
    do, dd = _dev(o.astype(np.float32)), _dev(d.astype(np.float32))
    p = torch.empty(n, dtype=torch.int32, device="cuda")
    t = torch.empty(n, dtype=torch.float32, device="cuda")
    ob = torch.empty(n, dtype=torch.int32, device="cuda")
    hl = torch.full((n,), 7, dtype=torch.int32, device="cuda") if want_leaf else None
    st = _dev(np.asarray(start, np.int32)) if start is not None else None
    torch.cuda.synchronize()
    cnt = scene.trace_closest_hinted_device(do.data_ptr(), dd.data_ptr(), n, p.data_ptr(), t.data_ptr(),
                                            st.data_ptr() if st is not None else 0,
                                            hl.data_ptr() if hl is not None else 0, ob.data_ptr(), counts=True,
                                            variant=variant, **params)
    return (p.cpu().numpy(), t.cpu().numpy(), ob.cpu().numpy(), None if hl is None else hl.cpu().numpy(), cnt)


def _parity(p, t, po, to):
    assert np.array_equal(p, po), f"{int((p != po).sum())} prim ids differ"
    assert np.array_equal(t.view(np.uint32), to.view(np.uint32)), f"{int((t != to).sum())} t differ"


def _check_hit_leaves(hl, p, t, o, d, leaves, xform=False):
    """-1 exactly for the misses; a leaf id otherwise, whose cell holds the
    hit point (the walk ends in the leaf where its entry point passes the
    hit) unless the walk ended at the visit cap or a missing rope."""
    assert np.array_equal(hl < 0, p < 0), "hit leaf -1 exactly for misses"
    hit = np.flatnonzero(p >= 0)
    assert (hl[hit] < len(leaves)).all()
    if xform:
        return
    x = o[hit] + t[hit, None] * d[hit]
    lf = leaves[hl[hit]]
    tol = 1e-4 * (1 + np.abs(x))
    inside = np.all((x >= lf["bmin"] - tol) & (x <= lf["bmax"] + tol), axis=1)
    say(f"hit points inside their hit leaf's cell: {inside.mean():.5f}")
    assert inside.mean() > 0.99


@pytest.fixture(scope="module")
def cluster_rays(gpu_ctx, cluster):
    tri_v, tree = cluster
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    o, d = W.camera_rays(W.CLUSTER_CAMERA, 1920, 1080)
    order = W.tile_order(1920, 1080, 8)
    o, d = o[order], d[order]
    return scene, tree, o, d


@pytest.mark.parametrize("variant", [4, 5])
def test_cluster_bounces_from_hit_leaves(gpu_ctx, cluster, cluster_rays, variant):
    tri_v, _ = cluster
    scene, tree, o, d = cluster_rays
    nodes, leaves, tris = tree.export()
    obj = tree.identity_object()
    say(f"cluster primaries with hit leaves, variant {variant}")
    p, t, _, hl, c_hint = _hinted(scene, o, d, variant, want_leaf=True)
    po, to = oracle.trace(nodes, leaves, tris, obj, o, d, threads=16)
    _parity(p, t, po, to)
    _check_hit_leaves(hl, p, t, o, d, leaves)
    bo, bd = W.bounce_rays(tri_v, o, d, p, t, 11)
    start = hl[p >= 0]
    assert len(start) == len(bo)
    say(f"cluster bounces ({len(bo)}) from hit leaves, variant {variant}")
    bp, bt, _, _, cb_hint = _hinted(scene, bo, bd, variant, start=start)
    bpo, bto = oracle.trace(nodes, leaves, tris, obj, bo, bd, threads=16)
    _parity(bp, bt, bpo, bto)
    _, _, _, _, cb_plain = _hinted(scene, bo, bd, variant)
    say(f"node entries with hints {cb_hint[0]}, from the root {cb_plain[0]}")
    # the same walks: the same hits, no more leaf visits or triangle tests (a walk started at its first leaf can
    # see a repeated state one visit sooner than one that recorded the root there)
    assert cb_hint[3] == cb_plain[3] and cb_hint[1] <= cb_plain[1] and cb_hint[2] <= cb_plain[2]
    assert cb_hint[1] > 0.999 * cb_plain[1]
    assert cb_hint[0] < 0.75 * cb_plain[0], "the hints save the first descents"


@pytest.mark.parametrize("variant", [4, 5])
def test_random_hints_change_nothing(gpu_ctx, cluster_rays, variant):
    scene, tree, o, d = cluster_rays
    nodes, leaves, tris = tree.export()
    n = 200_000
    rng = np.random.default_rng(5)
    sel = rng.choice(len(o), n, replace=False)
    # rays from points inside the scene box, in random directions
    lo, hi = tree.info["root_min"], tree.info["root_max"]
    ro = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    rd = d[sel]
    start = rng.integers(-3, len(leaves) + 100, n).astype(np.int32)
    say(f"random hints, variant {variant}")
    p, t, _, hl, _ = _hinted(scene, ro, rd, variant, start=start, want_leaf=True)
    po, to = oracle.trace(nodes, leaves, tris, tree.identity_object(), ro, rd, threads=16)
    _parity(p, t, po, to)
    _check_hit_leaves(hl, p, t, ro, rd, leaves)
    # the hit leaves as hints for rays leaving the same hit points in new directions
    hit = p >= 0
    x = ro[hit] + t[hit, None] * rd[hit]
    nd = rng.normal(size=(int(hit.sum()), 3)).astype(np.float32)
    nd /= np.linalg.norm(nd, axis=1, keepdims=True)
    p2, t2, _, _, _ = _hinted(scene, x, nd, variant, start=hl[hit])
    po2, to2 = oracle.trace(nodes, leaves, tris, tree.identity_object(), x, nd, threads=16)
    _parity(p2, t2, po2, to2)


@pytest.mark.parametrize("secondary", [False, True])
@pytest.mark.parametrize("variant", [4, 5])
def test_multi_object_hints(gpu_ctx, variant, secondary):
    a = re.icosphere(3, 1.0)
    b = re.random_tris(5, 3000, -1.5, 1.5, 0.05, 0.3)
    ta, tb = re.KDTree(a), re.KDTree(b)
    objs = [re.SceneObject(ta, location=(0.5, -0.2, 0.1), rotation_y=0.7),
            re.SceneObject(tb, location=(-0.3, 0.4, 0.0), rotation_y=-1.3, prim_offset=len(a)),
            re.SceneObject(ta, location=(0.2, 0.1, -0.4), rotation_y=2.1, is_character=True,
                           prim_offset=len(a) + len(b))]
    scene = gpu_ctx.upload_scene(objs)
    nodes, leaves, tris, obs = scene.export()
    o, d = re.camera_rays(re.Camera(pos=(0.3, 0.5, 6.0), view=(0.0, -0.1, -1.0)), 160, 120)
    say(f"multi-object primaries, variant {variant}, secondary {secondary}")
    p, t, ob, hl, _ = _hinted(scene, o, d, variant, want_leaf=True, secondary=secondary)
    po, to = oracle.trace(nodes, leaves, tris, obs, o, d, secondary=secondary)
    _parity(p, t, po, to)
    _check_hit_leaves(hl, p, t, o, d, leaves, xform=True)
    # every hit leaf belongs to the hit object: leaf ids are the scene's, objects' leaves consecutive
    first_leaf = np.cumsum([0] + [len(o_.tree.export()[1]) for o_ in objs])
    hit = p >= 0
    assert np.array_equal(np.searchsorted(first_leaf, hl[hit], side="right") - 1, ob[hit])
    # secondary rays from the hit points with the hit leaves, then with the leaves of other objects
    rng = np.random.default_rng(2)
    x = o[hit] + t[hit, None] * d[hit]
    nd = rng.normal(size=(int(hit.sum()), 3)).astype(np.float32)
    nd /= np.linalg.norm(nd, axis=1, keepdims=True)
    po2, to2 = oracle.trace(nodes, leaves, tris, obs, x, nd, secondary=secondary)
    for start in (hl[hit], (hl[hit] + first_leaf[1]) % len(leaves), rng.integers(-1, len(leaves), int(hit.sum()))):
        p2, t2, _, _, _ = _hinted(scene, x, nd, variant, start=start, secondary=secondary)
        _parity(p2, t2, po2, to2)


def test_hints_need_a_cooperative_variant(gpu_ctx, cluster_rays):
    scene, _, o, d = cluster_rays
    with pytest.raises(re.ReError):
        _hinted(scene, o[:64], d[:64], 3, start=np.zeros(64, np.int32))
    with pytest.raises(re.ReError):
        _hinted(scene, o[:64], d[:64], 1, want_leaf=True)


def test_variant5_default_block_fills_the_device(gpu_ctx, cluster_rays):
    """With block 0 variant 5 launches one-wave blocks; its persistent grid is
    sized for that block (it once was sized for 256-thread blocks, a quarter
    of the waves the device holds, and ran 2.4x slower)."""
    import torch
    scene, _, o, d = cluster_rays
    n = len(o)
    do, dd = _dev(o), _dev(d)
    p = torch.empty(n, dtype=torch.int32, device="cuda")
    t = torch.empty(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ms = {}
    for block in (0, 64, 0, 64):
        scene.trace_closest_device(do.data_ptr(), dd.data_ptr(), n, p.data_ptr(), t.data_ptr(), variant=5,
                                   block=block)
        for _ in range(5):
            scene.trace_closest_device(do.data_ptr(), dd.data_ptr(), n, p.data_ptr(), t.data_ptr(), sync=False,
                                       variant=5, block=block)
        torch.cuda.synchronize()
        ms.setdefault(block, []).append(float(np.mean(gpu_ctx.kernel_times(5))))
    say(f"variant 5 kernel ms, block 0: {ms[0]}, block 64: {ms[64]}")
    assert min(ms[0]) < 1.25 * min(ms[64])
