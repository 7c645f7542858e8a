"""GPU parity of the rope kd-tree closest-hit kernel against the CPU oracle.

Bar: primitive ids equal and hit t bit-identical (compared as uint32) on the
same seeded inputs.  Calls go through the C ABI (librayengine.so)."""
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


def _parity(p, t, po, to):
    assert np.array_equal(p, po), f"{int((p != po).sum())} prim ids differ"
    assert np.array_equal(t.view(np.uint32), to.view(np.uint32)), f"{int((t.view(np.uint32) != to.view(np.uint32)).sum())} t differ"


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("cap", [50, 1, 4096])
def test_small_scene_bit_exact(gpu_ctx, small, cap, variant):
    say(f"small scene, cap {cap}, variant {variant}")
    tri_v, tree, o, d = small
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    p, t = scene.trace_closest(o, d, max_leaf_visits=cap, variant=variant)
    po, to = oracle.trace(*tree.export(), tree.identity_object(), o, d, cap=cap)
    _parity(p, t, po, to)
    if cap >= 50:
        assert (p >= 0).sum() > 8000
    scene.release()


def test_small_scene_golden(gpu_ctx, small):
    from pathlib import Path
    g = np.load(Path(__file__).resolve().parent / "golden" / "small_scene.npz")
    tri_v, tree, o, d = small
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    p, t = scene.trace_closest(o, d)
    assert np.array_equal(p, g["prim"]) and np.array_equal(t.view(np.uint32), g["t_bits"])


def test_two_sided_and_block_shapes(gpu_ctx, small):
    tri_v, tree, o, d = small
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    po, to = oracle.trace(*tree.export(), tree.identity_object(), o, d, two_sided=True)
    for block in (64, 128, 256):
        for variant in (0, 1, 2, 3, 4, 5):
            say(f"two-sided, block {block}, variant {variant}")
            p, t = scene.trace_closest(o, d, one_sided=False, block=block, variant=variant)
            _parity(p, t, po, to)


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5])
def test_counters_match_oracle(gpu_ctx, small, variant):
    tri_v, tree, o, d = small
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    p, t, cnt = scene.trace_counts(o, d, variant=variant)
    po, to = oracle.trace(*tree.export(), tree.identity_object(), o, d)
    _parity(p, t, po, to)
    # the kernel stops a walk whose state repeats: its counters are those of
    # the oracle's shortened walk (whose results equal the full walk's)
    _, _, co = oracle.trace(*tree.export(), tree.identity_object(), o, d, counts=True, cut=True)
    assert cnt[:3] == [int(x) for x in co[:, :3].sum(0)]
    assert cnt[3] == int((po >= 0).sum())


def test_multi_object_transforms(gpu_ctx):
    """Several objects with Y rotation, translation, a character object, and
    both primary and secondary transforms (ray_functions.cuh:238-267,376-396)."""
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
    for secondary in (False, True):
        po, to = oracle.trace(nodes, leaves, tris, obs, o, d, secondary=secondary)
        for variant in (0, 1, 2, 3, 4, 5):
            say(f"multi-object, secondary={secondary}, variant {variant}")
            p, t, ob = scene.trace_closest(o, d, with_obj=True, secondary=secondary, variant=variant)
            _parity(p, t, po, to)
        assert set(np.unique(ob[p >= 0])) <= {0, 1, 2}
    assert (p >= 0).sum() > 1000


def test_degenerate_rays(gpu_ctx, small):
    """NaN, zero and axis-aligned directions, origins on leaf planes."""
    tri_v, tree, _, _ = small
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    nodes, leaves, tris = tree.export()
    rng = np.random.default_rng(3)
    n = 4096
    o = rng.uniform(-3, 3, (n, 3)).astype(np.float32)
    d = np.zeros((n, 3), np.float32)
    ax = rng.integers(0, 3, n)
    d[np.arange(n), ax] = rng.choice([-1.0, 1.0], n)
    d[:64] = 0.0
    d[64:128, 0] = np.nan
    o[128:512, 0] = leaves["bmin"][rng.integers(0, len(leaves), 384), 0]
    d[512:1024] = -0.0
    d[512:1024, 2] = -1.0
    po, to = oracle.trace(nodes, leaves, tris, tree.identity_object(), o, d)
    for variant in (0, 1, 4, 5):
        p, t = scene.trace_closest(o, d, variant=variant)
        _parity(p, t, po, to)


def test_empty_batch_and_empty_scene(gpu_ctx):
    tree = re.KDTree(np.zeros((0, 9), np.float32))
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    p, t = scene.trace_closest(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32))
    assert p.size == 0
    o, d = re.camera_rays(W.SMALL_CAMERA, 8, 8)
    for variant in (4, 5):
        p, t = scene.trace_closest(o, d, variant=variant)
        assert (p == -1).all() and (t == np.float32(1e20)).all()
    # a variant outside 0..5 is refused
    with pytest.raises(re.ReError):
        scene.trace_closest(o, d, variant=6)


def test_bad_arguments_report_status(gpu_ctx, small):
    tri_v, tree, o, d = small
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    with pytest.raises(re.ReError):
        scene.trace_closest(o, d, block=96)
    bogus = re.Scene(gpu_ctx, 987654, [])
    with pytest.raises(re.ReError):
        bogus.trace_closest(o[:4], d[:4])


def test_device_pointer_path(gpu_ctx, small):
    import torch
    tri_v, tree, o, d = small
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    to_ = torch.from_numpy(o).cuda()
    td = torch.from_numpy(d).cuda()
    n = o.shape[0]
    tp = torch.empty(n, dtype=torch.int32, device="cuda")
    tt = torch.empty(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    scene.trace_closest_device(to_.data_ptr(), td.data_ptr(), n, tp.data_ptr(), tt.data_ptr())
    p, t = scene.trace_closest(o, d)
    assert np.array_equal(tp.cpu().numpy(), p) and np.array_equal(tt.cpu().numpy().view(np.uint32), t.view(np.uint32))


def test_cluster_scene_every_ray_bit_exact(gpu_ctx, cluster):
    """The full cluster workload (2^20 triangles, 1920x1080 primaries and one
    bounce per hit, about 4 M rays): every ray's prim id and t from every
    kernel variant equal the CPU oracle's walk."""
    tri_v, tree = cluster
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    o, d = re.camera_rays(W.CLUSTER_CAMERA, 1920, 1080)
    say("cluster primary")
    p, t = scene.trace_closest(o, d)
    ro, rd = W.cluster_rays(tri_v, p, t)
    say(f"cluster all rays: {len(ro)}")
    pa, ta = scene.trace_closest(ro, rd)
    assert np.array_equal(pa[:len(o)], p) and np.array_equal(ta[:len(o)].view(np.uint32), t.view(np.uint32))
    nodes, leaves, tris = tree.export()
    say("oracle over every ray")
    po, to = oracle.trace(nodes, leaves, tris, tree.identity_object(), ro, rd)
    _parity(pa, ta, po, to)
    for variant in (0, 1, 2, 4, 5):
        say(f"cluster all rays, variant {variant}")
        pv, tv = scene.trace_closest(ro, rd, variant=variant)
        _parity(pv, tv, po, to)
    # work counters of the whole launch against the oracle's shortened walks
    _, _, co = oracle.trace(nodes, leaves, tris, tree.identity_object(), ro, rd, counts=True, cut=True)
    for variant in (3, 4, 5):
        _, _, cnt = scene.trace_counts(ro, rd, variant=variant)
        assert cnt[:3] == [int(x) for x in co[:, :3].sum(0)], f"variant {variant} counters"
    # work counters of the whole launch against the oracle's shortened walks
    _, _, cnt = scene.trace_counts(ro, rd)
    _, _, co = oracle.trace(nodes, leaves, tris, tree.identity_object(), ro, rd, counts=True, cut=True)
    assert cnt[:3] == [int(x) for x in co[:, :3].sum(0)]
    # size-independent properties over all rays: hits are in front, misses are exact
    hit = pa >= 0
    assert (ta[hit] > 1e-4).all() and (ta[~hit] == np.float32(1e20)).all()
    assert hit[len(o):].mean() > 0.3


def _attr_parity(got, want):
    for name, g, w in zip(("normal", "uv", "bary", "hit_point"), got, want):
        gb, wb = g.view(np.uint32), w.view(np.uint32)
        same = (gb == wb) | (np.isnan(g) & np.isnan(w))
        assert same.all(), f"{name}: {int((~same).sum())} values differ"


def _prim_geometry(tris, n_prims):
    geom = np.zeros((n_prims, 9), np.float32)
    geom[tris["prim"], 0:3] = tris["v0"]
    geom[tris["prim"], 3:6] = tris["e1"]
    geom[tris["prim"], 6:9] = tris["e2"]
    return geom


def test_hit_attributes_bit_exact(gpu_ctx, small):
    """Normals, uvs, barycentrics and hit points of traced rays
    (ray_functions.cuh:146-166,204-213) equal the oracle's bit for bit, with
    and without vertex uvs; misses give zeros."""
    tri_v, tree, o, d = small
    n_prims = len(tri_v)
    rng = np.random.default_rng(8)
    normals = rng.normal(size=(n_prims, 9)).astype(np.float32)
    uvs = rng.uniform(0, 1, (n_prims, 6)).astype(np.float32)
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    nodes, leaves, tris, obs = scene.export()
    geom = _prim_geometry(tris, n_prims)
    p, t = scene.trace_closest(o, d)
    for with_uv in (True, False):
        say(f"hit attributes, uvs={with_uv}")
        scene.set_attributes(normals, uvs if with_uv else None)
        got = scene.hit_attributes(o, d, p)
        want = oracle.hit_attributes(geom, normals, uvs if with_uv else None, obs, o, d, p)
        _attr_parity(got, want)
        assert np.all(got[0][p < 0] == 0) and not np.isnan(got[0][p >= 0]).any()
    # the hit point lies at the traced t along the unit direction
    h = p >= 0
    dn = d / np.linalg.norm(d, axis=1, keepdims=True)
    assert np.abs(got[3][h] - (o[h] + t[h, None] * dn[h])).max() < 1e-4
    # prim ids outside the scene give NaN, not a read past the attribute arrays
    bad = p.copy()
    bad[:10] = n_prims + 5
    nrm, _, _, _ = scene.hit_attributes(o, d, bad)
    assert np.isnan(nrm[:10]).all()
    scene.release()


def test_hit_attributes_multi_object(gpu_ctx):
    a = re.icosphere(3, 1.0)
    b = re.random_tris(5, 3000, -1.5, 1.5, 0.05, 0.3)
    ta, tb = re.KDTree(a), re.KDTree(b)
    objs = [re.SceneObject(ta, location=(0.5, -0.2, 0.1), rotation_y=0.7),
            re.SceneObject(tb, location=(-0.3, 0.4, 0.0), rotation_y=-1.3, prim_offset=len(a)),
            re.SceneObject(ta, location=(0.2, 0.1, -0.4), rotation_y=2.1, is_character=True,
                           prim_offset=len(a) + len(b))]
    n_prims = 2 * len(a) + len(b)
    scene = gpu_ctx.upload_scene(objs)
    nodes, leaves, tris, obs = scene.export()
    geom = _prim_geometry(tris, n_prims)
    rng = np.random.default_rng(9)
    normals = rng.normal(size=(n_prims, 9)).astype(np.float32)
    uvs = rng.uniform(0, 1, (n_prims, 6)).astype(np.float32)
    scene.set_attributes(normals, uvs)
    o, d = re.camera_rays(re.Camera(pos=(0.3, 0.5, 6.0), view=(0.0, -0.1, -1.0)), 160, 120)
    for secondary in (False, True):
        say(f"hit attributes, multi-object, secondary={secondary}")
        p, t, ob = scene.trace_closest(o, d, with_obj=True, secondary=secondary)
        got = scene.hit_attributes(o, d, p, obj=ob, secondary=secondary)
        want = oracle.hit_attributes(geom, normals, uvs, obs, o, d, p, obj=ob, secondary=secondary)
        _attr_parity(got, want)
        assert not np.isnan(got[0][p >= 0]).any()
    scene.release()


def test_median_tree_bit_exact(gpu_ctx, small):
    """The reference's spatial-median tree (split_rule 1, bounding-box
    assignment) traced on the device equals the oracle's walk of it."""
    tri_v, _, o, d = small
    tree = re.KDTree(tri_v, split_rule=1, clip=False)
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    po, to = oracle.trace(*tree.export(), tree.identity_object(), o, d)
    for variant in (0, 3, 4, 5):
        say(f"median tree, variant {variant}")
        p, t = scene.trace_closest(o, d, variant=variant)
        _parity(p, t, po, to)
    assert (p >= 0).sum() > 8000
    scene.release()
    with pytest.raises(re.ReError):
        re.KDTree(tri_v, split_rule=1, ctx=gpu_ctx)   # the device build has the SAH rule only


def test_trace_status_matches_oracle_end_reasons(gpu_ctx, small):
    """re_trace_status: prim and t as the closest-hit query, and each ray's
    end reason equal to the oracle's (the walk that stops on a repeated
    state, cut=True); non-finite or zero directions flagged with 0x80."""
    tri_v, tree, o, d = small
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    nodes, leaves, tris = tree.export()
    rng = np.random.default_rng(12)
    extra_o = rng.uniform(-3, 3, (512, 3)).astype(np.float32)
    extra_d = np.zeros((512, 3), np.float32)
    ax = rng.integers(0, 3, 512)
    extra_d[np.arange(512), ax] = rng.choice([-1.0, 1.0], 512)
    extra_d[:32] = 0.0
    extra_d[32:64, 1] = np.nan
    extra_o[64:96, 2] = np.inf
    oo, dd = np.concatenate([o, extra_o]), np.concatenate([d, extra_d])
    for cap in (50, 3):
        say(f"trace status, cap {cap}")
        po, to, co = oracle.trace(nodes, leaves, tris, tree.identity_object(), oo, dd, cap=cap, counts=True, cut=True)
        p, t, st = scene.trace_status(oo, dd, max_leaf_visits=cap)
        _parity(p, t, po, to)
        assert np.array_equal(st & 0x7F, co[:, 3].astype(np.uint8))
        bad = ~(np.isfinite(oo).all(1) & np.isfinite(dd).all(1) & (dd != 0).any(1))
        assert np.array_equal((st & 0x80) != 0, bad) and bad.sum() == 96
        assert set(np.unique(st & 0x7F)) >= {0, 5}
    assert (st & 0x7F == 1).sum() > 0          # some walks reach a cap of 3
    scene.release()
