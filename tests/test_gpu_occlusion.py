"""GPU any-hit (shadow) queries against the oracle's trace_shadow
restatement: the occlusion flag of every ray equal, through the C ABI."""
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


def test_small_scene_shadow_rays(gpu_ctx, small):
    tri_v, tree, o, d = small
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    nodes, leaves, tris = tree.export()
    obj = tree.identity_object()
    p, t = scene.trace_closest(o, d)
    so, sd = W.shadow_rays(tri_v, o, d, p, t, (3.0, 4.0, 5.0))
    for two_sided in (False, True):
        for cap in (50, 1, 0):
            say(f"shadow rays, two_sided={two_sided}, cap {cap}")
            got = scene.occluded(so, sd, one_sided=not two_sided, max_leaf_visits=cap)
            ref = oracle.occluded(nodes, leaves, tris, obj, so, sd, cap=cap, two_sided=two_sided)
            assert np.array_equal(got, ref), f"{int((got != ref).sum())} flags differ"
    assert 0 < got.sum() < len(got)
    # primary rays (secondary=False) and their closest-hit relation
    got = scene.occluded(o, d, secondary=False)
    assert np.array_equal(got, (p >= 0).astype(np.uint8))
    scene.release()


def test_multi_object_shadow_rays(gpu_ctx):
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
    for secondary in (True, False):
        say(f"multi-object occlusion, secondary={secondary}")
        got = scene.occluded(o, d, secondary=secondary)
        ref = oracle.occluded(nodes, leaves, tris, obs, o, d, secondary=secondary)
        assert np.array_equal(got, ref)
    assert got.sum() > 1000
    scene.release()


def test_cluster_shadow_rays_and_device_path(gpu_ctx, cluster):
    import torch
    tri_v, tree = cluster
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    o, d = re.camera_rays(W.CLUSTER_CAMERA, 1920, 1080)
    p, t = scene.trace_closest(o, d)
    so, sd = W.shadow_rays(tri_v, o, d, p, t, W.CLUSTER_LIGHT)
    say(f"cluster shadow rays: {len(so)}")
    got = scene.occluded(so, sd)
    n = len(so)
    do, dd = torch.from_numpy(so).cuda(), torch.from_numpy(sd).cuda()
    dout = torch.empty(n, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    scene.occluded_device(do.data_ptr(), dd.data_ptr(), n, dout.data_ptr())
    assert np.array_equal(dout.cpu().numpy(), got)
    sel = np.random.default_rng(3).choice(n, 50000, replace=False)
    nodes, leaves, tris = tree.export()
    ref = oracle.occluded(nodes, leaves, tris, tree.identity_object(), so[sel], sd[sel])
    assert np.array_equal(got[sel], ref)
    assert 0 < got.mean() < 1
    assert len(scene.occluded(so[:0], sd[:0])) == 0
    scene.release()


def test_occlusion_bad_arguments(gpu_ctx, small):
    tri_v, tree, o, d = small
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    with pytest.raises(re.ReError):
        scene.occluded(o, d, block=96)
    bogus = re.Scene(gpu_ctx, 987654, [])
    with pytest.raises(re.ReError):
        bogus.occluded(o[:4], d[:4])
    scene.release()
