"""Batched mixed workload: one persistent work-queue kernel in a captured
hipGraph must give, ray for ray, the same results as the single-workload
entry points (bit-exact), on every replay."""
import sys

import numpy as np
import pytest

import rayengine_amd as re
from rayengine_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small_batch(gpu_ctx):
    soups, dims, seeds, ro, rd, sr, tr = W.batch_workload(n_scenes=6, n_tiles=3, tri_lo=1e3, tri_hi=2e4, ray_lo=2e3,
                                                           ray_hi=3e4, tile_lo=64, tile_hi=96, seed=3)
    trees = [re.KDTree(s) for s in soups]
    vols = [gpu_ctx.noise_volume(dim=(dm, dm, dm), spacing=(1.0 / dm,) * 3, seed=sd) for dm, sd in zip(dims, seeds)]
    return trees, vols, ro, rd, sr, tr


def _bits(x):
    return np.ascontiguousarray(x).view(np.uint32)


@pytest.mark.timeout(120, method="thread")
@pytest.mark.parametrize("item_rays,block", [(256, 256), (64, 128), (1000, 64)])
def test_batch_equals_single_calls(gpu_ctx, small_batch, item_rays, block):
    trees, vols, ro, rd, sr, tr = small_batch
    modes = [0, 1, 0]
    say(f"test_batch_equals_single_calls: create item_rays={item_rays} block={block} rays={len(ro)}")
    b = re.Batch(gpu_ctx, [(t, s, c) for t, (s, c) in zip(trees, sr)],
                 [(v, m, s, c) for v, m, (s, c) in zip(vols, modes, tr)], ro, rd, item_rays=item_rays, block=block)
    say(f"batch items={b.items()} item_rays={item_rays} block={block}")
    plan, _ = re.batch_plan(sr, [(m, s, c) for m, (s, c) in zip(modes, tr)], len(ro), item_rays)
    assert b.items() == len(plan)            # the device queue holds the host plan's items
    refs = []
    for tree, (s, c) in zip(trees, sr):
        scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
        p, t = scene.trace_closest(ro[s:s + c], rd[s:s + c])
        refs.append((s, c, p, t))
        scene.release()
    for v, m, (s, c) in zip(vols, modes, tr):
        if m == 0:
            t, h, _, _ = re.hybrid_trace(v, None, ro[s:s + c], rd[s:s + c])
        else:
            t, h, _ = v.trace(ro[s:s + c], rd[s:s + c])
        refs.append((s, c, np.where(h == 1, 0, -1).astype(np.int32), t))
    for rep in range(3):
        say(f"replay {rep}")
        b.run()
        p, t = b.results()
        for s, c, pr, tr_ in refs:
            assert np.array_equal(p[s:s + c], pr), f"replay {rep}: prim mismatch in [{s}, {s + c})"
            assert np.array_equal(_bits(t[s:s + c]), _bits(tr_)), f"replay {rep}: t mismatch"
        ms = b.completion_ms()
        assert (ms > 0).all()
        # eager work between replays
        gpu_ctx.noise_volume(dim=(16, 16, 16), spacing=(1 / 16,) * 3).release()
    b.release()


def test_batch_rejects_bad_ranges(gpu_ctx, small_batch):
    trees, vols, ro, rd, sr, tr = small_batch
    with pytest.raises(re.ReError):
        re.Batch(gpu_ctx, [(trees[0], len(ro) - 5, 10)], [], ro, rd)
    with pytest.raises(re.ReError):
        re.Batch(gpu_ctx, [], [(vols[0], 2, 0, 10)], ro, rd)          # surrogate without a network
    with pytest.raises(re.ReError):
        re.Batch(gpu_ctx, [], [(vols[0], 0, 0, 10)], ro, rd, hybrid=dict(max_iters=0))   # unbounded march
    with pytest.raises(re.ReError):
        re.Batch(gpu_ctx, [], [(vols[0], 1, 0, 10)], ro, rd, sdf=dict(max_steps=0))
    for bad in (-0.1, 0.95, float("nan")):
        with pytest.raises(re.ReError):
            re.Batch(gpu_ctx, [(trees[0], 0, 10)], [], ro, rd, side_share=bad)


@pytest.mark.timeout(120, method="thread")
@pytest.mark.parametrize("side_share", [0.0, 0.5])
def test_batch_with_surrogate_tile_equals_single_calls(gpu_ctx, small_batch, side_share):
    """A surrogate tile runs on its own branch of the graph beside the
    persistent kernel (which may leave part of the device to it); every ray
    still equals the single-call result."""
    trees, vols, ro, rd, sr, tr = small_batch
    dv = re.distance_volume(vols[1])
    net = re.Surrogate(gpu_ctx, seed=2)
    net.train(dv, n_samples=1 << 16, batch=4096, steps=200, lr=1e-3, seed=3)
    dv.release()
    hyb = dict(tau_voxels=4.0, margin_voxels=1.0, k_exact=0.1, n_bisect=4)
    mar = dict(k_exact=0.05, h_min_voxels=1.0, n_bisect=1)
    modes = [0, 2, 1]
    say("test_batch_with_surrogate_tile_equals_single_calls: create")
    b = re.Batch(gpu_ctx, [(t, s, c) for t, (s, c) in zip(trees, sr)],
                 [(v, m, s, c) + ((net,) if m == 2 else ()) for v, m, (s, c) in zip(vols, modes, tr)], ro, rd,
                 item_rays=256, block=128, hybrid=hyb, march=mar, side_share=side_share)
    refs = []
    for tree, (s, c) in zip(trees, sr):
        scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
        p, t = scene.trace_closest(ro[s:s + c], rd[s:s + c])
        refs.append((s, c, p, t))
        scene.release()
    for v, m, (s, c) in zip(vols, modes, tr):
        if m == 0:
            t, h, _, _ = re.hybrid_trace(v, None, ro[s:s + c], rd[s:s + c], **mar)
        elif m == 2:
            t, h, ev, _ = re.hybrid_trace(v, net, ro[s:s + c], rd[s:s + c], **hyb)
            assert ev.sum() > 0
        else:
            t, h, _ = v.trace(ro[s:s + c], rd[s:s + c])
        refs.append((s, c, np.where(h == 1, 0, -1).astype(np.int32), t))
    for rep in range(3):
        say(f"replay {rep}")
        b.run()
        p, t = b.results()
        for s, c, pr, tr_ in refs:
            assert np.array_equal(p[s:s + c], pr), f"replay {rep}: prim mismatch in [{s}, {s + c})"
            assert np.array_equal(_bits(t[s:s + c]), _bits(tr_)), f"replay {rep}: t mismatch in [{s}, {s + c})"
        assert (b.completion_ms() > 0).all()
        gpu_ctx.noise_volume(dim=(16, 16, 16), spacing=(1 / 16,) * 3).release()
    b.release()
    net.release()
