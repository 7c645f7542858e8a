"""Many scenes and terrain tiles traced one launch per item over two streams
(rayengine_amd.mixed): every triangle ray's prim id and t equal the CPU
oracle's walk bit for bit, every terrain ray equals the single-call
exact-field march or, on a tile with a network, the single-call surrogate
hybrid, and the completion times are measured from each stream's first
launch (gaps included) and bounded by the wall time of the run."""
import sys
import time

import numpy as np
import pytest

import oracle
import rayengine_amd as re
from rayengine_amd import workloads as W
from rayengine_amd.mixed import MixedRun, TILE_HYBRID, TILE_MARCH

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_mixed_run_matches_oracle_and_single_calls(gpu_ctx):
    import torch
    soups, dims, seeds, ro, rd, sr, tr = W.batch_workload(n_scenes=5, n_tiles=2, tri_lo=500, tri_hi=5000,
                                                           ray_lo=500, ray_hi=5000, tile_lo=32, tile_hi=48, seed=11)
    # the second tile is traced by the surrogate hybrid with a network trained on it
    dm1, sd1 = dims[1], seeds[1]
    v1 = gpu_ctx.noise_volume(dim=(dm1, dm1, dm1), spacing=(1.0 / dm1,) * 3, seed=sd1)
    dv1 = re.distance_volume(v1)
    net = re.Surrogate(gpu_ctx, seed=2)
    net.train(dv1, n_samples=1 << 16, batch=4096, steps=200, lr=1e-3, seed=3)
    params = net.params()[0]
    for x in (net, dv1, v1):
        x.release()
    tiles = [(dims[0], seeds[0]), (dm1, sd1, params)]
    say(f"mixed: {len(soups)} scenes, tiles {dims} (the second with a network), {len(ro)} rays")
    trees = [re.KDTree(s) for s in soups]
    ctxs = [gpu_ctx, re.Context(0)]
    dev = torch.device("cuda:0")
    n = len(ro)
    d_o, d_d = torch.from_numpy(ro).to(dev), torch.from_numpy(rd).to(dev)
    d_p = torch.full((n,), -7, dtype=torch.int32, device=dev)
    d_t = torch.zeros(n, dtype=torch.float32, device=dev)
    d_h = torch.full((n,), 7, dtype=torch.uint8, device=dev)
    run = MixedRun(ctxs, trees, tiles, sr, tr, d_o.data_ptr(), d_d.data_ptr(), d_p.data_ptr(), d_t.data_ptr(),
                   d_h.data_ptr())
    try:
        for rep in range(2):
            torch.cuda.synchronize()          # the fills above run on torch's stream
            t0 = time.perf_counter()
            run.run(block=64, variant=4)
            torch.cuda.synchronize()
            wall_ms = (time.perf_counter() - t0) * 1e3
            # read the run's event times before any other launch on these contexts
            done = run.completion_ms()
            ksum = run.kernel_ms_sum()
            p, t, h = d_p.cpu().numpy(), d_t.cpu().numpy(), d_h.cpu().numpy()
            for tree, (b, c) in zip(trees, sr):
                po, to = oracle.trace(*tree.export(), tree.identity_object(), ro[b:b + c], rd[b:b + c])
                assert np.array_equal(p[b:b + c], po), f"run {rep}: prim ids of rays [{b}, {b + c})"
                assert np.array_equal(t[b:b + c].view(np.uint32), to.view(np.uint32)), f"run {rep}: t"
            for tile, (b, c) in zip(tiles, tr):
                dm, sd = tile[0], tile[1]
                v = gpu_ctx.noise_volume(dim=(dm, dm, dm), spacing=(1.0 / dm,) * 3, seed=sd)
                if len(tile) > 2:
                    ref_net = re.Surrogate(gpu_ctx)
                    ref_net.set_params(tile[2])
                    ts, hs, ev, _ = re.hybrid_trace(v, ref_net, ro[b:b + c], rd[b:b + c], **TILE_HYBRID)
                    ref_net.release()
                    assert ev.sum() > 0, "the surrogate tile evaluates its network"
                else:
                    ts, hs, _, _ = re.hybrid_trace(v, None, ro[b:b + c], rd[b:b + c], **TILE_MARCH)
                v.release()
                assert np.array_equal(h[b:b + c], hs), f"run {rep}: tile hits of rays [{b}, {b + c})"
                assert np.array_equal(t[b:b + c].view(np.uint32), ts.view(np.uint32)), f"run {rep}: tile t"
            assert done.shape == (len(run.items),) and (done > 0).all()
            per_stream_last = []
            for s in range(len(ctxs)):
                mine = done[[i for i in range(len(run.items)) if run.owner[i] == s]]
                assert (np.diff(mine) > 0).all(), "completion times grow along a stream"
                per_stream_last.append(mine[-1])
            # a stream's last completion covers its kernels (gaps included) and
            # lies inside the wall time of the whole run
            assert max(per_stream_last) <= wall_ms * 1.001 + 0.05
            assert sum(per_stream_last) >= ksum * (1 - 1e-4)
            say(f"run {rep}: identical; completion {done.max():.3f} ms, kernels {ksum:.3f} ms, wall {wall_ms:.3f} ms")
            d_p.fill_(-7)
            d_h.fill_(7)
    finally:
        run.release()


def test_mixed_run_items_in_ascending_ray_count(gpu_ctx):
    soups, dims, seeds, ro, rd, sr, tr = W.batch_workload(n_scenes=3, n_tiles=2, tri_lo=200, tri_hi=400,
                                                           ray_lo=100, ray_hi=2000, tile_lo=16, tile_hi=16, seed=5)
    trees = [re.KDTree(s) for s in soups]
    run = MixedRun([gpu_ctx], trees, list(zip(dims, seeds)), sr, tr, 0, 0, 0, 0, 0)
    try:
        counts = [it[3] for it in run.items]
        assert counts == sorted(counts)
        assert sorted((it[0], it[1]) for it in run.items) == sorted(
            [("scene", k) for k in range(3)] + [("tile", k) for k in range(2)])
        assert run.owner == [0] * 5
    finally:
        run.release()


def test_tune_mixed_picks_a_measured_configuration(gpu_ctx):
    import torch
    from rayengine_amd.autotune import tune_mixed
    soups, dims, seeds, ro, rd, sr, tr = W.batch_workload(n_scenes=3, n_tiles=1, tri_lo=500, tri_hi=2000,
                                                           ray_lo=500, ray_hi=3000, tile_lo=32, tile_hi=32, seed=4)
    dev = torch.device("cuda:0")
    n = len(ro)
    d_o, d_d = torch.from_numpy(ro).to(dev), torch.from_numpy(rd).to(dev)
    d_p = torch.empty(n, dtype=torch.int32, device=dev)
    d_t = torch.empty(n, dtype=torch.float32, device=dev)
    d_h = torch.empty(n, dtype=torch.uint8, device=dev)
    ptrs = (d_o.data_ptr(), d_d.data_ptr(), d_p.data_ptr(), d_t.data_ptr(), d_h.data_ptr())
    cfg, meas = tune_mixed([gpu_ctx], soups, list(zip(dims, seeds)), sr, tr, ptrs, torch.cuda.synchronize,
                           leaf_sizes=(4, 8), blocks=(64, 128), variants=(4, 5), reps=2, log=say)
    assert len(meas) == 8
    assert {(m["max_leaf_tris"], m["block"], m["variant"]) for m in meas} == {
        (lf, b, v) for lf in (4, 8) for b in (64, 128) for v in (4, 5)}
    best = min(meas, key=lambda m: m["wall_ms"])
    assert (cfg["max_leaf_tris"], cfg["block"], cfg["variant"]) == (best["max_leaf_tris"], best["block"],
                                                                     best["variant"])
    assert cfg["rays"] == n and all(m["wall_ms"] > 0 for m in meas)
