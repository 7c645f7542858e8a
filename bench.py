"""Benchmark of the rope kd-tree closest-hit path on one MI355X.

Workload (the cluster scene): 2^20 FP32 triangles in 64 Gaussian clusters
(sigma 0.3, centres U[-10,10]^3, edge lengths LogNormal(-3, 0.5), seed 1), SAH
kd-tree with ropes (leaf <= 8, depth <= 24); 1920x1080 primary rays from
(0,0,30) towards the origin (FOV 60), issued in 8x8 pixel tiles, plus one
cosine-weighted diffuse bounce ray per primary hit, in the order of their
primaries.  One step = one closest-hit launch over every ray, with rays and
results resident in HBM.

Prints ONE JSON line.  `value` is rays traced per second (Mrays/s) over the
timed steps.  `roofline` uses the algorithmic bytes of a launch, from the
kernel's own work counters: per ray 24 B in + 8 B out, 8 B per node entry,
56 B per leaf record, 40 B per triangle record; divided by the mean kernel
duration from HIP events recorded on the launch stream during the timed
steps.  `cpu_baseline` times the CPU oracle on a bounded sample of the same
rays.  `terrain` reports the terrain path (noise volume, surrogate training
and inference, and Mrays/s with hit depth error of the reference sphere
tracer, the exact-field march and the surrogate hybrid).  `mixed_workload`
runs the many-scene, many-tile workload one launch per scene or tile.
"""
from __future__ import annotations

import argparse
import json
import re as re_mod
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s closest-hit; terrain Mrays/s surrogate vs exact SDF at ≤0.5-voxel err"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(tree, ro, rd, n_primary, threads, budget_cpu_s=20.0):
    """Oracle traversal (the reference's walk, no shortcut) timed on the host
    cores: a stratified sample (primary and bounce rays in the workload's
    proportion) or, when the whole workload takes less, whole passes over it,
    sized to about budget_cpu_s of CPU time (threads x wall)."""
    import oracle
    nodes, leaves, tris = tree.export()
    obj = tree.identity_object()
    rng = np.random.default_rng(0)
    n = len(ro)
    frac_b = (n - n_primary) / n
    probe = rng.choice(n, 20000, replace=False)
    t0 = time.perf_counter()
    oracle.trace(nodes, leaves, tris, obj, ro[probe], rd[probe], threads=threads)
    rate = 20000 / max(time.perf_counter() - t0, 1e-6)
    want = rate * budget_cpu_s / threads
    whole = want >= n
    if whole:
        so, sd = np.ascontiguousarray(ro), np.ascontiguousarray(rd)
        m, mb = n, n - n_primary
    else:
        m = int(max(20000, want))
        mb = int(round(m * frac_b))
        sel = np.concatenate([rng.choice(n_primary, m - mb, replace=False),
                              n_primary + rng.choice(n - n_primary, mb, replace=False)])
        so, sd = np.ascontiguousarray(ro[sel]), np.ascontiguousarray(rd[sel])
    reps = 0
    t0 = time.perf_counter()
    while True:
        oracle.trace(nodes, leaves, tris, obj, so, sd, threads=threads)
        reps += 1
        dt = time.perf_counter() - t0
        if not whole or dt * threads >= budget_cpu_s:
            break
    what = (f"the whole workload ({n} rays: {n_primary} primary + {n - n_primary} bounce) x {reps} passes"
            if whole else f"{m} rays ({m - mb} primary + {mb} bounce, uniform sample of the workload)")
    return {"value": round(m * reps / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{what}, CPU oracle (FP32 restatement of the reference traversal, full walks to the cap), "
                      f"OpenMP dynamic/64 over {threads} threads, {dt:.1f} s wall"}


# Terrain settings (hit depth error is measured against REFERENCE_DEPTH, the
# fine exact-field march that tests/test_gpu_terrain.py pins to the FP64
# first crossing within 0.02 voxel at the 99th percentile).
REFERENCE_DEPTH = dict(k_exact=0.02, h_min_voxels=0.25, n_bisect=20, max_iters=1 << 16)
TERRAIN_METHODS = {
    # the reference sphere tracer (renderer.cu:142-235) at its own settings
    "reference_march": ("plain", dict(step_scale=0.8, eps=1e-3, max_steps=256)),
    # the same tracer at the smallest step and tolerance tried: its most accurate setting
    "reference_march_fine": ("plain", dict(step_scale=0.02, eps=1e-5, max_steps=8192)),
    # exact-field bracketing march (no network)
    "exact_march": ("exact", dict(k_exact=0.05, h_min_voxels=1.0, n_bisect=1)),
    # surrogate far field + exact near field
    "surrogate_hybrid": ("hybrid", dict(tau_voxels=4.0, margin_voxels=1.0, k_exact=0.1, n_bisect=4)),
}
MLP_FLOP_PER_EVAL = 2 * (99 * 64 + 3 * 64 * 64 + 64)
# forward + backward (input and weight gradients) per training sample, about
# three forward passes
MLP_TRAIN_FLOP_PER_SAMPLE = 3 * MLP_FLOP_PER_EVAL
NOISE_FLOP_PER_VOXEL = 5 * 70
FP32_PEAK_TFLOPS = 157.3
# Surrogate training: 2^22 samples (half within 2 voxels of the surface), Adam
# lr 1e-3, batch 65,536, 20k steps.
TRAIN_STEPS = 20000
BF16_PEAK_TFLOPS = 2500.0


def terrain_section(re, W, torch, ctx):
    """Terrain path on one GPU: the 256^3 fBm volume (seed 2016) over the unit
    cube, 1920x1080 rays from the terrain camera; noise generation, training,
    network throughput, and hit depth error / Mrays/s of each tracer."""
    dev = torch.device("cuda:0")
    out = {"workload": "256^3 fBm terrain (5 octaves, seed 2016) on the unit cube, 1920x1080 rays"}
    vol = ctx.noise_volume()
    noise = []
    for _ in range(3):
        v2 = ctx.noise_volume()
        noise.append(ctx.last_kernel_ms())
        v2.release()
    out["noise_ms"] = round(float(np.mean(noise)), 4)
    out["noise_gvox_s"] = round(256 ** 3 / (np.mean(noise) * 1e-3) / 1e9, 2)
    # VALU-bound: about 5 octaves x 70 FP32 operations per voxel, and a 4-byte write
    nsec = float(np.mean(noise)) * 1e-3
    out["noise_roofline"] = {"bound": "valu", "flop_per_voxel": NOISE_FLOP_PER_VOXEL,
                             "achieved_tflops": round(256 ** 3 * NOISE_FLOP_PER_VOXEL / nsec / 1e12, 2),
                             "peak_tflops": FP32_PEAK_TFLOPS,
                             "frac": round(256 ** 3 * NOISE_FLOP_PER_VOXEL / nsec / 1e12 / FP32_PEAK_TFLOPS, 4),
                             "write_gbs": round(256 ** 3 * 4 / nsec / 1e9, 1)}
    # brush edit (SPHERE_ADD) over the whole 256^3 volume and its material
    # channel: 8 bytes read and 8 written per voxel
    ev = ctx.noise_volume()
    ev.edit("sphere_add", center=(0.5, 0.45, 0.5), extent=(0.1, 0, 0), material=1.0)
    ems = []
    for k in range(5):
        ev.edit("sphere_add", center=(0.3 + 0.1 * k, 0.45, 0.5), extent=(0.1, 0, 0), material=1.0)
        ems.append(ctx.last_kernel_ms())
    ev.release()
    em = float(np.mean(ems))
    out["edit"] = {"kernel": "volume_edit_kernel", "kernel_ms": round(em, 4), "bound": "hbm",
                   "achieved_gbs": round(256 ** 3 * 16 / (em * 1e-3) / 1e9, 1), "peak_gbs": HBM_PEAK_GBS,
                   "frac": round(256 ** 3 * 16 / (em * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    dv = re.distance_volume(vol)
    net = re.Surrogate(ctx, seed=1)
    t0 = time.perf_counter()
    hist = net.train(dv, n_samples=1 << 22, batch=65536, steps=TRAIN_STEPS, lr=1e-3, seed=1, over_weight=4.0)
    step_ms = (time.perf_counter() - t0) * 1e3 / TRAIN_STEPS
    train_tf = MLP_TRAIN_FLOP_PER_SAMPLE * 65536 / (step_ms * 1e-3) / 1e12
    out["train"] = {"steps": TRAIN_STEPS, "batch": 65536, "ms_per_step": round(step_ms, 4),
                    "loss_first": float(hist[:10].mean()), "loss_last": float(hist[-100:].mean()),
                    "flop_per_sample": MLP_TRAIN_FLOP_PER_SAMPLE, "tflops": round(train_tf, 2),
                    "frac_of_bf16_peak": round(train_tf / BF16_PEAK_TFLOPS, 4)}
    # network throughput (re_mlp_forward on 4M device points)
    npts = 1 << 22
    pts = torch.rand(npts, 3, device=dev)
    val = torch.empty(npts, device=dev)
    torch.cuda.synchronize()
    for _ in range(2):
        net.forward_device(pts.data_ptr(), npts, val.data_ptr())
    for _ in range(5):
        net.forward_device(pts.data_ptr(), npts, val.data_ptr(), sync=False)
    torch.cuda.synchronize()
    fms = float(np.mean(ctx.kernel_times(5)))
    tf = npts * MLP_FLOP_PER_EVAL / (fms * 1e-3) / 1e12
    out["mlp_forward"] = {"kernel": "mlp_infer_kernel", "gevals_s": round(npts / (fms * 1e-3) / 1e9, 3),
                          "tflops": round(tf, 2), "bound": "mfma", "peak_tflops": BF16_PEAK_TFLOPS,
                          "frac": round(tf / BF16_PEAK_TFLOPS, 4), "flop_per_eval": MLP_FLOP_PER_EVAL}
    o, d = W.terrain_rays(1920, 1080)
    n = len(o)
    do, dd = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    dt = torch.empty(n, dtype=torch.float32, device=dev)
    dh = torch.empty(n, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()

    def run(kind, params, sync=True):
        if kind == "plain":
            vol.trace_device(do.data_ptr(), dd.data_ptr(), n, dt.data_ptr(), dh.data_ptr(), sync=sync, **params)
        else:
            re.hybrid_trace_device(vol, net if kind == "hybrid" else None, do.data_ptr(), dd.data_ptr(), n,
                                   dt.data_ptr(), dh.data_ptr(), sync=sync, **params)

    run("exact", REFERENCE_DEPTH)
    ref_t, ref_h = dt.cpu().numpy().astype(np.float64), dh.cpu().numpy()
    vox = 1.0 / 256
    methods = {}
    for name, (kind, params) in TERRAIN_METHODS.items():
        run(kind, params)
        t, h = dt.cpu().numpy().astype(np.float64), dh.cpu().numpy()
        both = (h == 1) & (ref_h == 1)
        e = np.abs(t[both] - ref_t[both]) / vox
        for _ in range(5):
            run(kind, params, sync=False)
        torch.cuda.synchronize()
        ms = float(np.mean(ctx.kernel_times(5)))
        # field samples per ray (untimed host-pointer call): 32 bytes of the
        # trilinear cell per sample + 32 bytes of ray in and result out
        if kind == "plain":
            _, _, steps = vol.trace(o, d, **params)
        else:
            _, _, evals, steps = re.hybrid_trace(vol, net if kind == "hybrid" else None, o, d, **params)
        alg = float(32 * n + 32 * np.sum(steps, dtype=np.int64))
        roof = {"samples_per_ray": round(float(np.mean(steps)), 2),
                "algorithmic_gbs": round(alg / (ms * 1e-3) / 1e9, 1),
                "hbm_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        if kind == "hybrid":
            roof["network_evals_per_ray"] = round(float(np.mean(evals)), 2)
        methods[name] = {"mrays_s": round(n / (ms * 1e-3) / 1e6, 1), "kernel_ms": round(ms, 4), "roofline": roof,
                         "err_p50_vox": round(float(np.median(e)), 4), "err_p99_vox": round(float(np.percentile(e, 99)), 4),
                         "err_max_vox": round(float(e.max()), 3) if e.size else None,
                         "frac_err_gt_half_vox": round(float(np.mean(e > 0.5)), 5),
                         "hit_agree": round(float(np.mean(h == ref_h)), 5), "params": params}
    out["methods"] = methods
    # frame output of the reference tracer's frame (sky for misses, shaded hits, RGBA8)
    dn = torch.empty(n * 3, dtype=torch.float32, device=dev)
    dpx = torch.empty(n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    vol.trace_device(do.data_ptr(), dd.data_ptr(), n, dt.data_ptr(), dh.data_ptr(), nrm_ptr=dn.data_ptr(),
                     **TERRAIN_METHODS["reference_march"][1])
    fr = []
    for _ in range(6):
        re.frame_shade_device(ctx, dd.data_ptr(), dh.data_ptr(), dn.data_ptr(), 1920, 1080, dpx.data_ptr())
        fr.append(ctx.last_kernel_ms())
    fms = float(np.mean(fr[1:]))
    out["frame"] = {"kernel": "frame_kernel", "kernel_ms": round(fms, 4), "mpixels_s": round(n / (fms * 1e-3) / 1e6, 1),
                    "sky_pixels": int(n - int(dh.sum().item()))}
    hy, fine, ex = methods["surrogate_hybrid"], methods["reference_march_fine"], methods["exact_march"]
    out["surrogate_vs_reference_march"] = round(hy["mrays_s"] / fine["mrays_s"], 3)
    out["surrogate_vs_exact_march"] = round(hy["mrays_s"] / ex["mrays_s"], 3)
    reach = "reaches" if fine["err_p99_vox"] <= 0.5 else "does not reach"
    out["note"] = (f"reference_march_fine is the reference tracer's most accurate setting tried; it {reach} 0.5 voxel "
                   f"at p99. exact_march (no network) is the ablation of the surrogate.")
    for x in (net, dv, vol):
        x.release()
    return out


def mixed_workload_section(re, W, torch, ctx, reps=3, n_streams=4):
    """The batched mixed workload (workloads.batch_workload, seed 7: 64
    triangle scenes of 1k-500k triangles with 16k-2M incoherent rays each, 16
    noise-terrain tiles of 64^3-256^3 with their own rays), run unbatched by
    rayengine_amd.mixed: one closest-hit launch per scene and, per tile, the
    exact-field march (even tiles) or the surrogate hybrid with a network
    trained on that tile (odd tiles), in ascending ray count, every input resident in HBM, on
    one stream and dealt round-robin over n_streams contexts (one HIP stream
    each).  Leaf size and block are picked first by autotune.tune_mixed on the
    n_streams issue.  Aggregate Mrays/s over the wall time of the whole
    workload; an item's completion time is the device time from the start of
    its stream's first launch to the end of its own."""
    from rayengine_amd.autotune import tune_mixed
    from rayengine_amd.mixed import MixedRun
    t0 = time.perf_counter()
    soups, dims, seeds, ro, rd, sr, tr = W.batch_workload()
    # every other tile is traced by the surrogate hybrid with a network of its
    # own (2,000 Adam steps on 2^20 samples of that tile's distance field)
    t_net = time.perf_counter()
    tiles = []
    for k, (dm, sd) in enumerate(zip(dims, seeds)):
        if k % 2 == 1:
            v = ctx.noise_volume(dim=(dm, dm, dm), spacing=(1.0 / dm,) * 3, seed=sd)
            dv = re.distance_volume(v)
            net = re.Surrogate(ctx, seed=1)
            net.train(dv, n_samples=1 << 20, batch=65536, steps=2000, lr=1e-3, seed=1, over_weight=4.0)
            tiles.append((dm, sd, net.params()[0]))
            net.release()
            dv.release()
            v.release()
        else:
            tiles.append((dm, sd))
    net_s = time.perf_counter() - t_net
    ctxs = [ctx] + [re.Context(0) for _ in range(n_streams - 1)]
    dev = torch.device("cuda:0")
    n = len(ro)
    d_o, d_d = torch.from_numpy(ro).to(dev), torch.from_numpy(rd).to(dev)
    d_p = torch.empty(n, dtype=torch.int32, device=dev)
    d_t = torch.empty(n, dtype=torch.float32, device=dev)
    d_h = torch.empty(n, dtype=torch.uint8, device=dev)
    ptrs = (d_o.data_ptr(), d_d.data_ptr(), d_p.data_ptr(), d_t.data_ptr(), d_h.data_ptr())
    cfg, meas = tune_mixed(ctxs, soups, tiles, sr, tr, ptrs, torch.cuda.synchronize, log=log)
    trees = [re.KDTree(s, max_leaf_tris=cfg["max_leaf_tris"], ctx=ctx) for s in soups]
    n_scene_rays = int(sum(c for _, c in sr))
    out = {"workload": "64 cluster-soup scenes (1k-500k triangles, log-uniform) with 16k-2M incoherent rays each "
                       "(log-uniform), 16 noise-terrain tiles 64^3-256^3 (seeds 2016+k) with 16k-2M rays each; "
                       "seed 7",
           "mode": "unbatched: one launch per scene or tile, ascending ray count (the batched persistent launch "
                   "is switched off)",
           "triangles": int(sum(len(s) for s in soups)), "rays": int(n), "scene_rays": n_scene_rays,
           "tile_rays": int(n - n_scene_rays),
           "tiles_traced_with": "even tiles: exact-field march (k 0.05, h_min 1 voxel); odd tiles: surrogate hybrid "
                                "(tau 4 voxels, margin 1, k 0.1, 4 bisections) with a network per tile, 2,000 Adam "
                                "steps on 2^20 samples of its field",
           "tile_networks_train_s": round(net_s, 2),
           "launches": len(sr) + len(tr),
           "autotune": {"chosen": {"max_leaf_tris": cfg["max_leaf_tris"], "block": cfg["block"]},
                        "issue": f"{n_streams} streams",
                        "wall_ms": {f"leaf{m['max_leaf_tris']}_block{m['block']}": m["wall_ms"] for m in meas}}}
    first = None
    for streams in (1, n_streams):
        run = MixedRun(ctxs[:streams], trees, tiles, sr, tr, *ptrs)
        run.run(block=cfg["block"], variant=cfg["variant"])
        torch.cuda.synchronize()
        walls, comps, ksums = [], [], []
        for _ in range(reps):
            torch.cuda.synchronize()
            ts = time.perf_counter()
            run.run(block=cfg["block"], variant=cfg["variant"])
            torch.cuda.synchronize()
            walls.append(time.perf_counter() - ts)
            comps.append(run.completion_ms())
            ksums.append(run.kernel_ms_sum())
        wall = float(np.median(walls))
        done = np.mean(comps, axis=0)
        sd = np.array([done[i] for i, it in enumerate(run.items) if it[0] == "scene"])
        td = np.array([done[i] for i, it in enumerate(run.items) if it[0] == "tile"])
        rec = {"wall_ms": round(wall * 1e3, 3), "mrays_s": round(n / wall / 1e6, 1),
               "kernel_ms_sum": round(float(np.mean(ksums)), 3),
               "scene_completion_ms": {"p50": round(float(np.percentile(sd, 50)), 3),
                                       "p99": round(float(np.percentile(sd, 99)), 3)},
               "tile_completion_ms": {"p50": round(float(np.percentile(td, 50)), 3),
                                      "p99": round(float(np.percentile(td, 99)), 3)}}
        got = (d_p.cpu().numpy(), d_t.cpu().numpy().view(np.uint32), d_h.cpu().numpy())
        if first is None:
            first = got
        else:
            rec["identical_to_one_stream"] = all(np.array_equal(x, y) for x, y in zip(first, got))
        out[f"streams_{streams}"] = rec
        run.release()
    out["prep_and_tune_s"] = round(time.perf_counter() - t0, 1)
    return out


def kd_build_section(re, ctx, tri_v, host_tree, max_leaf, host_build_s, reps=3):
    """Device kd-tree build of the cluster soup (re_kd_build_gpu) beside the
    host build, and whether the two trees are identical."""
    dev = None
    ms, walls = [], []
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        dev = re.KDTree(tri_v, max_leaf_tris=max_leaf, ctx=ctx)
        walls.append(time.perf_counter() - t0)
        ms.append(dev.build_ms)
    same = all(np.array_equal(a.view(np.uint8), b.view(np.uint8)) for a, b in zip(host_tree.export(), dev.export()))
    return {"triangles": int(len(tri_v)), "device_ms": round(float(np.mean(ms[1:])), 2),
            "call_ms": round(float(np.mean(walls[1:])) * 1e3, 1), "host_build_s": round(host_build_s, 2),
            "speedup_vs_host_call": round(host_build_s / float(np.mean(walls[1:])), 1), "identical_to_host": bool(same),
            "note": "call_ms includes the host input check, the upload of the soup and the download of the arrays"}


def shadow_section(W, torch, ctx, scene, tri_v, po, pd, pp, pt, order, steps=10):
    """Any-hit (shadow) queries: one ray from each primary hit toward a point
    light (pixels in the tiled order of the closest-hit line), timed on the
    occlusion kernel, beside the closest-hit kernel on the same rays."""
    so, sd = W.shadow_rays(tri_v, po[order], pd[order], pp[order], pt[order], W.CLUSTER_LIGHT)
    n = len(so)
    dev = torch.device("cuda:0")
    d_o = torch.from_numpy(so).to(dev)
    d_d = torch.from_numpy(sd).to(dev)
    d_occ = torch.empty(n, dtype=torch.uint8, device=dev)
    d_p = torch.empty(n, dtype=torch.int32, device=dev)
    d_t = torch.empty(n, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    out = {"rays": n, "light": list(W.CLUSTER_LIGHT)}
    for name, fn in (("any_hit", lambda s: scene.occluded_device(d_o.data_ptr(), d_d.data_ptr(), n,
                                                                 d_occ.data_ptr(), sync=s)),
                     ("closest_hit", lambda s: scene.trace_closest_device(d_o.data_ptr(), d_d.data_ptr(), n,
                                                                          d_p.data_ptr(), d_t.data_ptr(),
                                                                          sync=s, secondary=True))):
        fn(True)
        for _ in range(steps):
            fn(False)
        torch.cuda.synchronize()
        kms = float(np.mean(ctx.kernel_times(steps)))
        out[name] = {"kernel_ms": round(kms, 4), "Mrays_per_s": round(n / (kms * 1e-3) / 1e6, 1)}
    occ = d_occ.cpu().numpy()
    out["occluded_frac"] = round(float(occ.mean()), 4)
    out["agrees_with_closest_hit"] = bool(np.array_equal(occ, (d_p.cpu().numpy() >= 0).astype(np.uint8)))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-terrain", action="store_true", help="skip the terrain section")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-workload section")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run 5 device copies of BYTES bytes (counter calibration)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--autotune", action="store_true",
                    help="pick leaf size, ray tile, block and variant by measurement before timing")
    args = ap.parse_args()
    if args.calibrate_copy:
        # A device-to-device copy of known size: the reference point that turns
        # the FETCH_SIZE / WRITE_SIZE counters into bytes.
        import torch
        n = args.calibrate_copy // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(rays would shard across GPUs with no exchange between them)", file=sys.stderr)
        sys.exit(2)

    import torch
    import rayengine_amd as re
    from rayengine_amd import workloads as W

    tri_v = W.cluster_soup_default()
    ctx = re.Context(0)
    # Leaf size, ray-sort granularity (primary rays in tile x tile pixel
    # blocks, bounces following their primaries: the order a tiled renderer
    # issues them in), block and variant.  The defaults are within 1.3 % of
    # the autotuner's best on this workload (leaf 4, same tile, block and
    # variant), inside run-to-run noise; leaf 8 keeps the leaf bound of the
    # small-scene configuration.  --autotune measures them again.
    tune = {"max_leaf_tris": 8, "tile": 8, "block": 64, "variant": 3}
    tuned = None
    if args.autotune:
        from rayengine_amd.autotune import tune_trace
        tuned, _ = tune_trace(ctx, tri_v, lambda sc, tile: W.cluster_workload_rays(sc, tri_v, tile), 1920, 1080,
                              device_name=torch.cuda.get_device_name(0), cache=None, log=log)
        tune = {k: tuned[k] for k in tune}
    t0 = time.time()
    tree = re.KDTree(tri_v, max_leaf_tris=tune["max_leaf_tris"])
    log(f"scene: {tri_v.shape[0]} tris, tree {tree.info['n_leaves']} leaves, "
        f"{tree.info['n_trirecs']} tri records, build {time.time() - t0:.1f} s")
    host_build_s = time.time() - t0
    kd_gpu = kd_build_section(re, ctx, tri_v, tree, tune["max_leaf_tris"], host_build_s)
    scene = ctx.upload_scene([re.SceneObject(tree)])
    po, pd = re.camera_rays(W.CLUSTER_CAMERA, 1920, 1080)
    pp, pt = scene.trace_closest(po, pd)
    order = W.tile_order(1920, 1080, tune["tile"])
    ro, rd = W.cluster_rays(tri_v, pp, pt, order=order)
    tp = {"block": tune["block"], "variant": tune["variant"]}
    n = len(ro)
    n_primary = len(po)
    log(f"rays: {n} ({n_primary} primary + {n - n_primary} bounce)")

    dev = torch.device("cuda:0")
    d_o = torch.from_numpy(ro).to(dev)
    d_d = torch.from_numpy(rd).to(dev)
    d_p = torch.empty(n, dtype=torch.int32, device=dev)
    d_t = torch.empty(n, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    counts = scene.trace_counts_device(d_o.data_ptr(), d_d.data_ptr(), n, d_p.data_ptr(), d_t.data_ptr(), **tp)
    alg_bytes = 32 * n + 8 * counts[0] + 56 * counts[1] + 40 * counts[2]
    # the same walks in the reference's layout: a 128-byte node per node visit
    # (inner or leaf), 52 bytes per triangle test (index + 3 float4 fetches)
    ref_layout_bytes = 32 * n + 128 * (counts[0] + counts[1]) + 52 * counts[2]

    for _ in range(args.warmup):
        scene.trace_closest_device(d_o.data_ptr(), d_d.data_ptr(), n, d_p.data_ptr(), d_t.data_ptr(), sync=False,
                                   **tp)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        scene.trace_closest_device(d_o.data_ptr(), d_d.data_ptr(), n, d_p.data_ptr(), d_t.data_ptr(), sync=False,
                                   **tp)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ktimes = ctx.kernel_times(args.steps)
    kms = float(np.mean(ktimes))
    achieved = alg_bytes / (kms * 1e-3) / 1e9

    # HBM bytes per launch of the traversal kernel, from the FETCH_SIZE and
    # WRITE_SIZE counter runs recorded in the profile summary (calibrated
    # against a copy kernel of known size).
    traffic = None
    tfile = ROOT / "profiles" / "trace_closest_cluster_traffic.md"
    if tfile.exists():
        m = re_mod.search(r"hbm_bytes_per_launch:\s*([0-9]+)", tfile.read_text())
        traffic = int(m.group(1)) if m else None

    value = n * args.steps / wall / 1e6
    res = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"cluster scene: 2^20-triangle Gaussian-cluster soup (64 clusters, seed 1), "
                               f"SAH rope kd-tree leaf<={tune['max_leaf_tris']} depth<=24, 1920x1080 primary rays "
                               f"({tune['tile']}x{tune['tile']} pixel tiles) + 1 diffuse bounce per hit, closest hit, "
                               f"leaf-visit cap 50",
                   "launch": {"block": tune["block"], "variant": tune["variant"], "autotuned": tuned is not None},
                   "rays": n, "primary_rays": n_primary, "bounce_rays": n - n_primary,
                   "triangles": int(tri_v.shape[0]), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "trace_closest_o6_kernel" if tune["variant"] == 3 else
                               ("trace_persistent_kernel" if tune["variant"] == 1 else "trace_closest_kernel"),
                     "kernel_ms": round(kms, 4),
                     "algorithmic_bytes": int(alg_bytes),
                     "reference_layout_bytes": int(ref_layout_bytes),
                     "reference_layout_gbs": round(ref_layout_bytes / (kms * 1e-3) / 1e9, 1),
                     "work": {"node_entries": counts[0], "leaves": counts[1], "tri_tests": counts[2],
                              "rays_hit": counts[3]}},
    }
    # the same call with rays and results in pageable host memory (copied
    # through the context's staging buffer): the PCIe-inclusive rate of the
    # host-pointer boundary, reported beside `value`, never as it
    scene.trace_closest(ro, rd, **tp)
    hw = []
    for _ in range(3):
        th = time.perf_counter()
        scene.trace_closest(ro, rd, **tp)
        hw.append(time.perf_counter() - th)
    res["host_pointer_call"] = {"ms": round(float(np.median(hw)) * 1e3, 3),
                                "mrays_s": round(n / float(np.median(hw)) / 1e6, 1),
                                "bytes_over_pcie": int(32 * n)}
    res["kd_build"] = kd_gpu
    res["shadow"] = shadow_section(W, torch, ctx, scene, tri_v, po, pd, pp, pt, order)
    if not args.no_mixed:
        try:
            res["mixed_workload"] = mixed_workload_section(re, W, torch, ctx)
        except Exception as e:  # the closest-hit line is still reported
            res["mixed_workload"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_terrain:
        try:
            res["terrain"] = terrain_section(re, W, torch, ctx)
        except Exception as e:  # the closest-hit line is still reported
            res["terrain"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        res["cpu_baseline"] = cpu_baseline(tree, ro, rd, n_primary, threads)
        one = cpu_baseline(tree, ro, rd, n_primary, 1, budget_cpu_s=8.0)
        res["cpu_baseline"]["one_core"] = {"value": one["value"], "sample": one["sample"]}
        res["cpu_baseline"]["cpu_model"] = _cpu_model()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
