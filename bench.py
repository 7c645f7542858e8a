"""Benchmark of the rope kd-tree closest-hit path on one MI355X.

Workload (the cluster scene): 2^20 FP32 triangles in 64 Gaussian clusters
(sigma 0.3, centres U[-10,10]^3, edge lengths LogNormal(-3, 0.5), seed 1), SAH
kd-tree with ropes (leaf <= 8, depth <= 24); 1920x1080 primary rays from
(0,0,30) towards the origin (FOV 60) plus one cosine-weighted diffuse bounce
ray per primary hit.  One step = one closest-hit launch over every ray, with
rays and results resident in HBM.

Prints ONE JSON line.  `value` is rays traced per second (Mrays/s) over the
timed steps.  `roofline` uses the algorithmic bytes of a launch, from the
kernel's own work counters: per ray 24 B in + 8 B out, 8 B per node entry,
56 B per leaf record, 40 B per triangle record; divided by the mean kernel
duration from HIP events recorded on the launch stream during the timed
steps.  `cpu_baseline` times the CPU oracle on a bounded sample of the same
rays.
"""
from __future__ import annotations

import argparse
import json
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


def cpu_baseline(tree, ro, rd, n_primary, threads, budget_s=12.0):
    """Oracle traversal on a stratified sample (primary and bounce rays in the
    workload's proportion), sized to about budget_s of CPU time."""
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
    m = int(min(n, max(20000, rate * budget_s)))
    mb = int(round(m * frac_b))
    sel = np.concatenate([rng.choice(n_primary, m - mb, replace=False),
                          n_primary + rng.choice(n - n_primary, mb, replace=False)])
    so, sd = np.ascontiguousarray(ro[sel]), np.ascontiguousarray(rd[sel])
    t0 = time.perf_counter()
    oracle.trace(nodes, leaves, tris, obj, so, sd, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(m / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{m} rays ({m - mb} primary + {mb} bounce, uniform sample of the workload), "
                      f"CPU oracle (FP32 restatement of the reference traversal), OpenMP dynamic/64, "
                      f"{dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(rays shard across GPUs with no exchange, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    import rayengine_amd as re
    from rayengine_amd import workloads as W

    t0 = time.time()
    tri_v = W.cluster_soup_default()
    tree = re.KDTree(tri_v)
    log(f"scene: {tri_v.shape[0]} tris, tree {tree.info['n_leaves']} leaves, "
        f"{tree.info['n_trirecs']} tri records, build {time.time() - t0:.1f} s")
    ctx = re.Context(0)
    scene = ctx.upload_scene([re.SceneObject(tree)])
    po, pd = re.camera_rays(W.CLUSTER_CAMERA, 1920, 1080)
    pp, pt = scene.trace_closest(po, pd)
    ro, rd = W.cluster_rays(tri_v, pp, pt)
    n = len(ro)
    n_primary = len(po)
    log(f"rays: {n} ({n_primary} primary + {n - n_primary} bounce)")

    dev = torch.device("cuda:0")
    d_o = torch.from_numpy(ro).to(dev)
    d_d = torch.from_numpy(rd).to(dev)
    d_p = torch.empty(n, dtype=torch.int32, device=dev)
    d_t = torch.empty(n, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    counts = scene.trace_counts_device(d_o.data_ptr(), d_d.data_ptr(), n, d_p.data_ptr(), d_t.data_ptr())
    alg_bytes = 32 * n + 8 * counts[0] + 56 * counts[1] + 40 * counts[2]

    for _ in range(args.warmup):
        scene.trace_closest_device(d_o.data_ptr(), d_d.data_ptr(), n, d_p.data_ptr(), d_t.data_ptr(), sync=False)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        scene.trace_closest_device(d_o.data_ptr(), d_d.data_ptr(), n, d_p.data_ptr(), d_t.data_ptr(), sync=False)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ktimes = ctx.kernel_times(args.steps)
    kms = float(np.mean(ktimes))
    achieved = alg_bytes / (kms * 1e-3) / 1e9

    traffic = None
    tfile = ROOT / "profiles" / "trace_closest_cluster_traffic.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

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
        "config": {"workload": "cluster scene: 2^20-triangle Gaussian-cluster soup (64 clusters, seed 1), "
                               "SAH rope kd-tree leaf<=8 depth<=24, 1920x1080 primary rays + 1 diffuse bounce "
                               "per hit, closest hit, leaf-visit cap 50",
                   "rays": n, "primary_rays": n_primary, "bounce_rays": n - n_primary,
                   "triangles": int(tri_v.shape[0]), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "trace_closest_kernel", "kernel_ms": round(kms, 4),
                     "algorithmic_bytes": int(alg_bytes),
                     "work": {"node_entries": counts[0], "leaves": counts[1], "tri_tests": counts[2],
                              "rays_hit": counts[3]}},
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        res["cpu_baseline"] = cpu_baseline(tree, ro, rd, n_primary, threads)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
