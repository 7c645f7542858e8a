"""Benchmark of the RayEngine ray-query path on one MI355X.

Headline workload (the cluster scene): 2^20 FP32 triangles in 64 Gaussian
clusters (sigma 0.3, centres U[-10,10]^3, edge lengths LogNormal(-3, 0.5),
seed 1), SAH kd-tree with ropes (leaf <= 8, depth <= 24); 1920x1080 primary
rays from a camera 0.8 units outside the cluster nearest the origin, looking
at its centre (FOV 60; 95 % of the primaries hit), issued in 8x8 pixel tiles,
plus one cosine-weighted diffuse bounce ray per primary hit, in the order of
their primaries: 4,038,016 rays, 1,964,416 of them bounces.  One step = the
closest-hit launch over the primaries (kernel variant 4, the cooperative
walk), then the launch over their bounce rays (variant 5, the cooperative walk
with lane refill), back to back on one stream, as a renderer issues them; rays
and results resident in HBM.

Prints ONE JSON line.  `value` is rays traced per second (Mrays/s) over the
timed steps.  `roofline` is the dominant kernel's, the bounce launch: its
algorithmic bytes, from the kernel's own work counters (per ray 24 B in + 8 B
out, 8 B per node entry, 56 B per leaf record, 40 B per triangle record),
divided by its mean duration from HIP events recorded on the launch stream
during the timed steps (`roofline.step` has the same for the whole step).
`roofline.traffic` is measured by this run: two rocprofv3 counter passes
(FETCH_SIZE, then WRITE_SIZE) over a child process that launches the same
kernel on the bounce rays beside a device copy of known size, which
calibrates the counters.  `single_launch` is every ray in one launch of
variant 4; `parts` times each ray set with variants 4 and 5.  `cpu_baseline` times the CPU
oracle (the reference's traversal restated in C++) on the box's host cores.
Further sections: the small icosphere scene (`small_scene`), any-hit queries
(`shadow`), the device kd build (`kd_build`), the many-scene mixed workload
unbatched and as one persistent work-queue launch in a captured hipGraph
(`mixed_workload`), and the terrain path (`terrain`: noise volume, sphere
tracers, surrogate training and inference, FP32-vs-FP64 depth error, CPU
baselines).
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
METRIC = "Mrays/s closest-hit; terrain Mrays/s surrogate vs exact SDF at ≤0.5-voxel err"
TRACE_KERNEL = {0: "trace_closest_kernel", 1: "trace_persistent_kernel", 2: "trace_closest_kernel",
                3: "trace_closest_o6_kernel", 4: "trace_coop_kernel", 5: "trace_coop_refill_kernel"}
DEFAULT_TUNE = {"max_leaf_tris": 8, "tile": 8, "block": 64, "variant": 4, "bounce_variant": 5}
COPY_PROBE_BYTES = 1 << 28
# rays of the headline workload at the default tree (leaf 8): 2,073,600
# primaries + 1,964,416 bounces; checked against the built workload below
CLUSTER_RAYS = 4038016


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# --------------------------------------------------------------------------- CPU side

def cpu_info():
    """What the CPU baselines ran on: the thread count used (the box's CPU
    share, OMP_NUM_THREADS when set), what nproc and the affinity mask show,
    and the lscpu model name."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    if model is None:
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    aff = len(os.sched_getaffinity(0))
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(aff, omp) if omp > 0 else aff
    return {"threads": threads, "nproc": os.cpu_count(), "affinity_cpus": aff,
            "omp_num_threads": omp or None, "model": model}


def time_cpu(fn, warmup=1, reps=5):
    """1 warm-up call, then the median wall time of `reps` calls (seconds)."""
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), [round(t, 4) for t in ts]


def cpu_rate(fn, units, unit, threads, what, kind="port", one_core_fn=None, one_core_units=None,
             one_core_what=None):
    """A cpu_baseline object: `units` per second of fn() (1 warm-up + median
    of 5) on `threads` threads, and the same on one core (optionally on a
    smaller sample)."""
    med, runs = time_cpu(lambda: fn(threads))
    rec = {"value": round(units / med / 1e6, 4), "unit": unit, "cores": threads, "kind": kind,
           "sample": what, "median_s": round(med, 4), "runs_s": runs, "protocol": "1 warm-up + median of 5"}
    f1 = one_core_fn or fn
    u1 = one_core_units or units
    med1, runs1 = time_cpu(lambda: f1(1))
    rec["one_core"] = {"value": round(u1 / med1 / 1e6, 4), "median_s": round(med1, 4), "runs_s": runs1,
                       "sample": one_core_what or what}
    return rec


# --------------------------------------------------------------------------- traffic (rocprofv3)

def _dispatch_counters(csv_dir: Path, counter: str):
    """[(dispatch id, kernel name, grid size, value)] of one counter from
    rocprofv3's counter_collection CSV, in dispatch order."""
    rows = {}
    for f in sorted(csv_dir.rglob("*counter_collection.csv")):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") != counter:
                    continue
                key = int(r["Dispatch_Id"])
                name = r.get("Kernel_Name", "")
                grid = int(float(r.get("Grid_Size", 0) or 0))
                rows.setdefault(key, [name, grid, 0.0])
                rows[key][2] += float(r["Counter_Value"])
    return [(k, rows[k][0], rows[k][1], rows[k][2]) for k in sorted(rows)]


def measure_traffic(tune, timeout_s=150):
    """HBM-side bytes per launch of the step's dominant kernel (the bounce
    rays' launch, with the bounce variant) on the headline workload, from two
    rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) over `bench.py
    --traffic-probe --probe-set bounce`, each calibrated by a device copy of
    known size in the same process.  Runs before this process touches the GPU
    (the probes are child processes).  Returns a dict, or one with "error"."""
    exe = subprocess.run(["which", "rocprofv3"], capture_output=True, text=True).stdout.strip()
    if not exe:
        return {"error": "rocprofv3 not found"}
    kname = TRACE_KERNEL[tune["bounce_variant"]]
    out = {"kernel": kname, "launch": "the bounce rays", "counters": {}}
    env = dict(os.environ, TMPDIR="/tmp")
    total = 0.0
    for counter, bytes_per_copy in (("FETCH_SIZE", COPY_PROBE_BYTES), ("WRITE_SIZE", COPY_PROBE_BYTES)):
        d = Path(tempfile.mkdtemp(prefix=f"re_pmc_{counter.lower()}_", dir="/tmp"))
        cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", str(d), "-o", "run", "--",
               sys.executable, str(ROOT / "bench.py"), "--traffic-probe", "--probe-leaf", str(tune["max_leaf_tris"]),
               "--probe-tile", str(tune["tile"]), "--probe-block", str(tune["block"]),
               "--probe-variant", str(tune["bounce_variant"]), "--probe-set", "bounce"]
        t0 = time.time()
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s, cwd="/tmp", env=env)
        except subprocess.TimeoutExpired:
            return {"error": f"{counter} pass timed out after {timeout_s} s"}
        rows = _dispatch_counters(d, counter)
        # the probe's only launches of this kernel are the bounce launches (its
        # workload set-up traces the primaries with the default kernel)
        trace = [v for _, n, g, v in rows if kname in n]
        copy = [v for _, n, g, v in rows if "copyBuffer" in n or ("copy" in n.lower() and kname not in n)]
        if len(trace) < 3 or len(copy) < 3:
            return {"error": f"{counter} pass: rc {r.returncode}, {len(trace)} trace and {len(copy)} copy dispatches "
                             f"found; stderr tail: {r.stderr[-300:]}"}
        trace, copy = trace[-3:], copy[-3:]            # the probe's measured launches
        factor = bytes_per_copy / float(np.mean(copy))  # bytes per counter unit
        b = float(np.mean(trace)) * factor
        total += b
        out["counters"][counter] = {"per_dispatch": [round(x, 1) for x in trace], "copy_per_dispatch":
                                    [round(x, 1) for x in copy], "copy_bytes": bytes_per_copy,
                                    "bytes_per_unit": round(factor, 4), "bytes_per_launch": int(b),
                                    "pass_s": round(time.time() - t0, 1)}
    out["bytes_per_launch"] = int(total)
    out["note"] = ("counts of traffic below the L2 (L2 misses, served by the Infinity Cache or HBM), so an upper "
                   "bound on HBM bytes; calibration: a device-to-device copy of known size in the same process")
    return out


def traffic_probe(args):
    """Child process of measure_traffic (and the program the committed counter
    profiles run): the headline workload; the closest-hit kernel launched 3
    times over every ray, 3 times over the primary rays alone and 3 times over
    the bounce rays alone, in that order (--probe-set each), or 3 times over one
    of those sets (all, primary, bounce); then 3 device copies of
    COPY_PROBE_BYTES."""
    import torch
    import rayengine_amd as re
    from rayengine_amd import workloads as W
    tri_v = W.cluster_soup_default()
    ctx = re.Context(0)
    tree = re.KDTree(tri_v, max_leaf_tris=args.probe_leaf, ctx=ctx)   # identical to the host build
    scene = ctx.upload_scene([re.SceneObject(tree)])
    ro, rd = W.cluster_workload_rays(scene, tri_v, args.probe_tile)
    dev = torch.device("cuda:0")
    n = len(ro)
    d_o, d_d = torch.from_numpy(ro).to(dev), torch.from_numpy(rd).to(dev)
    d_p = torch.empty(n, dtype=torch.int32, device=dev)
    d_t = torch.empty(n, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    n_primary = 1920 * 1080
    sets = {"all": (0, n), "primary": (0, n_primary), "bounce": (n_primary, n - n_primary)}
    chosen = list(sets.values()) if args.probe_set == "each" else [sets[args.probe_set]]
    for b, c in chosen:
        for _ in range(3):
            scene.trace_closest_device(d_o.data_ptr() + 12 * b, d_d.data_ptr() + 12 * b, c, d_p.data_ptr() + 4 * b,
                                       d_t.data_ptr() + 4 * b, block=args.probe_block, variant=args.probe_variant)
    src = torch.ones(COPY_PROBE_BYTES // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "ok", "rays": n}), flush=True)


def noise_probe(args):
    """The program the committed noise-fill counter profile runs: the 256^3
    terrain volume of the terrain section generated 5 times."""
    import torch  # noqa: F401  (one HIP runtime: torch's)
    import rayengine_amd as re
    ctx = re.Context(0)
    for _ in range(5):
        ctx.noise_volume(dim=(256, 256, 256), spacing=(1 / 256,) * 3, seed=2016).release()
    print(json.dumps({"probe": "ok", "noise_ms": round(ctx.last_kernel_ms(), 4)}), flush=True)


def train_probe(args):
    """The program the committed surrogate counter profile runs: the terrain
    section's network trained for 50 Adam steps at batch 65,536 on 2^22
    samples of the 256^3 terrain's distance field, then 3 forward passes over
    2^22 device-resident points."""
    import torch
    import rayengine_amd as re
    ctx = re.Context(0)
    vol = ctx.noise_volume()
    dv = re.distance_volume(vol)
    net = re.Surrogate(ctx, seed=1)
    net.train(dv, n_samples=1 << 22, batch=65536, steps=50, lr=1e-3, seed=2)
    n = 1 << 22
    pts = torch.rand((n, 3), dtype=torch.float32, device="cuda:0", generator=torch.Generator(device="cuda:0").manual_seed(3))
    out = torch.empty(n, dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    for _ in range(3):
        net.forward_device(pts.data_ptr(), n, out.data_ptr())
    print(json.dumps({"probe": "ok", "forward_ms": round(ctx.last_kernel_ms(), 4)}), flush=True)


# --------------------------------------------------------------------------- sections

def launch_stats(ctx, fn, steps):
    """Mean HIP-event kernel time (ms) of `steps` asynchronous launches."""
    import torch
    fn(True)
    for _ in range(steps):
        fn(False)
    torch.cuda.synchronize()
    return float(np.mean(ctx.kernel_times(steps)))


def part_record(n, counts, kms):
    alg = 32 * n + 8 * counts[0] + 56 * counts[1] + 40 * counts[2]
    gbs = alg / (kms * 1e-3) / 1e9
    return {"rays": int(n), "kernel_ms": round(kms, 4), "mrays_s": round(n / (kms * 1e-3) / 1e6, 1),
            "algorithmic_bytes": int(alg), "bytes_per_ray": round(alg / n, 1), "achieved_gbs": round(gbs, 1),
            "frac_of_hbm": round(gbs / HBM_PEAK_GBS, 4),
            "work": {"node_entries": counts[0], "leaves": counts[1], "tri_tests": counts[2], "rays_hit": counts[3]}}


def small_scene_section(re, W, torch, ctx, cpu, steps=20):
    """The small icosphere scene: 5,120 + 1,024 triangles, SAH leaf <= 8,
    depth <= 24, 256x256 primary rays from (0,0,5); GPU closest hit against
    the CPU oracle on the same rays."""
    import oracle
    tri_v, tree, o, d = W.small_scene()
    scene = ctx.upload_scene([re.SceneObject(tree)])
    n = len(o)
    dev = torch.device("cuda:0")
    d_o, d_d = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    d_p = torch.empty(n, dtype=torch.int32, device=dev)
    d_t = torch.empty(n, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    counts = scene.trace_counts_device(d_o.data_ptr(), d_d.data_ptr(), n, d_p.data_ptr(), d_t.data_ptr())
    kms = launch_stats(ctx, lambda s: scene.trace_closest_device(d_o.data_ptr(), d_d.data_ptr(), n, d_p.data_ptr(),
                                                                 d_t.data_ptr(), sync=s), steps)
    rec = part_record(n, counts, kms)
    nodes, leaves, tris = tree.export()
    obj = tree.identity_object()
    po, to = oracle.trace(nodes, leaves, tris, obj, o, d, threads=cpu["threads"])
    rec["bit_exact_vs_oracle"] = bool(np.array_equal(d_p.cpu().numpy(), po) and
                                      np.array_equal(d_t.cpu().numpy().view(np.uint32), to.view(np.uint32)))
    rec["workload"] = "icosphere (5,120 tris, 4x subdivided) + 1,024 random tris (seed 0), 256x256 rays from (0,0,5)"
    rec["cpu_baseline"] = cpu_rate(lambda th: oracle.trace(nodes, leaves, tris, obj, o, d, threads=th), n,
                                   "Mrays/s", cpu["threads"], f"all {n} rays, CPU oracle (full walks to the cap)")
    scene.release()
    return rec


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
        kms = launch_stats(ctx, fn, steps)
        out[name] = {"kernel_ms": round(kms, 4), "Mrays_per_s": round(n / (kms * 1e-3) / 1e6, 1)}
    occ = d_occ.cpu().numpy()
    out["occluded_frac"] = round(float(occ.mean()), 4)
    out["agrees_with_closest_hit"] = bool(np.array_equal(occ, (d_p.cpu().numpy() >= 0).astype(np.uint8)))
    return out


# Terrain settings (hit depth error is measured against REFERENCE_DEPTH, the
# fine exact-field march that tests/test_gpu_terrain.py pins to the FP64
# first crossing within 0.02 voxel at the 99th percentile).
REFERENCE_DEPTH = dict(k_exact=0.02, h_min_voxels=0.25, n_bisect=20, max_iters=1 << 16)
TERRAIN_METHODS = {
    # the reference sphere tracer (renderer.cu:142-235) at its own settings
    "reference_march": ("plain", dict(step_scale=0.8, eps=1e-3, max_steps=256)),
    # the same tracer at the smallest step and tolerance tried: its most accurate setting
    "reference_march_fine": ("plain", dict(step_scale=0.02, eps=1e-5, max_steps=8192)),
    # exact-field bracketing march (no network): the surrogate's ablation
    "exact_march": ("exact", dict(k_exact=0.05, h_min_voxels=1.0, n_bisect=1)),
    # surrogate far field + exact near field
    "surrogate_hybrid": ("hybrid", dict(tau_voxels=4.0, margin_voxels=1.0, k_exact=0.1, n_bisect=4)),
}
MLP_FLOP_PER_EVAL = 2 * (99 * 64 + 3 * 64 * 64 + 64)
# forward + backward (input and weight gradients) per training sample, about
# three forward passes
MLP_TRAIN_FLOP_PER_SAMPLE = 3 * MLP_FLOP_PER_EVAL
NOISE_FLOP_PER_VOXEL = 5 * 70
# Surrogate training: 2^22 samples (half within 2 voxels of the surface), Adam
# lr 1e-3, batch 65,536, 20k steps.
TRAIN_STEPS = 20000


def terrain_cpu(cpu, o, d, grid32, net_params):
    """CPU baselines of the terrain path: the host noise fill (oracle,
    voxels/s), the host reference sphere tracer, the host exact-field march
    and the host surrogate hybrid (the trained network in FP32) on the same
    rays (every k-th ray where the whole set would take minutes)."""
    import oracle
    perm, _ = oracle.perlin_table(2016)
    out = {}
    slab = dict(dim=(256, 256, 16))
    out["noise_fill"] = cpu_rate(lambda th: oracle.noise_fill(perm, threads=th), 256 ** 3, "Mvoxels/s",
                                 cpu["threads"], "the whole 256^3 volume, CPU oracle (FP32, 5 octaves)",
                                 one_core_fn=lambda th: oracle.noise_fill(perm, threads=th, **slab),
                                 one_core_units=256 * 256 * 16,
                                 one_core_what="z-slices 0-15 of the 256^3 volume (256x256x16 voxels)")
    params = TERRAIN_METHODS["reference_march"][1]
    n = len(o)
    out["reference_march"] = cpu_rate(
        lambda th: oracle.sdf_trace(grid32, (1 / 256,) * 3, o, d, threads=th, **params), n, "Mrays/s",
        cpu["threads"], f"all {n} terrain rays, CPU oracle sphere trace (step 0.8, eps 1e-3, <= 256 steps)")
    sp = (1 / 256,) * 3
    ex = oracle.hybrid_world_params(1 / 256, **_hybrid_voxel_params(TERRAIN_METHODS["exact_march"][1]))
    out["exact_march"] = cpu_rate(
        lambda th: oracle.exact_march(grid32, sp, o, d, ex["k_exact"], ex["h_min"], ex["n_bisect"], ex["max_iters"],
                                      threads=th), n, "Mrays/s", cpu["threads"],
        f"all {n} terrain rays, CPU oracle exact-field march", one_core_fn=lambda th: oracle.exact_march(
            grid32, sp, o[::8], d[::8], ex["k_exact"], ex["h_min"], ex["n_bisect"], ex["max_iters"], threads=th),
        one_core_units=len(o[::8]), one_core_what=f"every 8th terrain ray ({len(o[::8])} rays)")
    hy = oracle.hybrid_world_params(1 / 256, **_hybrid_voxel_params(TERRAIN_METHODS["surrogate_hybrid"][1]))
    os_, ds_ = o[::64], d[::64]
    o1, d1 = o[::1024], d[::1024]
    out["surrogate_hybrid"] = cpu_rate(
        lambda th: oracle.hybrid_trace(grid32, sp, net_params, os_, ds_, threads=th, **hy), len(os_), "Mrays/s",
        cpu["threads"], f"every 64th terrain ray ({len(os_)} rays), CPU oracle hybrid, the trained network in FP32",
        one_core_fn=lambda th: oracle.hybrid_trace(grid32, sp, net_params, o1, d1, threads=th, **hy),
        one_core_units=len(o1), one_core_what=f"every 1024th terrain ray ({len(o1)} rays)")
    return out


def _hybrid_voxel_params(kw):
    """re_hybrid_params fields (voxel units) with the defaults filled in."""
    import rayengine_amd as re
    p = re.hybrid_params(**kw)
    return dict(tau_voxels=p.tau_voxels, margin_voxels=p.margin_voxels, k_exact=p.k_exact,
                h_min_voxels=p.h_min_voxels, rejoin_voxels=p.rejoin_voxels, n_bisect=p.n_bisect,
                max_iters=p.max_iters)


def terrain_fp64(o, d, grid32, t_gpu, h_gpu, cpu):
    """The FP32 reference sphere trace (GPU) against its FP64 twin on the FP64
    noise grid (CPU oracle), on every terrain ray."""
    import oracle
    perm, _ = oracle.perlin_table(2016)
    g64 = oracle.noise_fill_f64(perm, threads=cpu["threads"])
    params = TERRAIN_METHODS["reference_march"][1]
    t64, h64, _ = oracle.sdf_trace_f64(g64, (1 / 256,) * 3, o, d, threads=cpu["threads"], **params)
    both = (h64 == 1) & (h_gpu == 1)
    e = np.abs(t_gpu[both].astype(np.float64) - t64[both])
    gmax = float(np.abs(grid32.astype(np.float64) - g64).max())
    return {"tracer": "reference sphere trace (step 0.8, eps 1e-3, <= 256 steps)", "rays_both_hit": int(both.sum()),
            "hit_agreement": round(float(np.mean(h64 == h_gpu)), 6),
            "abs_t_err_p50": float(np.median(e)), "abs_t_err_p99": float(np.percentile(e, 99)),
            "abs_t_err_max": float(e.max()), "err_p99_voxels": round(float(np.percentile(e, 99)) * 256, 5),
            "frac_within_1e-4": round(float(np.mean(e <= 1e-4)), 6), "grid_max_abs_err": gmax}


def terrain_section(re, W, torch, ctx, cpu, with_cpu=True):
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
    nsec = float(np.mean(noise)) * 1e-3
    out["noise_roofline"] = {"bound": "valu", "flop_per_voxel": NOISE_FLOP_PER_VOXEL,
                             "achieved_tflops": round(256 ** 3 * NOISE_FLOP_PER_VOXEL / nsec / 1e12, 2),
                             "peak_tflops": FP32_PEAK_TFLOPS,
                             "frac": round(256 ** 3 * NOISE_FLOP_PER_VOXEL / nsec / 1e12 / FP32_PEAK_TFLOPS, 4),
                             "write_gbs": round(256 ** 3 * 4 / nsec / 1e9, 1)}
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
    ref_gpu = None
    for name, (kind, params) in TERRAIN_METHODS.items():
        run(kind, params)
        t, h = dt.cpu().numpy().astype(np.float64), dh.cpu().numpy()
        if name == "reference_march":
            ref_gpu = (dt.cpu().numpy().copy(), h.copy())
        both = (h == 1) & (ref_h == 1)
        e = np.abs(t[both] - ref_t[both]) / vox
        for _ in range(5):
            run(kind, params, sync=False)
        torch.cuda.synchronize()
        ms = float(np.mean(ctx.kernel_times(5)))
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
            roof["mfma_tflops"] = round(float(np.sum(evals, dtype=np.int64)) * MLP_FLOP_PER_EVAL / (ms * 1e-3) / 1e12,
                                        1)
        methods[name] = {"mrays_s": round(n / (ms * 1e-3) / 1e6, 1), "kernel_ms": round(ms, 4), "roofline": roof,
                         "err_p50_vox": round(float(np.median(e)), 4),
                         "err_p99_vox": round(float(np.percentile(e, 99)), 4),
                         "err_max_vox": round(float(e.max()), 3) if e.size else None,
                         "frac_err_gt_half_vox": round(float(np.mean(e > 0.5)), 5),
                         "hit_agree": round(float(np.mean(h == ref_h)), 5), "params": params}
    out["methods"] = methods
    hy, ex = methods["surrogate_hybrid"], methods["exact_march"]
    # C3's ratio: the surrogate hybrid against its ablation, the exact-field
    # march without the network (both within 0.5 voxel at p99)
    out["surrogate_vs_exact_march"] = round(hy["mrays_s"] / ex["mrays_s"], 3)
    out["surrogate_vs_exact_march_note"] = (
        "the C3 ratio: surrogate hybrid Mrays/s / exact-field march Mrays/s (both meet p99 <= 0.5 voxel); the "
        "reference sphere tracer at its own settings misses the accuracy bar (see methods.reference_march)")
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
    if with_cpu:
        grid32 = vol.read()
        try:
            out["fp32_vs_fp64"] = terrain_fp64(o, d, grid32, ref_gpu[0], ref_gpu[1], cpu)
        except Exception as e:
            out["fp32_vs_fp64"] = {"error": f"{type(e).__name__}: {e}"}
        try:
            out["cpu_baseline"] = terrain_cpu(cpu, o, d, grid32, net.params()[0])
        except Exception as e:
            out["cpu_baseline"] = {"error": f"{type(e).__name__}: {e}"}
    for x in (net, dv, vol):
        x.release()
    return out


def mixed_cpu(cpu, trees, vols, tiles, ro, rd, sr, tr, stride=32, one_core_stride=512):
    """CPU baseline of the mixed workload: the scenes and tiles one after
    another, each on all threads (CPU oracle traversal for scenes, exact-field
    march for even tiles, the hybrid with the tile's network in FP32 for odd
    tiles), on every stride-th ray of every item."""
    import oracle
    from rayengine_amd.mixed import TILE_HYBRID, TILE_MARCH
    exports = [(t.export(), t.identity_object()) for t in trees]
    grids = [v.read() for v in vols]
    hy = oracle.hybrid_world_params(1.0, **_hybrid_voxel_params(TILE_HYBRID))
    ex = oracle.hybrid_world_params(1.0, **_hybrid_voxel_params(TILE_MARCH))

    def world(P, dm):
        vox = np.float32(1.0 / dm)
        return {k: (float(np.float32(v) * vox) if k in ("tau", "margin", "h_min", "rejoin") else v)
                for k, v in P.items()}

    def run(th, step):
        n = 0
        for (arrays, obj), (b, c) in zip(exports, sr):
            o, d = ro[b:b + c:step], rd[b:b + c:step]
            oracle.trace(*arrays, obj, o, d, threads=th)
            n += len(o)
        for k, (g, (b, c)) in enumerate(zip(grids, tr)):
            dm = g.shape[0]
            o, d = ro[b:b + c:step], rd[b:b + c:step]
            sp = (1.0 / dm,) * 3
            if k % 2 == 0:
                P = world(ex, dm)
                oracle.exact_march(g, sp, o, d, P["k_exact"], P["h_min"], P["n_bisect"], P["max_iters"], threads=th)
            else:
                oracle.hybrid_trace(g, sp, tiles[k][2], o, d, threads=th, **world(hy, dm))
            n += len(o)
        return n

    n_s = run(cpu["threads"], stride)
    n_1 = sum(len(range(0, c, one_core_stride)) for _, c in list(sr) + list(tr))
    return cpu_rate(lambda th: run(th, stride), n_s, "Mrays/s", cpu["threads"],
                    f"every {stride}th ray of every scene and tile ({n_s} rays), items one after another",
                    one_core_fn=lambda th: run(th, one_core_stride), one_core_units=n_1,
                    one_core_what=f"every {one_core_stride}th ray of every scene and tile ({n_1} rays)")


def mixed_workload_section(re, W, torch, ctx, cpu=None, reps=3, n_streams=4):
    """The many-scene mixed workload (workloads.batch_workload, seed 7: 64
    triangle scenes of 1k-500k triangles with 16k-2M incoherent rays each, 16
    noise-terrain tiles of 64^3-256^3 with their own rays; even tiles traced
    by the exact-field march, odd tiles by the surrogate hybrid with a
    network trained on that tile), every input resident in HBM, run
    (a) as one persistent work-queue launch in a captured hipGraph
    (re_batch_*: scenes and exact tiles in the queue, surrogate tiles as
    hybrid launches in the same graph), leaf size, block and rays per work
    item picked by autotune.autotune; (b) unbatched, one launch per scene or
    tile in ascending ray count, on one stream and dealt round-robin over
    n_streams contexts (rayengine_amd.mixed), leaf size and block picked by
    autotune.tune_mixed.  Aggregate Mrays/s over the wall time; completion
    times per scene and per tile."""
    from rayengine_amd.autotune import autotune, batch_time_ms, tune_mixed
    from rayengine_amd.mixed import MixedRun, TILE_HYBRID, TILE_MARCH
    t0 = time.perf_counter()
    soups, dims, seeds, ro, rd, sr, tr = W.batch_workload()
    t_net = time.perf_counter()
    tiles, nets, vols = [], [], []
    for k, (dm, sd) in enumerate(zip(dims, seeds)):
        v = ctx.noise_volume(dim=(dm, dm, dm), spacing=(1.0 / dm,) * 3, seed=sd)
        vols.append(v)
        if k % 2 == 1:
            dv = re.distance_volume(v)
            net = re.Surrogate(ctx, seed=1)
            net.train(dv, n_samples=1 << 20, batch=65536, steps=2000, lr=1e-3, seed=1, over_weight=4.0)
            tiles.append((dm, sd, net.params()[0]))
            nets.append(net)
            dv.release()
        else:
            tiles.append((dm, sd))
            nets.append(None)
    net_s = time.perf_counter() - t_net
    n = len(ro)
    n_scene_rays = int(sum(c for _, c in sr))
    out = {"workload": "64 cluster-soup scenes (1k-500k triangles, log-uniform) with 16k-2M incoherent rays each "
                       "(log-uniform), 16 noise-terrain tiles 64^3-256^3 (seeds 2016+k) with 16k-2M rays each; "
                       "seed 7",
           "triangles": int(sum(len(s) for s in soups)), "rays": int(n), "scene_rays": n_scene_rays,
           "tile_rays": int(n - n_scene_rays),
           "tiles_traced_with": "even tiles: exact-field march (k 0.05, h_min 1 voxel); odd tiles: surrogate hybrid "
                                "(tau 4 voxels, margin 1, k 0.1, 4 bisections) with a network per tile, 2,000 Adam "
                                "steps on 2^20 samples of its field",
           "tile_networks_train_s": round(net_s, 2)}
    # (a) the batched persistent launch in one captured graph
    btiles = [(v, 0 if k % 2 == 0 else 2, b, c) + ((nets[k],) if nets[k] is not None else ())
              for k, (v, (b, c)) in enumerate(zip(vols, tr))]
    t_tune = time.perf_counter()
    bcfg, bmeas = autotune(ctx, soups, btiles, ro, rd, sr, device_name="mi355x", cache=None, log=log,
                           hybrid=TILE_HYBRID, march=TILE_MARCH)
    trees = [re.KDTree(s, max_leaf_tris=bcfg["max_leaf_tris"], ctx=ctx) for s in soups]
    batch = re.Batch(ctx, [(t, b, c) for t, (b, c) in zip(trees, sr)], btiles, ro, rd, item_rays=bcfg["item_rays"],
                     block=bcfg["block"], hybrid=TILE_HYBRID, march=TILE_MARCH,
                     side_share=bcfg.get("side_share", 0.0))
    replay_ms = batch_time_ms(ctx, batch, 5)
    walls = []
    for _ in range(reps):
        torch.cuda.synchronize()
        ts = time.perf_counter()
        batch.run()
        walls.append(time.perf_counter() - ts)
    bwall = float(np.median(walls))
    comp = batch.completion_ms()
    bp, bt = batch.results()
    info = {"items": batch.items(), "graph_replay_ms": round(replay_ms, 3), "wall_ms": round(bwall * 1e3, 3),
            "mrays_s": round(n / (replay_ms * 1e-3) / 1e6, 1),
            "mrays_s_wall": round(n / bwall / 1e6, 1),
            "scene_completion_ms": {"p50": round(float(np.percentile(comp[:len(sr)], 50)), 3),
                                    "p99": round(float(np.percentile(comp[:len(sr)], 99)), 3)},
            "tile_completion_ms": {"p50": round(float(np.percentile(comp[len(sr):], 50)), 3),
                                   "p99": round(float(np.percentile(comp[len(sr):], 99)), 3)},
            "autotune": {"chosen": {k: bcfg.get(k) for k in ("max_leaf_tris", "block", "item_rays", "side_share")},
                         "seconds": round(time.perf_counter() - t_tune, 1),
                         "measurements": bmeas}}
    out["batched_graph"] = info
    batch.release()
    del trees
    # (b) unbatched, one launch per item
    ctxs = [ctx] + [re.Context(0) for _ in range(n_streams - 1)]
    dev = torch.device("cuda:0")
    d_o, d_d = torch.from_numpy(ro).to(dev), torch.from_numpy(rd).to(dev)
    d_p = torch.empty(n, dtype=torch.int32, device=dev)
    d_t = torch.empty(n, dtype=torch.float32, device=dev)
    d_h = torch.empty(n, dtype=torch.uint8, device=dev)
    ptrs = (d_o.data_ptr(), d_d.data_ptr(), d_p.data_ptr(), d_t.data_ptr(), d_h.data_ptr())
    cfg, meas = tune_mixed(ctxs, soups, tiles, sr, tr, ptrs, torch.cuda.synchronize, variants=(4, 5), log=log)
    utrees = [re.KDTree(s, max_leaf_tris=cfg["max_leaf_tris"], ctx=ctx) for s in soups]
    out["unbatched_autotune"] = {"chosen": {"max_leaf_tris": cfg["max_leaf_tris"], "block": cfg["block"],
                                            "variant": cfg["variant"]},
                                 "issue": f"{n_streams} streams",
                                 "wall_ms": {f"leaf{m['max_leaf_tris']}_block{m['block']}_variant{m['variant']}":
                                             m["wall_ms"] for m in meas}}
    first = None
    for streams in (1, n_streams):
        run = MixedRun(ctxs[:streams], utrees, tiles, sr, tr, *ptrs)
        run.run(block=cfg["block"], variant=cfg["variant"])
        torch.cuda.synchronize()
        walls, comps = [], []
        for _ in range(reps):
            torch.cuda.synchronize()
            ts = time.perf_counter()
            run.run(block=cfg["block"], variant=cfg["variant"])
            torch.cuda.synchronize()
            walls.append(time.perf_counter() - ts)
            comps.append(run.completion_ms())
        wall = float(np.median(walls))
        done = np.mean(comps, axis=0)
        sd_ = np.array([done[i] for i, it in enumerate(run.items) if it[0] == "scene"])
        td = np.array([done[i] for i, it in enumerate(run.items) if it[0] == "tile"])
        rec = {"wall_ms": round(wall * 1e3, 3), "mrays_s": round(n / wall / 1e6, 1),
               "scene_completion_ms": {"p50": round(float(np.percentile(sd_, 50)), 3),
                                       "p99": round(float(np.percentile(sd_, 99)), 3)},
               "tile_completion_ms": {"p50": round(float(np.percentile(td, 50)), 3),
                                      "p99": round(float(np.percentile(td, 99)), 3)}}
        got = (d_p.cpu().numpy(), d_t.cpu().numpy().view(np.uint32), d_h.cpu().numpy())
        if first is None:
            first = got
            # the batched results equal the unbatched ones ray for ray
            hit_u = np.where(np.arange(n) < n_scene_rays, first[0], np.where(first[2] == 1, 0, -1))
            out["batched_graph"]["identical_to_unbatched"] = bool(
                np.array_equal(bp, hit_u) and np.array_equal(bt.view(np.uint32), first[1]))
        else:
            rec["identical_to_one_stream"] = all(np.array_equal(x, y) for x, y in zip(first, got))
        out[f"unbatched_streams_{streams}"] = rec
        run.release()
    for c in ctxs[1:]:
        c.close()
    if cpu is not None:
        try:
            out["cpu_baseline"] = mixed_cpu(cpu, utrees, vols, tiles, ro, rd, sr, tr)
        except Exception as e:
            out["cpu_baseline"] = {"error": f"{type(e).__name__}: {e}"}
    for x in nets + vols:
        if x is not None:
            x.release()
    out["prep_and_tune_s"] = round(time.perf_counter() - t0, 1)
    return out


# --------------------------------------------------------------------------- main

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-terrain", action="store_true", help="skip the terrain section")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-workload section")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--autotune", action="store_true",
                    help="pick leaf size, ray tile, block and variant by measurement before timing")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--noise-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--train-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--probe-leaf", type=int, default=8, help=argparse.SUPPRESS)
    ap.add_argument("--probe-tile", type=int, default=8, help=argparse.SUPPRESS)
    ap.add_argument("--probe-block", type=int, default=64, help=argparse.SUPPRESS)
    ap.add_argument("--probe-variant", type=int, default=3, help=argparse.SUPPRESS)
    ap.add_argument("--probe-set", default="each", choices=("each", "all", "primary", "bounce"), help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(args)
        return
    if args.noise_probe:
        noise_probe(args)
        return
    if args.train_probe:
        train_probe(args)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(rays would shard across GPUs with no exchange between them)", file=sys.stderr)
        sys.exit(2)

    tune = dict(DEFAULT_TUNE)
    # counter passes first: child processes, before this process uses the GPU
    traffic = None
    if not args.no_traffic and not args.autotune:
        log("traffic: two rocprofv3 counter passes over the headline launch")
        traffic = measure_traffic(tune)
        log(f"traffic: {traffic.get('bytes_per_launch', traffic.get('error'))}")

    import torch
    import rayengine_amd as re
    from rayengine_amd import workloads as W

    cpu = cpu_info()
    tri_v = W.cluster_soup_default()
    ctx = re.Context(0)
    # Leaf size, ray-sort granularity (primary rays in tile x tile pixel
    # blocks, bounces following their primaries), block and variant; the
    # defaults are the autotuner's choice class on this workload, --autotune
    # measures them again.
    tuned = None
    if args.autotune:
        from rayengine_amd.autotune import tune_trace
        tuned, _ = tune_trace(ctx, tri_v, lambda sc, tile: W.cluster_workload_rays(sc, tri_v, tile), 1920, 1080,
                              device_name=torch.cuda.get_device_name(0), cache=None, log=log)
        tune = {k: tuned.get(k, tune[k]) for k in tune}
    t0 = time.time()
    tree = re.KDTree(tri_v, max_leaf_tris=tune["max_leaf_tris"])
    host_build_s = time.time() - t0
    log(f"scene: {tri_v.shape[0]} tris, tree {tree.info['n_leaves']} leaves, "
        f"{tree.info['n_trirecs']} tri records, build {host_build_s:.1f} s")
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
    if traffic is not None and n != CLUSTER_RAYS:
        traffic = {"error": f"the counter passes assumed {CLUSTER_RAYS} rays; this tree gives {n}"}

    dev = torch.device("cuda:0")
    d_o = torch.from_numpy(ro).to(dev)
    d_d = torch.from_numpy(rd).to(dev)
    d_p = torch.empty(n, dtype=torch.int32, device=dev)
    d_t = torch.empty(n, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    o_ptr, d_ptr, p_ptr, t_ptr = d_o.data_ptr(), d_d.data_ptr(), d_p.data_ptr(), d_t.data_ptr()
    n_bounce = n - n_primary
    prim_args = (o_ptr, d_ptr, n_primary, p_ptr, t_ptr)
    bnc_args = (o_ptr + 12 * n_primary, d_ptr + 12 * n_primary, n_bounce, p_ptr + 4 * n_primary, t_ptr + 4 * n_primary)
    if tuned is not None:
        # the kernel variant of each launch, measured on its own ray set
        for key, args_ in (("variant", prim_args), ("bounce_variant", bnc_args)):
            by_v = {v: launch_stats(ctx, lambda s: scene.trace_closest_device(*args_, sync=s, block=tune["block"],
                                                                              variant=v), 10) for v in (4, 5)}
            tune[key] = min(by_v, key=by_v.get)
        tp = {"block": tune["block"], "variant": tune["variant"]}
    tpb = {"block": tune["block"], "variant": tune["bounce_variant"]}
    counts_p = scene.trace_counts_device(*prim_args, **tp)
    counts_b = scene.trace_counts_device(*bnc_args, **tpb)
    counts = [a + b for a, b in zip(counts_p, counts_b)]
    alg_p = 32 * n_primary + 8 * counts_p[0] + 56 * counts_p[1] + 40 * counts_p[2]
    alg_b = 32 * n_bounce + 8 * counts_b[0] + 56 * counts_b[1] + 40 * counts_b[2]
    alg_bytes = alg_p + alg_b
    ref_layout_bytes = 32 * n + 128 * (counts[0] + counts[1]) + 52 * counts[2]

    # One step: the primaries, then their bounce rays, as a renderer issues
    # them (the bounces are made from the primaries' hits), each launch with
    # its kernel variant, back to back on the context's stream.
    def step():
        scene.trace_closest_device(*prim_args, sync=False, **tp)
        scene.trace_closest_device(*bnc_args, sync=False, **tpb)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ktimes = np.asarray(ctx.kernel_times(2 * args.steps))
    kms_p, kms_b = float(np.mean(ktimes[0::2])), float(np.mean(ktimes[1::2]))
    kms = kms_b                                      # the dominant kernel: the bounce launch
    achieved = alg_b / (kms * 1e-3) / 1e9
    res_p, res_t = d_p.cpu().numpy(), d_t.cpu().numpy()

    # the same rays in one launch over every ray, and each ray set with both
    # cooperative variants (4 and 5)
    single_ms = launch_stats(ctx, lambda s: scene.trace_closest_device(o_ptr, d_ptr, n, p_ptr, t_ptr, sync=s, **tp),
                             10)
    single = {"variant": int(tune["variant"]), "kernel_ms": round(single_ms, 4),
              "mrays_s": round(n / (single_ms * 1e-3) / 1e6, 1), "algorithmic_bytes": int(alg_bytes),
              "achieved_gbs": round(alg_bytes / (single_ms * 1e-3) / 1e9, 1),
              "frac_of_hbm": round(alg_bytes / (single_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    same_single = bool(np.array_equal(d_p.cpu().numpy(), res_p) and
                       np.array_equal(d_t.cpu().numpy().view(np.uint32), res_t.view(np.uint32)))
    parts = {}
    for name, args_, pc in (("primary", prim_args, counts_p), ("bounce", bnc_args, counts_b)):
        by_var = {}
        for v in (4, 5):
            tpv = dict(tp, variant=v)
            by_var[v] = launch_stats(ctx, lambda s: scene.trace_closest_device(*args_, sync=s, **tpv), 10)
        used = tune["variant"] if name == "primary" else tune["bounce_variant"]
        parts[name] = part_record(args_[2], pc, by_var[used])
        parts[name]["variant"] = int(used)
        parts[name]["kernel_ms_by_variant"] = {str(v): round(ms, 4) for v, ms in by_var.items()}
    parts["primary"]["kernel_ms_in_step"] = round(kms_p, 4)
    parts["bounce"]["kernel_ms_in_step"] = round(kms_b, 4)
    parts["bounce"]["note"] = ("the incoherent part: one cosine-weighted diffuse bounce per primary hit, from the "
                               "hit point into the hemisphere about the facing geometric normal")
    parts["results_equal_single_launch"] = same_single

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
                               f"({tune['tile']}x{tune['tile']} pixel tiles) from 0.8 units outside the cluster "
                               f"nearest the origin + 1 diffuse bounce per hit, closest hit, leaf-visit cap 50; "
                               f"one step = the primaries' launch, then the bounces' launch",
                   "camera": {"pos": list(W.CLUSTER_CAMERA.pos), "view": list(W.CLUSTER_CAMERA.view), "fov": [60, 60]},
                   "launch": {"block": tune["block"], "variant": tune["variant"],
                              "bounce_variant": tune["bounce_variant"], "autotuned": tuned is not None},
                   "rays": n, "primary_rays": n_primary, "bounce_rays": n_bounce,
                   "triangles": int(tri_v.shape[0]), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic.get("bytes_per_launch") if traffic else None,
                     "kernel": TRACE_KERNEL[tune["bounce_variant"]],
                     "launch": "the bounce rays (the step's dominant kernel)",
                     "kernel_ms": round(kms, 4),
                     "algorithmic_bytes": int(alg_b),
                     "work": {"node_entries": counts_b[0], "leaves": counts_b[1], "tri_tests": counts_b[2],
                              "rays_hit": counts_b[3]},
                     "step": {"kernel_ms": round(kms_p + kms_b, 4), "algorithmic_bytes": int(alg_bytes),
                              "achieved_gbs": round(alg_bytes / ((kms_p + kms_b) * 1e-3) / 1e9, 1),
                              "reference_layout_bytes": int(ref_layout_bytes),
                              "work": {"node_entries": counts[0], "leaves": counts[1], "tri_tests": counts[2],
                                       "rays_hit": counts[3]}}},
        "single_launch": single,
        "parts": parts,
    }
    if traffic is not None:
        res["traffic_measurement"] = traffic
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
    if not args.no_cpu:
        import oracle
        nodes, leaves, tris = tree.export()
        obj = tree.identity_object()
        log(f"cpu baseline: oracle on {cpu['threads']} threads and 1 core")
        res["cpu_baseline"] = cpu_rate(
            lambda th: oracle.trace(nodes, leaves, tris, obj, ro, rd, threads=th), n, "Mrays/s", cpu["threads"],
            f"the whole workload ({n} rays: {n_primary} primary + {n - n_primary} bounce), CPU oracle (FP32 "
            f"restatement of the reference traversal, full walks to the cap), OpenMP dynamic/64")
        res["cpu_baseline"]["cpu"] = cpu
        po_, to_ = oracle.trace(nodes, leaves, tris, obj, ro, rd, threads=cpu["threads"])
        res["bit_exact_vs_oracle_all_rays"] = bool(np.array_equal(po_, res_p) and
                                                   np.array_equal(to_.view(np.uint32), res_t.view(np.uint32)))
    try:
        res["small_scene"] = small_scene_section(re, W, torch, ctx, cpu)
    except Exception as e:
        res["small_scene"] = {"error": f"{type(e).__name__}: {e}"}
    scene.release()
    if not args.no_mixed:
        try:
            res["mixed_workload"] = mixed_workload_section(re, W, torch, ctx, cpu=None if args.no_cpu else cpu)
        except Exception as e:  # the closest-hit line is still reported
            res["mixed_workload"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_terrain:
        try:
            res["terrain"] = terrain_section(re, W, torch, ctx, cpu, with_cpu=not args.no_cpu)
        except Exception as e:  # the closest-hit line is still reported
            res["terrain"] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
