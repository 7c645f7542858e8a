"""Autotuner for the batched work-queue path.

Picks the kd-tree leaf size, the work-group size and the work-item
granularity (rays per queue item) of a batched workload by measurement:

1. leaf size: trees of a representative sample of the scenes (the ones with
   the most rays) are rebuilt for every candidate and the sample batch is
   timed;
2. block x item_rays: the full batch (trees at the chosen leaf size) is
   captured once per candidate pair and its graph replays are timed.

Timings are the HIP-event durations of the graph replays.  The result is
cached in a JSON file keyed by the device name and a signature of the
workload's sizes, so a repeated run reuses it.
"""
from __future__ import annotations

import concurrent.futures as cf
import hashlib
import json
import os
import time
from pathlib import Path

import numpy as np

from . import Batch, Context, KDTree

DEFAULT_CACHE = Path(os.environ.get("RAYENGINE_TUNE_CACHE", Path.home() / ".cache" / "rayengine_amd" / "autotune.json"))


def build_trees(soups, max_leaf_tris=8, threads=None, **kw):
    """Builds one tree per soup in parallel (the builder releases the GIL)."""
    threads = threads or min(16, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda s: KDTree(s, max_leaf_tris=max_leaf_tris, **kw), soups))


def _signature(soups, tiles, scene_ranges):
    h = hashlib.sha256()
    for s in soups:
        h.update(np.int64(len(s)).tobytes())
    for (_, c) in scene_ranges:
        h.update(np.int64(c).tobytes())
    for t in tiles:
        h.update(repr((list(t[0].dim), t[1], t[3])).encode())
    return h.hexdigest()[:16]


def _time_batch(ctx: Context, batch: Batch, reps: int) -> float:
    batch.run()
    for _ in range(reps):
        batch.run(sync=False)
    batch.run()  # waits
    return float(np.min(ctx.kernel_times(reps + 1)))


def autotune(ctx: Context, soups, tiles, ray_o, ray_d, scene_ranges, device_name: str = "gpu",
             leaf_sizes=(4, 8, 12), blocks=(64, 128, 256), item_rays=(64, 256, 1024), sample_scenes=8, reps=3,
             surrogate=None, cache: Path | None = DEFAULT_CACHE, log=print):
    """Returns (config dict, measurements list).  tiles: (Volume, mode, begin, count)."""
    key = f"{device_name}:{_signature(soups, tiles, scene_ranges)}"
    if cache is not None and cache.exists():
        try:
            data = json.loads(cache.read_text())
            if key in data:
                log(f"autotune: cached {data[key]['config']}")
                return data[key]["config"], data[key]["measurements"]
        except Exception:
            pass
    meas = []
    t0 = time.time()
    # 1. leaf size on the scenes with the most rays
    order = np.argsort([-c for (_, c) in scene_ranges])[:sample_scenes]
    sample = [soups[i] for i in order]
    sranges = [scene_ranges[i] for i in order]
    best_leaf, best_ms = None, np.inf
    for leaf in leaf_sizes:
        trees = build_trees(sample, max_leaf_tris=leaf)
        b = Batch(ctx, [(t, s, c) for t, (s, c) in zip(trees, sranges)], [], ray_o, ray_d)
        ms = _time_batch(ctx, b, reps)
        b.release()
        meas.append({"stage": "leaf", "leaf": leaf, "ms": ms})
        log(f"autotune: leaf {leaf}: {ms:.3f} ms")
        if ms < best_ms:
            best_leaf, best_ms = leaf, ms
    # 2. block x item granularity on the full batch
    trees = build_trees(soups, max_leaf_tris=best_leaf)
    best = None
    for blk in blocks:
        for it in item_rays:
            b = Batch(ctx, [(t, s, c) for t, (s, c) in zip(trees, scene_ranges)], tiles, ray_o, ray_d,
                      item_rays=it, block=blk, surrogate=surrogate)
            ms = _time_batch(ctx, b, reps)
            b.release()
            meas.append({"stage": "launch", "block": blk, "item_rays": it, "ms": ms})
            log(f"autotune: block {blk} item_rays {it}: {ms:.3f} ms")
            if best is None or ms < best[0]:
                best = (ms, blk, it)
    cfg = {"max_leaf_tris": int(best_leaf), "block": int(best[1]), "item_rays": int(best[2]),
           "tune_seconds": round(time.time() - t0, 1)}
    if cache is not None:
        try:
            cache.parent.mkdir(parents=True, exist_ok=True)
            data = json.loads(cache.read_text()) if cache.exists() else {}
            data[key] = {"config": cfg, "measurements": meas}
            cache.write_text(json.dumps(data, indent=1))
        except Exception:
            pass
    return cfg, meas
