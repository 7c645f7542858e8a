"""CPU simulation of the batched launch's queue protocol.

The persistent kernel (kernels/batch.hip) runs the items of the library's
own plan (re_batch_plan, the planner re_batch_create uses): every wave takes
the next item index from one counter (fetch-and-add), runs the item's rays,
and decrements its slot's remaining count; the wave that takes it to zero
stamps the slot's completion.  Surrogate tiles have no items: a branch beside
the kernel runs each and counts its one completion.  Here CPU threads play the
waves over the same plan, with the oracle as each item's ray routine, and the
checks are the protocol's invariants: the plan covers every target's rays
once, in ascending target size, with slot counts equal to the item counts;
every ray is written by exactly one item; every slot is stamped exactly once,
after all its items; and the results equal whole-range oracle calls.  The
checker is shown to catch a dropped, a duplicated and a shifted item."""
import threading

import numpy as np
import pytest

import oracle
import rayengine_amd as re
from rayengine_amd import workloads as W

MARCH = dict(k_exact=0.05, h_min_vox=1.0, n_bisect=1, max_iters=4096)


@pytest.fixture(scope="module")
def workload():
    soups, dims, seeds, ro, rd, sr, tr = W.batch_workload(n_scenes=6, n_tiles=4, tri_lo=1e3, tri_hi=2e4, ray_lo=2e3,
                                                          ray_hi=1.5e4, tile_lo=64, tile_hi=96, seed=3)
    trees = [re.KDTree(s) for s in soups]
    grids = [oracle.noise_fill(oracle.perlin_table(sd)[0], dim=(dm, dm, dm), spacing=(1.0 / dm,) * 3)
             for dm, sd in zip(dims, seeds)]
    modes = [0, 1, 2, 0]   # exact march, reference sphere trace, surrogate (side branch), exact march
    return trees, grids, dims, modes, ro, rd, sr, tr


def _run_target(w, slot, lo, hi):
    """The per-ray routine of a target over rays [lo, hi): (prim, t)."""
    trees, grids, dims, modes, ro, rd, sr, tr = w
    o, d = ro[lo:hi], rd[lo:hi]
    if slot < len(trees):
        tree = trees[slot]
        return oracle.trace(*tree.export(), tree.identity_object(), o, d)
    k = slot - len(trees)
    sp = (1.0 / dims[k],) * 3
    if modes[k] == 1:
        t, h, _ = oracle.sdf_trace(grids[k], sp, o, d)
    else:   # exact march (the surrogate branch's stand-in: the protocol, not the network, is under test)
        vox = np.float32(sp[0])
        t, h, _ = oracle.exact_march(grids[k], sp, o, d, MARCH["k_exact"], float(np.float32(MARCH["h_min_vox"] * vox)),
                                     MARCH["n_bisect"], MARCH["max_iters"])
    return np.where(h == 1, 0, -1).astype(np.int32), np.where(h == 1, t, np.float32(1e20)).astype(np.float32)


def _plan(w, item_rays):
    trees, grids, dims, modes, ro, rd, sr, tr = w
    tiles = [(m, b, c) for m, (b, c) in zip(modes, tr)]
    return re.batch_plan(sr, tiles, len(ro), item_rays)


def _check_plan(w, items, rem, item_rays):
    trees, grids, dims, modes, ro, rd, sr, tr = w
    ranges = list(sr) + list(tr)
    surrogate = {len(sr) + k for k, m in enumerate(modes) if m == 2}
    # targets in ascending ray count, ties in slot order, each contiguous
    seq = [int(s) for s in items["slot"]]
    order = [s for i, s in enumerate(seq) if i == 0 or seq[i - 1] != s]
    assert len(order) == len(set(order)), "a target's items are not contiguous"
    sizes = [ranges[s][1] for s in order]
    assert sizes == sorted(sizes) and all(
        order[i] < order[i + 1] for i in range(len(order) - 1) if sizes[i] == sizes[i + 1]), "target order"
    for s, (b, c) in enumerate(ranges):
        mine = items[items["slot"] == s]
        if s in surrogate:
            assert len(mine) == 0 and rem[s] == 1
            continue
        assert rem[s] == len(mine) == -(-c // item_rays)
        assert (mine["count"] <= item_rays).all() and (mine["count"] > 0).all()
        starts = b + np.concatenate([[0], np.cumsum(mine["count"])[:-1]]) if len(mine) else []
        assert np.array_equal(mine["begin"], starts) and mine["count"].sum() == c, f"slot {s} not covered in order"
        kind = 0 if s < len(sr) else (1 if modes[s - len(sr)] == 0 else 2)
        assert (mine["kind"] == kind).all() and (mine["target"] == (s if s < len(sr) else s - len(sr))).all()


def _simulate(w, items, rem, n_waves=8):
    """Threads as waves over the plan; returns (prim, t, cover, stamps,
    stamp_clock, item_clock)."""
    trees, grids, dims, modes, ro, rd, sr, tr = w
    n = len(ro)
    prim = np.full(n, -1, np.int32)
    t = np.full(n, np.float32(1e20), np.float32)
    cover = np.zeros(n, np.int32)
    remaining = np.array(rem, np.int64)
    stamps = np.zeros(len(rem), np.int32)
    stamp_clock = np.full(len(rem), -1, np.int64)
    item_clock = np.full(len(items), -1, np.int64)
    state = {"counter": 0, "clock": 0}
    lock = threading.Lock()

    def finish(slot, it=None):
        with lock:
            state["clock"] += 1
            if it is not None:
                item_clock[it] = state["clock"]
            before = remaining[slot]
            remaining[slot] -= 1
            if before == 1:
                stamps[slot] += 1
                stamp_clock[slot] = state["clock"]

    def wave():
        while True:
            with lock:
                it = state["counter"]
                state["counter"] += 1
            if it >= len(items):
                return
            kind, target, begin, count, slot = (int(items[it][f]) for f in ("kind", "target", "begin", "count", "slot"))
            p, tt = _run_target(w, slot, begin, begin + count)
            prim[begin:begin + count] = p
            t[begin:begin + count] = tt
            np.add.at(cover, np.arange(begin, begin + count), 1)
            finish(slot, it)

    def side_branch():
        for k, m in enumerate(modes):
            if m == 2:
                s = len(sr) + k
                b, c = tr[k]
                p, tt = _run_target(w, s, b, b + c)
                prim[b:b + c] = p
                t[b:b + c] = tt
                np.add.at(cover, np.arange(b, b + c), 1)
                finish(s)

    threads = [threading.Thread(target=wave) for _ in range(n_waves)] + [threading.Thread(target=side_branch)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    return prim, t, cover, stamps, stamp_clock, item_clock


def _check_run(w, items, rem, prim, t, cover, stamps, stamp_clock, item_clock, want):
    trees, grids, dims, modes, ro, rd, sr, tr = w
    ranges = list(sr) + list(tr)
    targeted = np.zeros(len(ro), bool)
    for b, c in ranges:
        targeted[b:b + c] = True
    assert (cover[targeted] == 1).all(), "a ray written by no item or by two"
    assert (cover[~targeted] == 0).all()
    for s, (b, c) in enumerate(ranges):
        if c == 0:
            continue
        assert stamps[s] == 1, f"slot {s} stamped {stamps[s]} times"
        mine = np.flatnonzero(items["slot"] == s)
        if len(mine):
            assert stamp_clock[s] >= item_clock[mine].max(), f"slot {s} stamped before its last item"
        p, tt = want[s]
        assert np.array_equal(prim[b:b + c], p) and np.array_equal(t[b:b + c].view(np.uint32), tt.view(np.uint32))


@pytest.fixture(scope="module")
def whole_range(workload):
    trees, grids, dims, modes, ro, rd, sr, tr = workload
    return [_run_target(workload, s, b, b + c) for s, (b, c) in enumerate(list(sr) + list(tr))]


@pytest.mark.parametrize("item_rays", [256, 64, 1000])
def test_queue_protocol_over_the_library_plan(workload, whole_range, item_rays):
    items, rem = _plan(workload, item_rays)
    _check_plan(workload, items, rem, item_rays)
    out = _simulate(workload, items, rem)
    _check_run(workload, items, rem, *out, want=whole_range)


@pytest.mark.parametrize("mistake", ["drop", "duplicate", "shift"])
def test_checker_catches_a_broken_plan(workload, whole_range, mistake):
    items, rem = _plan(workload, 256)
    items = items.copy()
    k = len(items) // 2
    if mistake == "drop":
        items = np.delete(items, k)
    elif mistake == "duplicate":
        items = np.insert(items, k, items[k])
    else:
        items[k]["begin"] += 1
    with pytest.raises(AssertionError):
        _check_plan(workload, items, rem, 256)
    if mistake != "shift":   # a shifted item would also read past its target; drop and duplicate run
        with pytest.raises(AssertionError):
            _check_run(workload, items, rem, *_simulate(workload, items, rem), want=whole_range)


def test_plan_refuses_bad_ranges():
    with pytest.raises(re.ReError):
        re.batch_plan([(10, 100)], [], 50)
    with pytest.raises(re.ReError):   # begin + count overflows int64
        re.batch_plan([(1 << 62, 1 << 62)], [], 50)
    with pytest.raises(re.ReError):
        re.batch_plan([], [(0, 1 << 62, 1 << 62)], 50)
    with pytest.raises(re.ReError):
        re.batch_plan([], [(3, 0, 10)], 50)
    with pytest.raises(re.ReError):
        re.batch_plan([(0, 10)], [], 50, item_rays=0)
