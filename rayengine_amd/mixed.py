"""Many triangle scenes and noise-terrain tiles, each with its own slice of
one ray array, traced one launch per item: a closest-hit launch per scene
(re_trace_closest) and, per tile, the exact-field march or, for a tile given
network parameters, the surrogate hybrid tracer (re_hybrid_trace).  Items
are issued in ascending ray count and dealt round-robin over the given
contexts, one HIP stream each, so that small launches overlap.  Inputs and
outputs are device buffers given as raw pointers.

This is the unbatched way to run the many-scene workload; re_batch_* runs
it as one persistent work-queue launch in a captured hipGraph.  No
counterpart in the reference, which traced one scene per frame
(RayEngine/src/cuda-src/renderer.cu:701-738).
"""
from __future__ import annotations

import numpy as np

from . import Context, KDTree, SceneObject, Surrogate, hybrid_trace_device

# exact-field march of the tiles: step max(0.05 f, 1 voxel), one bisection
TILE_MARCH = dict(k_exact=0.05, h_min_voxels=1.0, n_bisect=1)
# surrogate tiles: network steps while its distance exceeds 4 voxels, exact
# field near the surface
TILE_HYBRID = dict(tau_voxels=4.0, margin_voxels=1.0, k_exact=0.1, n_bisect=4)


class MixedRun:
    """trees: one KDTree per scene; tiles: (dim, seed) or (dim, seed, network
    parameters) per noise tile (a dim^3 volume over the unit cube; with
    parameters the tile is traced by the surrogate hybrid with that network);
    scene_ranges / tile_ranges: (first ray, ray count) of each item in the ray
    arrays.  Results: prim id and t per triangle ray, t and a 0/1 hit byte per
    terrain ray, written at the ray's index into prim_ptr (int32), t_ptr
    (float32) and hit_ptr (uint8)."""

    def __init__(self, ctxs: list[Context], trees: list[KDTree], tiles, scene_ranges, tile_ranges, o_ptr: int,
                 d_ptr: int, prim_ptr: int, t_ptr: int, hit_ptr: int, tile_march: dict | None = None,
                 tile_hybrid: dict | None = None):
        self.ctxs = list(ctxs)
        items = [("scene", k, int(b), int(c)) for k, (b, c) in enumerate(scene_ranges)]
        items += [("tile", k, int(b), int(c)) for k, (b, c) in enumerate(tile_ranges)]
        items.sort(key=lambda it: it[3])
        self.items = items
        self.owner = [i % len(self.ctxs) for i in range(len(items))]
        self.ptrs = (o_ptr, d_ptr, prim_ptr, t_ptr, hit_ptr)
        self.tile_march = dict(TILE_MARCH if tile_march is None else tile_march)
        self.tile_hybrid = dict(TILE_HYBRID if tile_hybrid is None else tile_hybrid)
        self.objs = []
        self.nets = []
        try:
            for i, (kind, k, _, _) in enumerate(items):
                cx = self.ctxs[self.owner[i]]
                if kind == "scene":
                    self.objs.append(cx.upload_scene([SceneObject(trees[k])]))
                    self.nets.append(None)
                else:
                    dm, seed = tiles[k][0], tiles[k][1]
                    params = tiles[k][2] if len(tiles[k]) > 2 else None
                    self.objs.append(cx.noise_volume(dim=(dm, dm, dm), spacing=(1.0 / dm,) * 3, seed=seed))
                    net = None
                    if params is not None:
                        net = Surrogate(cx)
                        net.set_params(params)
                    self.nets.append(net)
        except Exception:
            self.release()
            raise

    def run(self, block: int = 64, variant: int = 4):
        """Issues every item's launch (asynchronous on its context's stream)."""
        o_ptr, d_ptr, p_ptr, t_ptr, h_ptr = self.ptrs
        for (kind, _, b, c), obj, net in zip(self.items, self.objs, self.nets):
            if kind == "scene":
                obj.trace_closest_device(o_ptr + 12 * b, d_ptr + 12 * b, c, p_ptr + 4 * b, t_ptr + 4 * b, sync=False,
                                         block=block, variant=variant)
            else:
                hybrid_trace_device(obj, net, o_ptr + 12 * b, d_ptr + 12 * b, c, t_ptr + 4 * b, h_ptr + b,
                                    sync=False, **(self.tile_hybrid if net is not None else self.tile_march))

    def completion_ms(self) -> np.ndarray:
        """After a run has finished (and before any other launch on these
        contexts): per item (in self.items order) the device time from the
        start of its stream's first launch to the end of its own launch,
        from HIP events on that stream, idle gaps between launches
        included."""
        done = np.zeros(len(self.items))
        for s, cx in enumerate(self.ctxs):
            mine = [i for i in range(len(self.items)) if self.owner[i] == s]
            if mine:
                done[mine] = cx.launch_end_times(len(mine))
        return done

    def kernel_ms_sum(self) -> float:
        """After a run has finished: the summed kernel durations of all items."""
        return float(sum(np.sum(cx.kernel_times(sum(1 for o in self.owner if o == s)))
                         for s, cx in enumerate(self.ctxs)))

    def release(self):
        for o in self.objs:
            o.release()
        for n in self.nets:
            if n is not None:
                n.release()
        self.objs = []
        self.nets = []
