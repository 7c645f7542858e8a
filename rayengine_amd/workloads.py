"""Seeded synthetic workloads of the ray-query path.

* small_scene      - icosphere (subdivided 4x, 5120 triangles) + 1024 random
                     triangles (centroids U[-2,2]^3, edges U[0.02,0.2], seed 0),
                     256x256 pinhole rays from (0,0,5) towards the origin, FOV 60.
* cluster_scene    - 2^20 triangles in 64 Gaussian clusters (sigma 0.3, centres
                     U[-10,10]^3, edge lengths LogNormal(-3, 0.5), seed 1);
                     1920x1080 primary rays from 0.8 units outside the cluster
                     nearest the origin, looking at its centre (95 % of the
                     primaries hit), plus one cosine-weighted diffuse bounce ray
                     per primary hit: 4.04 M rays, 49 % of them bounces.
"""
from __future__ import annotations

import numpy as np

from . import Camera, KDTree, bounce_rays, camera_rays, cluster_soup, icosphere, random_tris

SMALL_CAMERA = Camera(pos=(0.0, 0.0, 5.0), view=(0.0, 0.0, -1.0))
# The cluster workload's camera: the soup's cluster nearest the origin has its
# centre near (-0.6805, 1.1570, 0.9132) (mean of the triangle centroids within
# 1 unit of it); the camera stands 0.8 units further out along the line from
# the origin through that centre and looks at it, so the cluster fills the
# view: 1,964,416 of the 2,073,600 primaries hit, and the bounce rays leave a
# dense cluster in every direction of the hemisphere (the incoherent part).
CLUSTER_CAMERA = Camera(pos=(-1.016, 1.727, 1.363), view=(0.3353, -0.5701, -0.45))
# The previous framing, from (0,0,30) towards the origin: 14 % of the
# primaries hit, so bounces were 9 % of its 2.28 M rays.
CLUSTER_CAMERA_FAR = Camera(pos=(0.0, 0.0, 30.0), view=(0.0, 0.0, -1.0))
CLUSTER_TRIS = 1 << 20
BOUNCE_SEED = 1


def small_soup() -> np.ndarray:
    return np.concatenate([icosphere(4, 1.0), random_tris(0, 1024, -2.0, 2.0, 0.02, 0.2)])


def small_scene(width: int = 256, height: int = 256):
    """(tri_v, tree, ray_o, ray_d) of the small icosphere workload."""
    tri_v = small_soup()
    tree = KDTree(tri_v, max_leaf_tris=8, max_depth=24)
    o, d = camera_rays(SMALL_CAMERA, width, height)
    return tri_v, tree, o, d


def cluster_soup_default(n: int = CLUSTER_TRIS, seed: int = 1) -> np.ndarray:
    return cluster_soup(seed, n, 64, -10.0, 10.0, 0.3, -3.0, 0.5)


def tile_order(width: int, height: int, tile: int = 8) -> np.ndarray:
    """Pixel indices (y * width + x) visited tile by tile: tile x tile pixel
    blocks in row order, pixels inside a block in row order; partial blocks
    at the right and bottom edges keep the same scheme."""
    ys, xs = np.divmod(np.arange(width * height, dtype=np.int64), width)
    tiles_x = (width + tile - 1) // tile
    key = ((ys // tile) * tiles_x + xs // tile) * (tile * tile) + (ys % tile) * tile + xs % tile
    return np.argsort(key, kind="stable")


def cluster_rays(tri_v, primary_prim, primary_t, width=1920, height=1080, order=None):
    """Primary rays (in pixel order, or permuted by `order`, e.g. tile_order)
    plus one bounce ray per primary hit, in the primaries' order.  The
    primary hits passed in are in pixel order."""
    o, d = camera_rays(CLUSTER_CAMERA, width, height)
    bo, bd = bounce_rays(tri_v, o, d, primary_prim, primary_t, BOUNCE_SEED)
    if order is not None:
        o, d = o[order], d[order]
        hit = np.flatnonzero(primary_prim >= 0)          # bounce k belongs to pixel hit[k]
        rank = np.empty(len(primary_prim), np.int64)
        rank[order] = np.arange(len(order))
        bsel = np.argsort(rank[hit], kind="stable")
        bo, bd = bo[bsel], bd[bsel]
    return np.concatenate([o, bo]), np.concatenate([d, bd])


def cluster_workload_rays(scene, tri_v, tile: int = 8, width: int = 1920, height: int = 1080):
    """The cluster workload's rays for a scene holding the cluster soup:
    primaries in tile x tile pixel blocks (1 = scanline order), then one
    bounce ray per primary hit in the primaries' order."""
    o, d = camera_rays(CLUSTER_CAMERA, width, height)
    p, t = scene.trace_closest(o, d)
    return cluster_rays(tri_v, p, t, width, height, order=tile_order(width, height, tile) if tile > 1 else None)


# Shadow rays: from each primary hit, offset 1e-4 along the geometric normal
# facing the incoming ray, towards a point light.
CLUSTER_LIGHT = (20.0, 30.0, 25.0)


def shadow_rays(tri_v, o, d, prim, t, light):
    """(origin, direction) float32 of one shadow ray per hit (prim >= 0)."""
    hit = np.flatnonzero(prim >= 0)
    v = np.asarray(tri_v, np.float32).reshape(-1, 3, 3)[prim[hit]]
    nrm = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    facing = np.where(((nrm * d[hit]).sum(1) > 0)[:, None], -nrm, nrm)
    org = (o[hit] + t[hit, None] * d[hit] + np.float32(1e-4) * facing).astype(np.float32)
    dirv = np.asarray(light, np.float32) - org
    dirv /= np.linalg.norm(dirv, axis=1, keepdims=True)
    return org, dirv.astype(np.float32)


# Terrain workload: 256^3 FP32 field over the unit cube (seed 2016), traced
# with 1920x1080 rays from an elevated camera looking across the terrain.
TERRAIN_CAMERA = Camera(pos=(0.5, 0.95, -0.35), view=(0.0, -0.55, 1.0))
TERRAIN_SPACING = (1.0 / 256, 1.0 / 256, 1.0 / 256)


def terrain_rays(width: int = 1920, height: int = 1080):
    return camera_rays(TERRAIN_CAMERA, width, height)


# Batched mixed workload: n_scenes cluster soups with log-uniform triangle
# counts in [tri_lo, tri_hi] (seeds 100 + k) and n_tiles terrain tiles with
# log-uniform sizes in [64, 256]^3 (permutation seeds 2016 + k), each target
# with a log-uniform incoherent ray batch in [ray_lo, ray_hi]: origins on a
# sphere around the target, directions towards uniform points inside it.
def _log_uniform(rng, lo, hi, n):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n))


def _incoherent_rays(rng, lo, hi, n):
    c = (lo + hi) / 2
    r = 0.75 * np.linalg.norm(hi - lo) + 1e-3
    u = rng.standard_normal((n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    o = c + r * u
    tgt = lo + (hi - lo) * rng.uniform(0, 1, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(np.float32), d.astype(np.float32)


def batch_workload(n_scenes=64, n_tiles=16, tri_lo=1e3, tri_hi=5e5, ray_lo=1.6e4, ray_hi=2e6, tile_lo=64,
                   tile_hi=256, seed=7):
    """Returns (scene_soups, tile_dims, tile_seeds, ray_o, ray_d, scene_ranges, tile_ranges)."""
    rng = np.random.default_rng(seed)
    n_tris = np.round(_log_uniform(rng, tri_lo, tri_hi, n_scenes)).astype(int)
    n_rays_s = np.round(_log_uniform(rng, ray_lo, ray_hi, n_scenes)).astype(int)
    dims = (np.round(_log_uniform(rng, tile_lo, tile_hi, n_tiles) / 8) * 8).astype(int)
    n_rays_t = np.round(_log_uniform(rng, ray_lo, ray_hi, n_tiles)).astype(int)
    soups, os_, ds_, scene_ranges, tile_ranges = [], [], [], [], []
    off = 0
    for k in range(n_scenes):
        soup = cluster_soup(100 + k, int(n_tris[k]))
        soups.append(soup)
        v = soup.reshape(-1, 3)
        o, d = _incoherent_rays(rng, v.min(0).astype(np.float64), v.max(0).astype(np.float64), int(n_rays_s[k]))
        os_.append(o)
        ds_.append(d)
        scene_ranges.append((off, len(o)))
        off += len(o)
    for k in range(n_tiles):
        o, d = _incoherent_rays(rng, np.zeros(3), np.ones(3), int(n_rays_t[k]))
        os_.append(o)
        ds_.append(d)
        tile_ranges.append((off, len(o)))
        off += len(o)
    return (soups, [int(x) for x in dims], [2016 + k for k in range(n_tiles)], np.concatenate(os_),
            np.concatenate(ds_), scene_ranges, tile_ranges)
