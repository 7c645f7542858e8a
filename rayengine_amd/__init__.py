"""rayengine_amd - MI355X-native ray-query path of RayEngine.

Python host mirror of the C ABI in include/rayengine.h.  The classes follow
the reference's host-side objects:

* KDTree        - RKDTreeCPU + RKDThreeGPU (RayEngine/src/Primitives/KDTree.h:39-99,
                  KDThreeGPU.h:9): build over a triangle soup, flattened for the GPU.
* SceneObject   - GPUSceneObject (RayEngine/src/cuda-src/gpu_structs.h:69-90).
* Context.upload_scene - copy_memory (RayEngine/src/cuda-src/cuda_memory_functions.cuh:58-211).
* Scene.trace_closest  - trace_primary_rays / trace_scene
                  (RayEngine/src/cuda-src/renderer.cu:389-410, ray_functions.cuh:376-396).

All compute runs in librayengine.so (HIP kernels for gfx950); this module
only moves arrays across the boundary.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (LEAF_DT, NODE_DT, OBJ_DT, RE_ASYNC, RE_DEVICE_PTRS, TRI_DT, ReError, check, lib, ptr)

__all__ = ["KDTree", "SceneObject", "Context", "Scene", "Volume", "Camera", "ReError", "perlin_table",
           "noise_params", "sdf_params", "icosphere", "random_tris",
           "cluster_soup", "camera_rays", "bounce_rays", "NODE_DT", "LEAF_DT", "TRI_DT", "OBJ_DT",
           "RE_DEVICE_PTRS", "RE_ASYNC", "K_EPSILON", "K_INFINITY"]

K_EPSILON = np.float32(1e-4)   # RayEngine/src/RayEngine/RayEngine.h:24
K_INFINITY = np.float32(1e20)  # RayEngine/src/RayEngine/RayEngine.h:25


def _f32c(a, shape_last=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape_last is not None and a.size % shape_last:
        raise ValueError(f"array size {a.size} is not a multiple of {shape_last}")
    return a


# ------------------------------------------------------------------ generators

def icosphere(subdiv: int, radius: float = 1.0) -> np.ndarray:
    """Subdivided icosphere as an (n, 9) float32 soup, outward CCW faces."""
    n = C.c_int64()
    check(lib().re_gen_icosphere(subdiv, radius, None, C.byref(n)))
    out = np.empty((n.value, 9), np.float32)
    check(lib().re_gen_icosphere(subdiv, radius, ptr(out), C.byref(n)))
    return out


def random_tris(seed: int, n: int, lo=-2.0, hi=2.0, len_lo=0.02, len_hi=0.2) -> np.ndarray:
    out = np.empty((n, 9), np.float32)
    check(lib().re_gen_random_tris(seed, n, lo, hi, len_lo, len_hi, ptr(out)))
    return out


def cluster_soup(seed: int, n: int, n_clusters=64, lo=-10.0, hi=10.0, sigma=0.3, ln_mu=-3.0,
                 ln_sigma=0.5) -> np.ndarray:
    out = np.empty((n, 9), np.float32)
    check(lib().re_gen_cluster_soup(seed, n, n_clusters, lo, hi, sigma, ln_mu, ln_sigma, ptr(out)))
    return out


@dataclass
class Camera:
    """RCamera (RayEngine/src/Primitives/Camera.h:7-16); defaults of RMovableCamera
    (RayEngine/src/Objects/MovableCamera.cpp:4-14,101-118)."""
    pos: tuple = (0.0, 0.0, 5.0)
    view: tuple = (0.0, 0.0, -1.0)
    camdown: tuple = (0.0, -1.0, 0.0)
    fov: tuple = (60.0, 60.0)
    focal_distance: float = 4.0

    @staticmethod
    def look_at(pos, target, **kw) -> "Camera":
        v = np.asarray(target, np.float64) - np.asarray(pos, np.float64)
        return Camera(pos=tuple(map(float, pos)), view=tuple(map(float, v)), **kw)

    def _c(self) -> _lib.re_camera:
        c = _lib.re_camera()
        c.pos[:] = self.pos
        c.view[:] = self.view
        c.camdown[:] = self.camdown
        c.fov[:] = self.fov
        c.focal_distance = self.focal_distance
        return c


def camera_rays(cam: Camera, width: int, height: int) -> tuple[np.ndarray, np.ndarray]:
    """Primary rays of generate_ray (RayEngine/src/cuda-src/ray_functions.cuh:13-55)."""
    o = np.empty((width * height, 3), np.float32)
    d = np.empty((width * height, 3), np.float32)
    c = cam._c()
    check(lib().re_gen_camera_rays(C.byref(c), width, height, ptr(o), ptr(d)))
    return o, d


def bounce_rays(tri_v, o, d, prim, t, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """One cosine-weighted diffuse bounce per hit (prim >= 0)."""
    tri_v = _f32c(tri_v, 9)
    o, d = _f32c(o, 3), _f32c(d, 3)
    prim = np.ascontiguousarray(prim, np.int32)
    t = _f32c(t)
    n = prim.size
    oo = np.empty((n, 3), np.float32)
    od = np.empty((n, 3), np.float32)
    m = C.c_int64()
    check(lib().re_gen_bounce_rays(ptr(tri_v), tri_v.size // 9, ptr(o), ptr(d), ptr(prim), ptr(t), n, seed,
                                   ptr(oo), ptr(od), C.byref(m)))
    return oo[:m.value].copy(), od[:m.value].copy()


# ------------------------------------------------------------------ kd-tree

class KDTree:
    """Rope kd-tree over a triangle soup ((n, 9) float32: v0, v1, v2)."""

    def __init__(self, tri_v, max_leaf_tris: int = 8, max_depth: int = 24, cost_traversal: float = 1.0,
                 cost_intersect: float = 1.5, empty_bonus: float = 0.8, n_bins: int = 32, clip: bool = True,
                 ctx: "Context | None" = None, split_rule: int = 0):
        """Host build (re_kd_build), or with ctx the device build
        (re_kd_build_gpu, the same tree; its device time in build_ms).
        split_rule 1: the reference's spatial median (host build only)."""
        self.tri_v = _f32c(tri_v, 9).reshape(-1, 9)
        p = _lib.re_kd_build_params(max_leaf_tris, max_depth, cost_traversal, cost_intersect, empty_bonus, n_bins,
                                    int(clip), int(split_rule))
        h = C.c_void_p()
        self.build_ms = None
        if ctx is None:
            check(lib().re_kd_build(ptr(self.tri_v), self.tri_v.shape[0], C.byref(p), C.byref(h)))
        else:
            ms = C.c_float()
            check(lib().re_kd_build_gpu(ctx.handle, ptr(self.tri_v), self.tri_v.shape[0], C.byref(p), C.byref(h),
                                        C.byref(ms)), ctx.handle)
            self.build_ms = ms.value
        self._h = h
        info = _lib.re_kd_info()
        check(lib().re_kd_get_info(h, C.byref(info)))
        self.info = {"n_tris": info.n_tris, "n_nodes": info.n_nodes, "n_leaves": info.n_leaves,
                     "n_trirecs": info.n_trirecs, "depth": info.depth, "max_leaf": info.max_leaf,
                     "root_min": np.array(info.root_min[:], np.float32),
                     "root_max": np.array(info.root_max[:], np.float32)}

    @property
    def handle(self):
        return self._h

    def export(self):
        """(nodes, leaves, tris) as structured numpy arrays (NODE_DT, LEAF_DT, TRI_DT)."""
        nodes = np.zeros(self.info["n_nodes"], NODE_DT)
        leaves = np.zeros(self.info["n_leaves"], LEAF_DT)
        tris = np.zeros(self.info["n_trirecs"], TRI_DT)
        check(lib().re_kd_export(self._h, ptr(nodes), ptr(leaves), ptr(tris)))
        return nodes, leaves, tris

    def identity_object(self) -> np.ndarray:
        """The object record of this tree alone at identity transform (OBJ_DT)."""
        ob = np.zeros(1, OBJ_DT)
        ob["cos_y"] = 1.0
        ob["root"] = 0
        ob["root_min"] = self.info["root_min"]
        ob["root_max"] = self.info["root_max"]
        return ob

    def __del__(self):
        try:
            h = getattr(self, "_h", None)
            if h and _lib._lib is not None:
                _lib._lib.re_kd_free(h)
                self._h = None
        except Exception:
            pass  # interpreter shutdown


@dataclass
class SceneObject:
    """GPUSceneObject: a tree, a location and a rotation about Y (radians)."""
    tree: KDTree
    location: tuple = (0.0, 0.0, 0.0)
    rotation_y: float = 0.0
    is_character: bool = False
    prim_offset: int = 0


# ------------------------------------------------------------------ device

class Context:
    """Owns a HIP stream and the device-resident scenes (re_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().re_create(device, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def last_kernel_ms(self) -> float:
        v = C.c_float()
        check(lib().re_last_kernel_ms(self._h, C.byref(v)), self._h)
        return v.value

    def debug_bounds_errors(self) -> int:
        """Bounds-check bits set by the trace kernels since the last call (1
        node entry, 2 leaf record, 4 triangle run), after the stream finishes;
        only in the debug library (RE_LIBRARY=debug), ReError otherwise."""
        v = C.c_uint32()
        check(lib().re_debug_bounds_errors(self._h, C.byref(v)), self._h)
        return int(v.value)

    def kernel_times(self, max_n: int = 256) -> np.ndarray:
        """Durations (ms) of the last launches, oldest first (HIP events)."""
        buf = np.zeros(max_n, np.float32)
        k = C.c_int32()
        check(lib().re_kernel_times(self._h, buf.ctypes.data_as(_lib._F), max_n, C.byref(k)), self._h)
        return buf[:k.value].copy()

    def launch_end_times(self, max_n: int = 256) -> np.ndarray:
        """For the last launches, oldest first: ms from the start of the first
        of them to the end of each (HIP events; gaps between launches count)."""
        buf = np.zeros(max_n, np.float32)
        k = C.c_int32()
        check(lib().re_launch_end_times(self._h, buf.ctypes.data_as(_lib._F), max_n, C.byref(k)), self._h)
        return buf[:k.value].copy()

    # ------------------------------------------------------------ terrain

    def noise_volume(self, **kw) -> "Volume":
        """Generates the fBm terrain volume on the device (re_noise_generate).
        Keywords: dim, spacing, seed, octaves, lacunarity, gain, base_freq, y_offset."""
        p = noise_params(**kw)
        vid = C.c_int32()
        check(lib().re_noise_generate(self._h, C.byref(p), C.byref(vid)), self._h)
        return Volume(self, vid.value)

    def load_rsdf(self, path) -> "Volume":
        """Reads a .rsdf file into a new volume (re_volume_load_rsdf)."""
        vid = C.c_int32()
        check(lib().re_volume_load_rsdf(self._h, str(path).encode(), C.byref(vid)), self._h)
        return Volume(self, vid.value)

    def upload_volume(self, grid: np.ndarray, spacing) -> "Volume":
        """Uploads a (dz, dy, dx) float32 grid (x fastest) with the given spacing."""
        g = np.ascontiguousarray(grid, np.float32)
        dim = np.array([g.shape[2], g.shape[1], g.shape[0]], np.int32)
        sp = np.array(spacing, np.float32).reshape(3)
        vid = C.c_int32()
        check(lib().re_volume_upload(self._h, ptr(dim), ptr(sp), ptr(g), C.byref(vid)), self._h)
        return Volume(self, vid.value)

    def upload_volume_pairs(self, grid: np.ndarray, material: np.ndarray, spacing) -> "Volume":
        """Uploads (dz, dy, dx) distance and material grids (x fastest) as one
        volume with a material channel (re_volume_upload_pairs)."""
        g = np.ascontiguousarray(grid, np.float32)
        m = np.ascontiguousarray(material, np.float32)
        if g.shape != m.shape or g.ndim != 3:
            raise ValueError("distance and material grids must have one (dz, dy, dx) shape")
        pairs = np.ascontiguousarray(np.stack([g, m], axis=-1))
        dim = np.array([g.shape[2], g.shape[1], g.shape[0]], np.int32)
        sp = np.array(spacing, np.float32).reshape(3)
        vid = C.c_int32()
        check(lib().re_volume_upload_pairs(self._h, ptr(dim), ptr(sp), ptr(pairs), C.byref(vid)), self._h)
        return Volume(self, vid.value)

    def upload_scene(self, objects: list[SceneObject]) -> "Scene":
        arr = (_lib.re_object * len(objects))()
        for i, ob in enumerate(objects):
            a = arr[i]
            a.tree = ob.tree.handle
            a.location[:] = ob.location
            a.sin_rot_y = float(np.float32(math.sin(ob.rotation_y)))
            a.cos_rot_y = float(np.float32(math.cos(ob.rotation_y)))
            a.is_character = int(ob.is_character)
            a.prim_offset = int(ob.prim_offset)
        sid = C.c_int32()
        check(lib().re_scene_upload(self._h, arr, len(objects), C.byref(sid)), self._h)
        return Scene(self, sid.value, list(objects))

    def close(self):
        if getattr(self, "_h", None):
            lib().re_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rsdf_info(path):
    """(dim, spacing) of a well-formed .rsdf file (no device needed); raises
    ReError for a missing, short, long or malformed file."""
    dim = (C.c_int32 * 3)()
    sp = (C.c_float * 3)()
    check(lib().re_rsdf_info(str(path).encode(), dim, sp), None)
    return tuple(dim), tuple(sp)


def perlin_table(seed: int = 2016) -> np.ndarray:
    """Doubled permutation table of RPerlinNoise(seed) (512 uint32)."""
    out = np.zeros(512, np.uint32)
    check(lib().re_perlin_table(seed, ptr(out)))
    return out


def noise_params(**kw) -> _lib.re_noise_params:
    p = _lib.re_noise_params()
    lib().re_noise_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "dim":
            p.dim[:] = [int(x) for x in (v if np.ndim(v) else (v, v, v))]
        elif k in ("spacing", "origin"):
            getattr(p, k)[:] = [float(x) for x in (v if np.ndim(v) else (v, v, v))]
        else:
            setattr(p, k, v)
    return p


def sdf_params(**kw) -> _lib.re_sdf_params:
    p = _lib.re_sdf_params()
    lib().re_sdf_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def frame_params(**kw) -> _lib.re_frame_params:
    """re_frame_params with the defaults, overridden by keywords (sun_dir,
    earth_radius, atmosphere_radius, hr, hm, sky_origin, base_color, gamma,
    vignetting, vignetting_k)."""
    p = _lib.re_frame_params()
    lib().re_frame_default_params(C.byref(p))
    for k, v in kw.items():
        if k in ("sun_dir", "sky_origin", "base_color"):
            getattr(p, k)[:] = [float(x) for x in v]
        else:
            setattr(p, k, type(getattr(p, k))(v))
    return p


def frame_shade(ctx: "Context", ray_d, hit, normal, width: int, height: int, **kw) -> np.ndarray:
    """Colours a traced frame (re_frame_shade): (height, width) uint32 RGBA8."""
    d = _f32c(ray_d, 3)
    nrm = _f32c(normal, 3)
    h = np.ascontiguousarray(hit, np.uint8)
    out = np.empty((height, width), np.uint32)
    p = frame_params(**kw)
    check(lib().re_frame_shade(ctx.handle, ptr(d), ptr(h), ptr(nrm), int(width), int(height), C.byref(p), ptr(out),
                               0), ctx.handle)
    return out


def frame_shade_device(ctx: "Context", d_ptr: int, hit_ptr: int, nrm_ptr: int, width: int, height: int, out_ptr: int,
                       sync=True, **kw):
    """re_frame_shade on device-resident arrays."""
    p = frame_params(**kw)
    flags = RE_DEVICE_PTRS | (0 if sync else RE_ASYNC)
    check(lib().re_frame_shade(ctx.handle, d_ptr, hit_ptr, nrm_ptr, int(width), int(height), C.byref(p), out_ptr,
                               flags), ctx.handle)


def write_ppm(path, rgba: np.ndarray):
    """Writes (height, width) RGBA8 pixels as a binary PPM (re_write_ppm)."""
    a = np.ascontiguousarray(rgba, np.uint32)
    check(lib().re_write_ppm(str(path).encode(), ptr(a), int(a.shape[1]), int(a.shape[0])))


BRUSH_TYPES = {"sphere_add": 0, "sphere_subtract": 1, "cube_add": 2, "cube_subtract": 3}


class Volume:
    """A device-resident terrain volume (bricked FP32)."""

    def __init__(self, ctx: Context, vid: int):
        self.ctx, self.id = ctx, vid
        dim = np.zeros(3, np.int32)
        sp = np.zeros(3, np.float32)
        nb = C.c_int64()
        check(lib().re_volume_info(ctx.handle, vid, ptr(dim), ptr(sp), C.byref(nb)), ctx.handle)
        self.dim, self.spacing, self.bytes = dim, sp, nb.value

    def read(self) -> np.ndarray:
        """The grid as a (dz, dy, dx) float32 array."""
        out = np.empty((self.dim[2], self.dim[1], self.dim[0]), np.float32)
        check(lib().re_volume_read(self.ctx.handle, self.id, ptr(out)), self.ctx.handle)
        return out

    def read_material(self) -> np.ndarray:
        """The material channel as a (dz, dy, dx) float32 array (zeros if none)."""
        out = np.empty((self.dim[2], self.dim[1], self.dim[0]), np.float32)
        check(lib().re_volume_read_material(self.ctx.handle, self.id, ptr(out)), self.ctx.handle)
        return out

    def save_rsdf(self, path):
        """Writes the volume as a .rsdf file (re_volume_save_rsdf)."""
        check(lib().re_volume_save_rsdf(self.ctx.handle, self.id, str(path).encode()), self.ctx.handle)

    def edit(self, kind="sphere_add", center=(0.0, 0.0, 0.0), extent=(1.0, 1.0, 1.0), material=0.0, snap=False,
             smooth_k=0.1):
        """Applies one brush (re_volume_edit).  kind: sphere_add,
        sphere_subtract, cube_add, cube_subtract."""
        b = _lib.re_brush()
        lib().re_brush_default(C.byref(b))
        b.type = BRUSH_TYPES[kind] if isinstance(kind, str) else int(kind)
        b.center[:] = [float(x) for x in center]
        b.extent[:] = [float(x) for x in extent]
        b.material, b.snap_to_grid, b.smooth_k = float(material), int(bool(snap)), float(smooth_k)
        check(lib().re_volume_edit(self.ctx.handle, self.id, C.byref(b)), self.ctx.handle)

    def trace(self, o, d, normals=False, **params):
        """Sphere trace: (t, hit, steps[, normal])."""
        o, d = _f32c(o, 3), _f32c(d, 3)
        n = o.size // 3
        t = np.empty(n, np.float32)
        hit = np.empty(n, np.uint8)
        steps = np.empty(n, np.int32)
        nrm = np.empty((n, 3), np.float32) if normals else None
        p = sdf_params(**params)
        check(lib().re_sdf_trace(self.ctx.handle, self.id, C.byref(p), ptr(o), ptr(d), n, ptr(t), ptr(hit),
                                 ptr(steps), ptr(nrm), 0), self.ctx.handle)
        return (t, hit, steps, nrm) if normals else (t, hit, steps)

    def trace_device(self, o_ptr, d_ptr, n, t_ptr, hit_ptr, steps_ptr=None, nrm_ptr=None, sync=True, **params):
        p = sdf_params(**params)
        flags = RE_DEVICE_PTRS | (0 if sync else RE_ASYNC)
        check(lib().re_sdf_trace(self.ctx.handle, self.id, C.byref(p), o_ptr, d_ptr, n, t_ptr, hit_ptr, steps_ptr,
                                 nrm_ptr, flags), self.ctx.handle)

    def release(self):
        if self.id is not None:
            check(lib().re_volume_release(self.ctx.handle, self.id), self.ctx.handle)
            self.id = None


class Surrogate:
    """The learned SDF surrogate network (re_mlp_*)."""

    N_PARAMS = 18945

    def __init__(self, ctx: Context, seed: int = 0):
        self.ctx = ctx
        mid = C.c_int32()
        check(lib().re_mlp_create(ctx.handle, seed, C.byref(mid)), ctx.handle)
        self.id = mid.value

    def params(self):
        p = np.empty(self.N_PARAMS, np.float32)
        m = np.empty(self.N_PARAMS, np.float32)
        v = np.empty(self.N_PARAMS, np.float32)
        st = C.c_int32()
        check(lib().re_mlp_get_params(self.ctx.handle, self.id, ptr(p), ptr(m), ptr(v), C.byref(st)), self.ctx.handle)
        return p, m, v, st.value

    def set_params(self, p):
        p = _f32c(p)
        assert p.size == self.N_PARAMS
        check(lib().re_mlp_set_params(self.ctx.handle, self.id, ptr(p)), self.ctx.handle)

    def forward(self, pts) -> np.ndarray:
        pts = _f32c(pts, 3)
        n = pts.size // 3
        out = np.empty(n, np.float32)
        check(lib().re_mlp_forward(self.ctx.handle, self.id, ptr(pts), n, ptr(out), 0), self.ctx.handle)
        return out

    def forward_device(self, pts_ptr: int, n: int, out_ptr: int, sync=True):
        """Forward pass over n device-resident points (float32 x, y, z)."""
        flags = RE_DEVICE_PTRS | (0 if sync else RE_ASYNC)
        check(lib().re_mlp_forward(self.ctx.handle, self.id, pts_ptr, n, out_ptr, flags), self.ctx.handle)

    def batch_step(self, samples, over_weight=4.0, adam=False, lr=1e-3):
        """(loss, grad) of one batch of (x, y, z, target) samples; optionally one Adam step."""
        s = _f32c(samples, 4)
        n = s.size // 4
        g = np.empty(self.N_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        check(lib().re_mlp_batch_step(self.ctx.handle, self.id, ptr(s), n, over_weight, int(adam), lr, ptr(g),
                                      ptr(loss)), self.ctx.handle)
        return float(loss[0]), g

    def train(self, dist_volume: "Volume", **kw) -> np.ndarray:
        p = _lib.re_mlp_train_params()
        lib().re_mlp_default_train_params(C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        hist = np.zeros(max(p.steps, 1), np.float32)
        check(lib().re_mlp_train(self.ctx.handle, self.id, dist_volume.id, C.byref(p), ptr(hist)), self.ctx.handle)
        return hist[:p.steps]

    def save(self, path):
        """Writes a training checkpoint (re_mlp_save)."""
        check(lib().re_mlp_save(self.ctx.handle, self.id, str(path).encode()), self.ctx.handle)

    @classmethod
    def load(cls, ctx: Context, path) -> "Surrogate":
        """A network restored from a checkpoint (re_mlp_load)."""
        mid = C.c_int32()
        check(lib().re_mlp_load(ctx.handle, str(path).encode(), C.byref(mid)), ctx.handle)
        s = cls.__new__(cls)
        s.ctx, s.id = ctx, mid.value
        return s

    def release(self):
        if self.id is not None:
            check(lib().re_mlp_release(self.ctx.handle, self.id), self.ctx.handle)
            self.id = None


def hybrid_params(**kw) -> _lib.re_hybrid_params:
    p = _lib.re_hybrid_params()
    lib().re_hybrid_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def hybrid_trace(vol: "Volume", net: "Surrogate | None", o, d, **kw):
    """Hybrid trace (net given) or the exact-field march alone (net None):
    (t, hit, evals, samples)."""
    o, d = _f32c(o, 3), _f32c(d, 3)
    n = o.size // 3
    t = np.empty(n, np.float32)
    hit = np.empty(n, np.uint8)
    ev = np.empty(n, np.int32)
    sm = np.empty(n, np.int32)
    p = hybrid_params(**kw)
    check(lib().re_hybrid_trace(vol.ctx.handle, vol.id, net.id if net is not None else -1, C.byref(p), ptr(o),
                                ptr(d), n, ptr(t), ptr(hit), ptr(ev), ptr(sm), 0), vol.ctx.handle)
    return t, hit, ev, sm


def hybrid_trace_device(vol: "Volume", net, o_ptr, d_ptr, n, t_ptr, hit_ptr, sync=True, **kw):
    p = hybrid_params(**kw)
    flags = RE_DEVICE_PTRS | (0 if sync else RE_ASYNC)
    check(lib().re_hybrid_trace(vol.ctx.handle, vol.id, net.id if net is not None else -1, C.byref(p), o_ptr,
                                d_ptr, n, t_ptr, hit_ptr, None, None, flags), vol.ctx.handle)


def distance_volume(vol: "Volume") -> "Volume":
    """Signed Euclidean distance to the zero set of a field volume (re_sdf_distance_volume)."""
    vid = C.c_int32()
    check(lib().re_sdf_distance_volume(vol.ctx.handle, vol.id, C.byref(vid)), vol.ctx.handle)
    return Volume(vol.ctx, vid.value)


def sample_points(dist_vol: "Volume", n: int, seed: int = 1, near_band: float = 2 / 256) -> np.ndarray:
    out = np.empty((n, 4), np.float32)
    check(lib().re_mlp_sample(dist_vol.ctx.handle, dist_vol.id, n, seed, near_band, ptr(out), 0),
          dist_vol.ctx.handle)
    return out


BATCH_ITEM_DT = np.dtype([("kind", "<i4"), ("target", "<i4"), ("begin", "<i8"), ("count", "<i4"), ("slot", "<i4")])


def batch_plan(scene_ranges, tiles, n_rays: int, item_rays: int = 256):
    """re_batch_plan: the work items a batch would be cut into, on the host.
    scene_ranges: (ray_begin, ray_count) per scene; tiles: (mode, ray_begin,
    ray_count) per tile.  Returns (items as BATCH_ITEM_DT, remaining item
    count per slot)."""
    sc = (_lib.re_batch_scene * max(len(scene_ranges), 1))()
    for i, (b, c) in enumerate(scene_ranges):
        sc[i].tree, sc[i].ray_begin, sc[i].ray_count = None, int(b), int(c)
    tl = (_lib.re_batch_tile * max(len(tiles), 1))()
    for i, (mode, b, c) in enumerate(tiles):
        tl[i].vol_id, tl[i].mode, tl[i].ray_begin, tl[i].ray_count, tl[i].mlp_id = -1, int(mode), int(b), int(c), -1
    n = C.c_int64()
    rem = np.zeros(max(len(scene_ranges) + len(tiles), 1), np.int32)
    check(lib().re_batch_plan(sc, len(scene_ranges), tl, len(tiles), int(n_rays), int(item_rays), None, 0, C.byref(n),
                              ptr(rem)))
    items = np.zeros(n.value, BATCH_ITEM_DT)
    check(lib().re_batch_plan(sc, len(scene_ranges), tl, len(tiles), int(n_rays), int(item_rays), ptr(items),
                              n.value, C.byref(n), None))
    return items, rem[:len(scene_ranges) + len(tiles)]


class Batch:
    """A batched mixed workload captured in a hipGraph (re_batch_*).

    scenes: list of (KDTree, ray_begin, ray_count); tiles: list of
    (Volume, mode, ray_begin, ray_count[, Surrogate]) with mode 0 exact march,
    1 reference sphere trace, 2 surrogate hybrid (with the tile's own network,
    or `surrogate` when the tile names none)."""

    def __init__(self, ctx: Context, scenes, tiles, ray_o, ray_d, item_rays=256, block=256, surrogate=None,
                 trace=None, sdf=None, hybrid=None, march=None, side_share=0.0):
        self.ctx = ctx
        o, d = _f32c(ray_o, 3), _f32c(ray_d, 3)
        self.n_rays = o.size // 3
        sc = (_lib.re_batch_scene * max(len(scenes), 1))()
        for i, (tree, b, c) in enumerate(scenes):
            sc[i].tree, sc[i].ray_begin, sc[i].ray_count = tree.handle, int(b), int(c)
        tl = (_lib.re_batch_tile * max(len(tiles), 1))()
        for i, tile in enumerate(tiles):
            vol, mode, b, c = tile[:4]
            net = tile[4] if len(tile) > 4 else None
            tl[i].vol_id, tl[i].mode, tl[i].ray_begin, tl[i].ray_count = vol.id, int(mode), int(b), int(c)
            tl[i].mlp_id = net.id if net is not None else -1
        cfg = _lib.re_batch_config()
        lib().re_batch_default_config(C.byref(cfg))
        cfg.item_rays, cfg.block = int(item_rays), int(block)
        cfg.side_share = float(side_share)
        cfg.mlp_id = surrogate.id if surrogate is not None else -1
        if trace:
            cfg.trace = _trace_params(**trace)
        if sdf:
            cfg.sdf = sdf_params(**sdf)
        if hybrid:
            cfg.hybrid = hybrid_params(**hybrid)
        if march:
            cfg.march = hybrid_params(**march)
        bid = C.c_int32()
        check(lib().re_batch_create(ctx.handle, sc, len(scenes), tl, len(tiles), ptr(o), ptr(d), self.n_rays,
                                    C.byref(cfg), C.byref(bid)), ctx.handle)
        self.id = bid.value
        self.n_slots = len(scenes) + len(tiles)

    def run(self, sync=True):
        check(lib().re_batch_run(self.ctx.handle, self.id, 0 if sync else RE_ASYNC), self.ctx.handle)

    def results(self):
        p = np.empty(self.n_rays, np.int32)
        t = np.empty(self.n_rays, np.float32)
        check(lib().re_batch_results(self.ctx.handle, self.id, ptr(p), ptr(t)), self.ctx.handle)
        return p, t

    def completion_ms(self):
        ms = np.empty(self.n_slots, np.float32)
        check(lib().re_batch_completion_ms(self.ctx.handle, self.id, ptr(ms)), self.ctx.handle)
        return ms

    def items(self):
        n_items = C.c_int64()
        check(lib().re_batch_info(self.ctx.handle, self.id, None, C.byref(n_items), None), self.ctx.handle)
        return n_items.value

    def release(self):
        if self.id is not None:
            check(lib().re_batch_release(self.ctx.handle, self.id), self.ctx.handle)
            self.id = None


def _trace_params(max_leaf_visits=50, one_sided=True, secondary=False, block=0, variant=4):
    return _lib.re_trace_params(int(max_leaf_visits), int(bool(one_sided)), int(bool(secondary)), int(block),
                                int(variant))


class Scene:
    def __init__(self, ctx: Context, sid: int, objects):
        self.ctx, self.id, self.objects = ctx, sid, objects

    def bytes(self) -> int:
        v = C.c_int64()
        check(lib().re_scene_bytes(self.ctx.handle, self.id, C.byref(v)), self.ctx.handle)
        return v.value

    def export(self):
        sizes = (C.c_int64 * 4)()
        check(lib().re_scene_info(self.ctx.handle, self.id, sizes), self.ctx.handle)
        nodes = np.zeros(sizes[0], NODE_DT)
        leaves = np.zeros(sizes[1], LEAF_DT)
        tris = np.zeros(sizes[2], TRI_DT)
        objs = np.zeros(sizes[3], OBJ_DT)
        check(lib().re_scene_export(self.ctx.handle, self.id, ptr(nodes), ptr(leaves), ptr(tris), ptr(objs)),
              self.ctx.handle)
        return nodes, leaves, tris, objs

    def trace_closest(self, o, d, with_obj=False, **params):
        """Closest hit of each ray: (prim_id int32, t float32[, obj_id])."""
        o, d = _f32c(o, 3), _f32c(d, 3)
        n = o.size // 3
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        obj = np.empty(n, np.int32) if with_obj else None
        p = _trace_params(**params)
        check(lib().re_trace_closest(self.ctx.handle, self.id, ptr(o), ptr(d), n, C.byref(p), ptr(prim), ptr(t),
                                     ptr(obj), 0), self.ctx.handle)
        return (prim, t, obj) if with_obj else (prim, t)

    def occluded(self, o, d, **params):
        """Any-hit (shadow) query: uint8 per ray, 1 when some object is hit
        (trace_shadow; traverses with secondary=True unless given)."""
        o, d = _f32c(o, 3), _f32c(d, 3)
        n = o.size // 3
        out = np.empty(n, np.uint8)
        params.setdefault("secondary", True)
        p = _trace_params(**params)
        check(lib().re_trace_occluded(self.ctx.handle, self.id, ptr(o), ptr(d), n, C.byref(p), ptr(out), 0),
              self.ctx.handle)
        return out

    def occluded_device(self, o_ptr: int, d_ptr: int, n: int, out_ptr: int, sync=True, **params):
        params.setdefault("secondary", True)
        p = _trace_params(**params)
        flags = RE_DEVICE_PTRS | (0 if sync else RE_ASYNC)
        check(lib().re_trace_occluded(self.ctx.handle, self.id, o_ptr, d_ptr, n, C.byref(p), out_ptr, flags),
              self.ctx.handle)

    def trace_closest_device(self, o_ptr: int, d_ptr: int, n: int, prim_ptr: int, t_ptr: int, sync=True, **params):
        """Same on device-resident buffers (raw device pointers)."""
        p = _trace_params(**params)
        flags = RE_DEVICE_PTRS | (0 if sync else RE_ASYNC)
        check(lib().re_trace_closest(self.ctx.handle, self.id, o_ptr, d_ptr, n, C.byref(p), prim_ptr, t_ptr, None,
                                     flags), self.ctx.handle)

    def trace_closest_hinted_device(self, o_ptr: int, d_ptr: int, n: int, prim_ptr: int, t_ptr: int,
                                    start_leaf_ptr: int = 0, hit_leaf_ptr: int = 0, obj_ptr: int = 0, counts=False,
                                    sync=True, **params):
        """re_trace_closest_hinted on device buffers (raw pointers, 0 for
        none): walks may start at start_leaf's leaves, hit_leaf receives
        each ray's hit leaf.  Variants 4 and 5.  Returns the work counters
        [entries, leaves, tris, rays_hit] when counts is true."""
        p = _trace_params(**params)
        cnt = (C.c_uint64 * 4)() if counts else None
        flags = RE_DEVICE_PTRS | (0 if sync else RE_ASYNC)
        check(lib().re_trace_closest_hinted(self.ctx.handle, self.id, o_ptr, d_ptr, n, C.byref(p),
                                            start_leaf_ptr or None, prim_ptr, t_ptr, obj_ptr or None,
                                            hit_leaf_ptr or None, cnt, flags), self.ctx.handle)
        return list(cnt) if counts else None

    def set_attributes(self, normals, uvs=None):
        """Per-prim vertex normals ((n_prims, 9): n0, n1, n2) and optional
        vertex uvs ((n_prims, 6): uv0, uv1, uv2) for hit_attributes."""
        nr = _f32c(normals, 9)
        uv = None if uvs is None else _f32c(uvs, 6)
        check(lib().re_scene_set_attributes(self.ctx.handle, self.id, ptr(nr), ptr(uv), nr.size // 9), self.ctx.handle)

    def hit_attributes(self, o, d, prim, obj=None, **params):
        """(normal (n,3), uv (n,2), bary (n,2), hit_point (n,3)) of traced
        rays with their prim ids (and object ids for several objects)."""
        o, d = _f32c(o, 3), _f32c(d, 3)
        n = o.size // 3
        prim = np.ascontiguousarray(prim, np.int32)
        obj_a = None if obj is None else np.ascontiguousarray(obj, np.int32)
        nrm = np.empty((n, 3), np.float32)
        uv = np.empty((n, 2), np.float32)
        bary = np.empty((n, 2), np.float32)
        hp = np.empty((n, 3), np.float32)
        p = _trace_params(**params)
        check(lib().re_hit_attributes(self.ctx.handle, self.id, ptr(o), ptr(d), n, C.byref(p), ptr(prim), ptr(obj_a),
                                      ptr(nrm), ptr(uv), ptr(bary), ptr(hp), 0), self.ctx.handle)
        return nrm, uv, bary, hp

    def trace_counts(self, o, d, **params):
        """(prim, t, counts) with counts = [entries, leaves, tris, rays_hit] summed over rays."""
        o, d = _f32c(o, 3), _f32c(d, 3)
        n = o.size // 3
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        cnt = (C.c_uint64 * 4)()
        p = _trace_params(**params)
        check(lib().re_trace_counts(self.ctx.handle, self.id, ptr(o), ptr(d), n, C.byref(p), ptr(prim), ptr(t),
                                    cnt, 0), self.ctx.handle)
        return prim, t, list(cnt)

    def trace_status(self, o, d, **params):
        """(prim, t, status) with each ray's status byte (re_trace_status): the
        end of its last walk (0 normal, 1 visit cap, 2 leaf slab miss, 3 no
        exit face, 4 rope -1, 5 root box missed, 6 repeated state) | 0x80
        for non-finite input."""
        o, d = _f32c(o, 3), _f32c(d, 3)
        n = o.size // 3
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        st = np.empty(n, np.uint8)
        p = _trace_params(**params)
        check(lib().re_trace_status(self.ctx.handle, self.id, ptr(o), ptr(d), n, C.byref(p), ptr(prim), ptr(t),
                                    ptr(st), 0), self.ctx.handle)
        return prim, t, st

    def trace_counts_device(self, o_ptr: int, d_ptr: int, n: int, prim_ptr: int, t_ptr: int, **params):
        cnt = (C.c_uint64 * 4)()
        p = _trace_params(**params)
        check(lib().re_trace_counts(self.ctx.handle, self.id, o_ptr, d_ptr, n, C.byref(p), prim_ptr, t_ptr, cnt,
                                    RE_DEVICE_PTRS), self.ctx.handle)
        return list(cnt)

    def release(self):
        if self.id is not None:
            check(lib().re_scene_release(self.ctx.handle, self.id), self.ctx.handle)
            self.id = None
