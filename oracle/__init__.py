"""CPU oracle of the RayEngine ray-query path - TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker; the product library (rayengine_amd) never imports it.

The C++ restatement lives in oracle/oracle.cpp; each function there cites the
reference routine it follows and names what pins it.  Build with
`python oracle/build.py`.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_P = C.c_void_p
_SIGS = {
    "orc_trace": (None, [_P, _P, _P, _P, C.c_int, _P, _P, C.c_int64, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P,
                         C.c_int, C.c_int]),
    "orc_brute": (None, [_P, C.c_int64, _P, _P, _P, _P, C.c_int64, C.c_int, _P, _P, C.c_int]),
    "orc_occluded": (None, [_P, _P, _P, _P, C.c_int, _P, _P, C.c_int64, C.c_int, C.c_int, C.c_int, _P, C.c_int]),
    "orc_brute_f64": (None, [_P, C.c_int64, _P, _P, C.c_int64, C.c_int, _P, _P, C.c_int]),
    "orc_ray_tri": (C.c_int, [_P, _P, _P, _P, _P, C.c_int, _P, _P]),
    "orc_hit_attributes": (None, [_P, _P, _P, _P, _P, _P, C.c_int64, _P, _P, C.c_int, C.c_int, _P, _P, _P, _P]),
    "orc_ray_box": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "orc_neighbour_face": (C.c_int, [_P, _P, _P]),
    "orc_perlin_table": (None, [C.c_uint32, _P, _P]),
    "orc_mt19937_nth": (C.c_uint32, [C.c_uint32, C.c_int64]),
    "orc_grad_dot": (C.c_float, [C.c_uint32, C.c_float, C.c_float, C.c_float]),
    "orc_noise_eval": (C.c_float, [_P, C.c_float, C.c_float, C.c_float]),
    "orc_noise_eval_f64": (C.c_double, [_P, C.c_double, C.c_double, C.c_double]),
    "orc_exact_march": (None, [_P, _P, _P, _P, _P, C.c_int64, C.c_float, C.c_float, C.c_int, C.c_int, _P, _P, _P,
                               C.c_int]),
    "orc_hybrid_trace": (None, [_P, _P, _P, _P, C.c_float, _P, _P, C.c_int64, C.c_float, C.c_float, C.c_float,
                                C.c_float, C.c_float, C.c_int, C.c_int, _P, _P, _P, _P, C.c_int]),
    "orc_mlp_eval_f32": (C.c_float, [_P, C.c_float, C.c_float, C.c_float]),
    "orc_noise_fill": (None, [_P, _P, _P, _P, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, _P, C.c_int]),
    "orc_noise_fill_f64": (None, [_P, _P, _P, _P, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, _P, C.c_int]),
    "orc_trilinear": (C.c_float, [_P, _P, _P, C.c_float, C.c_float, C.c_float]),
    "orc_sky": (None, [_P, _P, C.c_float, _P, C.c_int64, _P]),
    "orc_sky_f64": (None, [_P, _P, C.c_float, _P, C.c_int64, _P]),
    "orc_sky_tone": (C.c_float, [C.c_float]),
    "orc_frame_shade": (None, [_P, _P, _P, C.c_int32, C.c_int32, _P, _P]),
    "orc_volume_edit": (None, [_P, _P, _P, _P, C.c_int, _P, _P, C.c_float, C.c_int, C.c_float]),
    "orc_sdf_trace": (None, [_P, _P, _P, _P, _P, C.c_int64, C.c_int, C.c_float, C.c_float, C.c_int, _P, _P, _P, _P,
                             C.c_int]),
    "orc_sdf_trace_f64": (None, [_P, _P, _P, _P, _P, C.c_int64, C.c_int, C.c_float, C.c_float, C.c_int, _P, _P,
                                 _P, C.c_int]),
}

# Terrain defaults: 256^3 over extent 1, 5 octaves, lacunarity 2, gain 0.5,
# base frequency 4, surface offset 0.4, permutation seed 2016.
TERRAIN = dict(dim=(256, 256, 256), spacing=(1 / 256, 1 / 256, 1 / 256), octaves=5, lacunarity=2.0, gain=0.5,
               base_freq=4.0, y_offset=0.4)

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            from . import build as _b  # noqa: F401  (oracle/build.py)
            _b.build()
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _f(a, k=None):
    a = np.ascontiguousarray(a, np.float32)
    return a


def trace(nodes, leaves, tris, objs, o, d, cap=50, two_sided=False, secondary=False, counts=False, threads=0,
          cut=False):
    """FP32 rope traversal; returns (prim, t[, counts]).  counts is (n, 4)
    uint64: node entries, leaf records, triangle tests, and how the walk ended
    (0 t_entry >= t_exit, 1 cap, 2 leaf slab miss, 3 no face within K_EPSILON,
    4 rope -1, 5 root box missed, 6 walk state repeated).  cut stops a walk
    whose state repeats (same results; the counters of the shortened walk)."""
    o, d = _f(o), _f(d)
    n = o.size // 3
    prim = np.empty(n, np.int32)
    t = np.empty(n, np.float32)
    cnt = np.empty((n, 4), np.uint64) if counts else None
    lib().orc_trace(_p(nodes), _p(leaves), _p(tris), _p(objs), len(objs), _p(o), _p(d), n, cap, int(two_sided),
                    int(secondary), _p(prim), _p(t), None, _p(cnt), threads, int(bool(cut)))
    return (prim, t, cnt) if counts else (prim, t)


def occluded(nodes, leaves, tris, objs, o, d, cap=50, two_sided=False, secondary=True, threads=0):
    """trace_shadow: uint8 per ray, 1 when some object's walk detects a hit."""
    o, d = _f(o), _f(d)
    n = o.size // 3
    out = np.empty(n, np.uint8)
    lib().orc_occluded(_p(nodes), _p(leaves), _p(tris), _p(objs), len(objs), _p(o), _p(d), n, cap, int(two_sided),
                       int(secondary), _p(out), threads)
    return out


def brute(tri_v, root_min, root_max, o, d, two_sided=False, threads=0):
    tri_v = _f(tri_v)
    o, d = _f(o), _f(d)
    n = o.size // 3
    prim = np.empty(n, np.int32)
    t = np.empty(n, np.float32)
    rmin, rmax = _f(root_min), _f(root_max)
    lib().orc_brute(_p(tri_v), tri_v.size // 9, _p(rmin), _p(rmax), _p(o), _p(d), n, int(two_sided), _p(prim),
                    _p(t), threads)
    return prim, t


def brute_f64(tri_v, o, d, two_sided=False, threads=0):
    tri_v = _f(tri_v)
    o, d = _f(o), _f(d)
    n = o.size // 3
    prim = np.empty(n, np.int32)
    t = np.empty(n, np.float64)
    lib().orc_brute_f64(_p(tri_v), tri_v.size // 9, _p(o), _p(d), n, int(two_sided), _p(prim), _p(t), threads)
    return prim, t


def ray_tri(o, d, v0, e1, e2, two_sided=False):
    t = np.zeros(1, np.float32)
    uv = np.zeros(2, np.float32)
    args = [_f(x) for x in (o, d, v0, e1, e2)]
    hit = lib().orc_ray_tri(*[_p(a) for a in args], int(two_sided), _p(t), _p(uv))
    return bool(hit), float(t[0]), uv


def hit_attributes(geom, normals, uvs, objs, o, d, prim, obj=None, two_sided=False, secondary=False):
    """(normal (n,3), uv (n,2), bary (n,2), hit_point (n,3)) of traced rays:
    geom (n_prims, 9) v0, e1, e2; normals (n_prims, 9) n0, n1, n2; uvs
    (n_prims, 6) or None; objs an OBJ_DT array."""
    geom, normals = _f(geom), _f(normals)
    uvs = None if uvs is None else _f(uvs)
    o, d = _f(o), _f(d)
    n = o.size // 3
    prim = np.ascontiguousarray(prim, np.int32)
    obj = None if obj is None else np.ascontiguousarray(obj, np.int32)
    nrm = np.empty((n, 3), np.float32)
    uv = np.empty((n, 2), np.float32)
    bary = np.empty((n, 2), np.float32)
    hp = np.empty((n, 3), np.float32)
    lib().orc_hit_attributes(_p(geom), _p(normals), _p(uvs), _p(objs), _p(o), _p(d), n, _p(prim), _p(obj),
                             int(two_sided), int(secondary), _p(nrm), _p(uv), _p(bary), _p(hp))
    return nrm, uv, bary, hp


def ray_box(bmin, bmax, o, d):
    tn = np.zeros(1, np.float32)
    tf = np.zeros(1, np.float32)
    args = [_f(x) for x in (bmin, bmax, o, d)]
    hit = lib().orc_ray_box(*[_p(a) for a in args], _p(tn), _p(tf))
    return bool(hit), tn[0], tf[0]


def neighbour_face(bmin, bmax, p):
    args = [_f(x) for x in (bmin, bmax, p)]
    return lib().orc_neighbour_face(*[_p(a) for a in args])


def perlin_table(seed=2016):
    perm = np.zeros(512, np.uint32)
    grads = np.zeros((256, 3), np.float32)
    lib().orc_perlin_table(seed, _p(perm), _p(grads))
    return perm, grads


def mt19937_nth(seed, k):
    return int(lib().orc_mt19937_nth(seed, k))


# ------------------------------------------------------------------ terrain

def grad_dot(h, x, y, z):
    return lib().orc_grad_dot(h, x, y, z)


def noise_eval(perm, x, y, z):
    perm = np.ascontiguousarray(perm, np.uint32)
    return np.float32(lib().orc_noise_eval(_p(perm), x, y, z))


def noise_eval_f64(perm, x, y, z):
    perm = np.ascontiguousarray(perm, np.uint32)
    return lib().orc_noise_eval_f64(_p(perm), x, y, z)


def _terrain_args(perm, kw):
    p = dict(TERRAIN)
    p.update(kw)
    dim = np.array(p["dim"], np.int32)
    sp = np.array(p["spacing"], np.float32)
    origin = np.array(p.get("origin", (0.0, 0.0, 0.0)), np.float32)
    perm = np.ascontiguousarray(perm, np.uint32)
    return perm, dim, sp, origin, p


def noise_fill(perm, threads=0, **kw):
    """FP32 terrain grid as (dz, dy, dx)."""
    perm, dim, sp, origin, p = _terrain_args(perm, kw)
    out = np.empty((dim[2], dim[1], dim[0]), np.float32)
    lib().orc_noise_fill(_p(perm), _p(dim), _p(sp), _p(origin), p["octaves"], p["lacunarity"], p["gain"], p["base_freq"],
                         p["y_offset"], _p(out), threads)
    return out


def noise_fill_f64(perm, threads=0, **kw):
    perm, dim, sp, origin, p = _terrain_args(perm, kw)
    out = np.empty((dim[2], dim[1], dim[0]), np.float64)
    lib().orc_noise_fill_f64(_p(perm), _p(dim), _p(sp), _p(origin), p["octaves"], p["lacunarity"], p["gain"], p["base_freq"],
                             p["y_offset"], _p(out), threads)
    return out


def trilinear(grid, spacing, x, y, z):
    g = np.ascontiguousarray(grid, np.float32)
    dim = np.array([g.shape[2], g.shape[1], g.shape[0]], np.int32)
    sp = np.array(spacing, np.float32)
    return np.float32(lib().orc_trilinear(_p(g), _p(dim), _p(sp), x, y, z))


def sdf_trace(grid, spacing, o, d, max_steps=256, eps=1e-3, step_scale=0.8, step_rule=0, normals=False, threads=0):
    """FP32 sphere trace: (t, hit, steps[, normal])."""
    g = np.ascontiguousarray(grid, np.float32)
    dim = np.array([g.shape[2], g.shape[1], g.shape[0]], np.int32)
    sp = np.array(spacing, np.float32)
    o, d = _f(o), _f(d)
    n = o.size // 3
    t = np.empty(n, np.float32)
    hit = np.empty(n, np.uint8)
    steps = np.empty(n, np.int32)
    nrm = np.empty((n, 3), np.float32) if normals else None
    lib().orc_sdf_trace(_p(g), _p(dim), _p(sp), _p(o), _p(d), n, max_steps, eps, step_scale, step_rule, _p(t),
                        _p(hit), _p(steps), _p(nrm), threads)
    return (t, hit, steps, nrm) if normals else (t, hit, steps)


def sdf_trace_f64(grid64, spacing, o, d, max_steps=256, eps=1e-3, step_scale=0.8, step_rule=0, threads=0):
    g = np.ascontiguousarray(grid64, np.float64)
    dim = np.array([g.shape[2], g.shape[1], g.shape[0]], np.int32)
    sp = np.array(spacing, np.float32)
    o, d = _f(o), _f(d)
    n = o.size // 3
    t = np.empty(n, np.float64)
    hit = np.empty(n, np.uint8)
    steps = np.empty(n, np.int32)
    lib().orc_sdf_trace_f64(_p(g), _p(dim), _p(sp), _p(o), _p(d), n, max_steps, eps, step_scale, step_rule, _p(t),
                            _p(hit), _p(steps), threads)
    return t, hit, steps


# ------------------------------------------------------------------ surrogate MLP

def _mlp_sigs():
    L = lib()
    for name, res, args in [
        ("orc_mlp_n_params", C.c_int, []),
        ("orc_mlp_features", None, [_P, C.c_int64, _P]),
        ("orc_mlp_forward", None, [_P, _P, C.c_int64, _P]),
        ("orc_mlp_loss_grad", C.c_double, [_P, _P, _P, C.c_int64, C.c_float, _P]),
        ("orc_adam", None, [_P, _P, _P, _P, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float]),
    ]:
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    return L


def mlp_n_params():
    return _mlp_sigs().orc_mlp_n_params()


def mlp_features(pts):
    pts = _f(pts)
    n = pts.size // 3
    out = np.empty((n, 99), np.float64)
    _mlp_sigs().orc_mlp_features(_p(pts), n, _p(out))
    return out


def mlp_forward(params, pts):
    params, pts = _f(params), _f(pts)
    n = pts.size // 3
    out = np.empty(n, np.float64)
    _mlp_sigs().orc_mlp_forward(_p(params), _p(pts), n, _p(out))
    return out


def mlp_loss_grad(params, pts, target, over_weight=1.0):
    params, pts, target = _f(params), _f(pts), _f(target)
    n = pts.size // 3
    g = np.empty(params.size, np.float64)
    loss = _mlp_sigs().orc_mlp_loss_grad(_p(params), _p(pts), _p(target), n, over_weight, _p(g))
    return loss, g


def adam(params, m, v, grad, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """In-place FP64 Adam step on float32 arrays."""
    for a in (params, m, v):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    grad = _f(grad)
    _mlp_sigs().orc_adam(_p(params), _p(m), _p(v), _p(grad), params.size, step, lr, b1, b2, eps)


def truth_trace(grid64, spacing, o, d, h_voxels=0.125, threads=0):
    """FP64 first zero crossing per ray (fine march + bisection): (t, hit)."""
    L = lib()
    f = L.orc_truth_trace
    f.restype = None
    f.argtypes = [_P, _P, _P, _P, _P, C.c_int64, C.c_double, _P, _P, C.c_int]
    g = np.ascontiguousarray(grid64, np.float64)
    dim = np.array([g.shape[2], g.shape[1], g.shape[0]], np.int32)
    sp = np.array(spacing, np.float32)
    o, d = _f(o), _f(d)
    n = o.size // 3
    t = np.empty(n, np.float64)
    hit = np.empty(n, np.uint8)
    f(_p(g), _p(dim), _p(sp), _p(o), _p(d), n, h_voxels, _p(t), _p(hit), threads)
    return t, hit


BRUSH_TYPES = {"sphere_add": 0, "sphere_subtract": 1, "cube_add": 2, "cube_subtract": 3}


def volume_edit(dist, mat, spacing, kind, center, extent, material=0.0, snap=False, smooth_k=0.1):
    """insert_sphere_to_texture on (dz, dy, dx) float32 distance and material
    grids; returns edited copies."""
    dist = np.ascontiguousarray(dist, np.float32).copy()
    mat = np.ascontiguousarray(mat, np.float32).copy()
    dim = np.array(dist.shape[::-1], np.int32)
    sp = np.asarray(spacing, np.float32)
    c = np.asarray(center, np.float32)
    e = np.asarray(extent, np.float32)
    lib().orc_volume_edit(_p(dist), _p(mat), _p(dim), _p(sp), BRUSH_TYPES[kind], _p(c), _p(e), float(material),
                          int(bool(snap)), float(smooth_k))
    return dist, mat


class FrameParams(C.Structure):
    """re_frame_params layout."""
    _fields_ = [("sun_dir", C.c_float * 3), ("earth_radius", C.c_float), ("atmosphere_radius", C.c_float),
                ("hr", C.c_float), ("hm", C.c_float), ("sky_origin", C.c_float * 3), ("base_color", C.c_float * 3),
                ("gamma", C.c_int32), ("vignetting", C.c_int32), ("vignetting_k", C.c_float)]


def frame_params(**kw):
    """The frame defaults (sun (0, 1, 0), Atmosphere radii and scale heights,
    sky origin (0, 6370e3, 3e5), no gamma or vignetting), overridden by
    keywords."""
    p = FrameParams()
    p.sun_dir[:] = [0.0, 1.0, 0.0]
    p.earth_radius, p.atmosphere_radius, p.hr, p.hm = 6360e3, 6420e3, 79945.0, 1200.0
    p.sky_origin[:] = [0.0, 6370e3, 300000.0]
    p.base_color[:] = [0.0, 0.0, 0.0]
    p.gamma, p.vignetting, p.vignetting_k = 0, 0, 0.25
    for k, v in kw.items():
        if isinstance(v, (tuple, list)):
            getattr(p, k)[:] = [float(x) for x in v]
        else:
            setattr(p, k, v)
    return p


ATM_BETA_R = np.float32([3.8e-6, 13.5e-6, 33.1e-6])
ATM_BETA_M = np.float32(21e-6)


def sky(params, dirs, beta_r=ATM_BETA_R, beta_m=ATM_BETA_M, f64=False):
    """compute_incident_light for each direction (before the tone curve);
    params is an re_frame_params-layout ctypes structure."""
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    br = np.ascontiguousarray(beta_r, np.float32)
    out = np.empty((len(d), 3), np.float64 if f64 else np.float32)
    fn = lib().orc_sky_f64 if f64 else lib().orc_sky
    fn(C.byref(params), _p(br), float(beta_m), _p(d), len(d), _p(out))
    return out


def sky_tone(c):
    return np.float32(lib().orc_sky_tone(float(c)))


def frame_shade(params, ray_d, hit, normal, width, height):
    d = np.ascontiguousarray(ray_d, np.float32)
    h = np.ascontiguousarray(hit, np.uint8)
    nrm = np.ascontiguousarray(normal, np.float32)
    out = np.empty((height, width), np.uint32)
    lib().orc_frame_shade(_p(d), _p(h), _p(nrm), int(width), int(height), C.byref(params), _p(out))
    return out


def exact_march(grid, spacing, o, d, k_exact, h_min, n_bisect, max_iters, threads=0):
    """(t, hit, samples) of the exact-field march over a (dz, dy, dx) FP32
    grid; k_exact and h_min in world units as the library applies them."""
    g = np.ascontiguousarray(grid, np.float32)
    dim = np.array([g.shape[2], g.shape[1], g.shape[0]], np.int32)
    sp = np.array(spacing, np.float32).reshape(3)
    o, d = _f(o), _f(d)
    n = o.size // 3
    t = np.empty(n, np.float32)
    hit = np.empty(n, np.uint8)
    s = np.empty(n, np.int32)
    lib().orc_exact_march(_p(g), _p(dim), _p(sp), _p(o), _p(d), n, k_exact, h_min, n_bisect, max_iters, _p(t),
                          _p(hit), _p(s), threads)
    return t, hit, s


def hybrid_world_params(spacing, tau_voxels, margin_voxels, k_exact, h_min_voxels, rejoin_voxels, n_bisect,
                        max_iters):
    """re_hybrid_params in world units, rounded as the library rounds them
    (float32 products with the x spacing)."""
    vox = np.float32(np.float32(spacing[0]) if np.ndim(spacing) else spacing)
    f = np.float32
    return dict(tau=float(f(tau_voxels) * vox), margin=float(f(margin_voxels) * vox), k_exact=float(f(k_exact)),
                h_min=float(f(h_min_voxels) * vox), rejoin=float(f(rejoin_voxels) * vox), n_bisect=int(n_bisect),
                max_iters=int(max_iters))


def hybrid_trace(grid, spacing, net_params, o, d, tau, margin, k_exact, h_min, rejoin, n_bisect, max_iters,
                 const_s=0.0, threads=0):
    """(t, hit, evals, samples) of the hybrid tracer with the network in FP32
    (net_params: the flat parameter vector) or, with net_params None, a
    network of constant value const_s; distances in world units."""
    g = np.ascontiguousarray(grid, np.float32)
    dim = np.array([g.shape[2], g.shape[1], g.shape[0]], np.int32)
    sp = np.array(spacing, np.float32).reshape(3)
    net = None if net_params is None else np.ascontiguousarray(net_params, np.float32)
    o, d = _f(o), _f(d)
    n = o.size // 3
    t = np.empty(n, np.float32)
    hit = np.empty(n, np.uint8)
    ev = np.empty(n, np.int32)
    sm = np.empty(n, np.int32)
    lib().orc_hybrid_trace(_p(g), _p(dim), _p(sp), _p(net), const_s, _p(o), _p(d), n, tau, margin, k_exact, h_min,
                           rejoin, n_bisect, max_iters, _p(t), _p(hit), _p(ev), _p(sm), threads)
    return t, hit, ev, sm


def mlp_eval_f32(params, pts):
    p = np.ascontiguousarray(params, np.float32)
    pts = np.asarray(pts, np.float32).reshape(-1, 3)
    return np.array([lib().orc_mlp_eval_f32(_p(p), float(x), float(y), float(z)) for x, y, z in pts], np.float32)
