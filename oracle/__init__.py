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
                         C.c_int]),
    "orc_brute": (None, [_P, C.c_int64, _P, _P, _P, _P, C.c_int64, C.c_int, _P, _P, C.c_int]),
    "orc_brute_f64": (None, [_P, C.c_int64, _P, _P, C.c_int64, C.c_int, _P, _P, C.c_int]),
    "orc_ray_tri": (C.c_int, [_P, _P, _P, _P, _P, C.c_int, _P, _P]),
    "orc_ray_box": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "orc_neighbour_face": (C.c_int, [_P, _P, _P]),
    "orc_perlin_table": (None, [C.c_uint32, _P, _P]),
    "orc_mt19937_nth": (C.c_uint32, [C.c_uint32, C.c_int64]),
}

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


def trace(nodes, leaves, tris, objs, o, d, cap=50, two_sided=False, secondary=False, counts=False, threads=0):
    """FP32 rope traversal; returns (prim, t[, counts (n,3) uint64])."""
    o, d = _f(o), _f(d)
    n = o.size // 3
    prim = np.empty(n, np.int32)
    t = np.empty(n, np.float32)
    cnt = np.empty((n, 3), np.uint64) if counts else None
    lib().orc_trace(_p(nodes), _p(leaves), _p(tris), _p(objs), len(objs), _p(o), _p(d), n, cap, int(two_sided),
                    int(secondary), _p(prim), _p(t), None, _p(cnt), threads)
    return (prim, t, cnt) if counts else (prim, t)


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
