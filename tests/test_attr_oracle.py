"""Pins of the oracle's hit attributes (orc_hit_attributes) against closed forms.

The reference shades a hit with the Moller-Trumbore barycentrics of the
accepted triangle: the normal is normalize(w n0 + u n1 + v n2) with
w = 1 - (u + v) (gpu_get_tri_normal, RayEngine/src/cuda-src/ray_functions.cuh:146-155),
the texture coordinate w' uv0 + u uv1 + v uv2 (get_uvs, :157-166) or (u, v)
without vertex uvs (:206-213), and the hit point ro + t rd (:204).  Each case
below fixes the answer by geometry: a ray aimed at v0 + a e1 + b e2 has
barycentrics (a, b); the interpolated uvs and normals there are computed in
FP64; flat vertex normals give the face normal; object transforms are undone
with the inverse rotation.  A swapped (u, v), a wrong weight for n0, or a
transform with the wrong sign fails one of them.
"""
import numpy as np
import pytest

import oracle


def _object(loc=(0.0, 0.0, 0.0), rot=0.0):
    """One object record (OBJ_DT): rotation about Y by rot, then loc."""
    from rayengine_amd import OBJ_DT
    ob = np.zeros(1, OBJ_DT)
    ob["loc"] = loc
    ob["sin_y"] = np.float32(np.sin(rot))
    ob["cos_y"] = np.float32(np.cos(rot))
    return ob


def _identity():
    return _object()


def _random_tris(rng, n):
    v0 = rng.uniform(-1, 1, (n, 3))
    e1 = rng.uniform(-1, 1, (n, 3))
    e2 = rng.uniform(-1, 1, (n, 3))
    return v0, e1, e2


def _aimed_rays(rng, v0, e1, e2, a, b, dist=2.0):
    """Rays from the front side (the side a one-sided test accepts) aimed at
    v0 + a e1 + b e2."""
    p = v0 + a[:, None] * e1 + b[:, None] * e2
    n = np.cross(e1, e2)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    # one-sided Moller-Trumbore accepts det = dot(e1, cross(d, e2)) > 0,
    # i.e. dot(d, cross(e1, e2)) < 0: come from the side n points to
    tilt = rng.uniform(-0.3, 0.3, (len(p), 3))
    d = -(n + tilt)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = p - dist * d
    return o, d, p


def test_flat_normals_give_face_normal_and_bary_hit_point():
    rng = np.random.default_rng(3)
    n = 400
    v0, e1, e2 = _random_tris(rng, n)
    a = rng.uniform(0.05, 0.6, n)
    b = rng.uniform(0.05, 0.9 - a)
    o, d, p = _aimed_rays(rng, v0, e1, e2, a, b)
    face = np.cross(e1, e2)
    face /= np.linalg.norm(face, axis=1, keepdims=True)
    geom = np.concatenate([v0, e1, e2], axis=1).astype(np.float32)
    nrm = np.tile(face, 3).astype(np.float32)
    prim = np.arange(n, dtype=np.int32)
    normal, uv, bary, hp = oracle.hit_attributes(geom, nrm, None, _identity(), o, d, prim)
    assert np.abs(normal - face).max() < 2e-6
    assert np.abs(np.linalg.norm(normal, axis=1) - 1).max() < 1e-6
    assert np.abs(bary[:, 0] - a).max() < 1e-4 and np.abs(bary[:, 1] - b).max() < 1e-4
    # without vertex uvs the texture coordinate is the barycentric pair
    assert np.array_equal(uv, bary)
    assert np.abs(hp - p).max() < 1e-4


def test_interpolated_normals_and_uvs_match_fp64():
    rng = np.random.default_rng(11)
    n = 500
    v0, e1, e2 = _random_tris(rng, n)
    a = rng.uniform(0.02, 0.7, n)
    b = rng.uniform(0.02, 0.95 - a)
    o, d, p = _aimed_rays(rng, v0, e1, e2, a, b)
    n0, n1, n2 = (rng.normal(size=(n, 3)) for _ in range(3))
    c0, c1, c2 = (rng.uniform(0, 1, (n, 2)) for _ in range(3))
    geom = np.concatenate([v0, e1, e2], axis=1).astype(np.float32)
    nrm = np.concatenate([n0, n1, n2], axis=1).astype(np.float32)
    uvs = np.concatenate([c0, c1, c2], axis=1).astype(np.float32)
    prim = np.arange(n, dtype=np.int32)
    normal, uv, bary, hp = oracle.hit_attributes(geom, nrm, uvs, _identity(), o, d, prim)
    # FP64 closed forms at the aimed barycentrics, from the FP32 inputs
    f = lambda x: x.astype(np.float32).astype(np.float64)
    w = 1 - a - b
    want_n = w[:, None] * f(n0) + a[:, None] * f(n1) + b[:, None] * f(n2)
    want_n /= np.linalg.norm(want_n, axis=1, keepdims=True)
    want_uv = w[:, None] * f(c0) + a[:, None] * f(c1) + b[:, None] * f(c2)
    # the normal's error follows the barycentrics' over |interpolated normal|
    scale = np.linalg.norm(w[:, None] * f(n0) + a[:, None] * f(n1) + b[:, None] * f(n2), axis=1)
    ok = scale > 0.2
    assert ok.sum() > 400
    assert np.abs(normal - want_n)[ok].max() < 2e-3
    assert np.abs(uv - want_uv).max() < 5e-4
    assert np.abs(hp - p).max() < 1e-4
    # the weight of vertex 0 is 1 - (u + v): swapping n0 and n1 moves the answer
    swapped = np.concatenate([n1, n0, n2], axis=1).astype(np.float32)
    normal_s, _, _, _ = oracle.hit_attributes(geom, swapped, uvs, _identity(), o, d, prim)
    assert np.abs(normal_s - want_n)[ok].max() > 0.1


def test_vertex_hits_give_vertex_attributes():
    """Rays through the three corners pick up one vertex's normal and uv."""
    v0 = np.zeros((3, 3))
    e1 = np.tile([1.0, 0, 0], (3, 1))
    e2 = np.tile([0, 1.0, 0], (3, 1))
    geom = np.concatenate([v0, e1, e2], axis=1).astype(np.float32)
    n0, n1, n2 = np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0])
    nrm = np.tile(np.concatenate([n0, n1, n2]), (3, 1)).astype(np.float32)
    uvs = np.tile([0.1, 0.2, 0.3, 0.4, 0.5, 0.6], (3, 1)).astype(np.float32)
    corners = np.array([[1e-7, 1e-7, 0], [1 - 2e-7, 1e-7, 0], [1e-7, 1 - 2e-7, 0]])
    o = corners + [0, 0, 1]
    d = np.tile([0, 0, -1.0], (3, 1))
    normal, uv, bary, hp = oracle.hit_attributes(geom, nrm, uvs, _identity(), o, d, np.arange(3, dtype=np.int32))
    assert np.abs(normal - np.stack([n0, n1, n2])).max() < 1e-5
    assert np.abs(uv - np.array([[0.1, 0.2], [0.3, 0.4], [0.5, 0.6]])).max() < 1e-5
    assert np.abs(hp - corners).max() < 1e-6


def test_misses_are_zero_and_wrong_prim_is_nan():
    geom = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0], [5, 5, 5, 1, 0, 0, 0, 1, 0]], np.float32)
    nrm = np.tile([0, 0, 1.0] * 3, (2, 1)).astype(np.float32)
    o = np.array([[0.2, 0.2, 1], [0.2, 0.2, 1]], np.float32)
    d = np.array([[0, 0, -1], [0, 0, -1]], np.float32)
    normal, uv, bary, hp = oracle.hit_attributes(geom, nrm, None, _identity(), o, d, np.array([-1, 1], np.int32))
    for x in (normal, uv, bary, hp):
        assert np.all(x[0] == 0)
        assert np.all(np.isnan(x[1]))


@pytest.mark.parametrize("rot", [0.7, -2.3])
def test_object_transform_is_undone(rot):
    """A ray built in object space and taken to world space by the inverse of
    the object transform (rotate about Y, then translate) hits at the same
    barycentrics; the hit point is reported in object space."""
    rng = np.random.default_rng(5)
    n = 200
    v0, e1, e2 = _random_tris(rng, n)
    a = rng.uniform(0.05, 0.6, n)
    b = rng.uniform(0.05, 0.9 - a)
    o_obj, d_obj, p = _aimed_rays(rng, v0, e1, e2, a, b)
    loc = np.array([0.4, -0.3, 0.25])
    c, s = np.cos(rot), np.sin(rot)
    R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])  # object-space = R @ world + loc
    o_w = (o_obj - loc) @ R  # R^T applied to rows
    d_w = d_obj @ R
    objs = _object(loc, rot)
    geom = np.concatenate([v0, e1, e2], axis=1).astype(np.float32)
    nrm = np.tile([0, 0, 1.0] * 3, (n, 1)).astype(np.float32)
    normal, uv, bary, hp = oracle.hit_attributes(geom, nrm, None, objs, o_w, d_w, np.arange(n, dtype=np.int32),
                                                 obj=np.zeros(n, np.int32))
    assert not np.isnan(bary).any()
    assert np.abs(bary[:, 0] - a).max() < 2e-4 and np.abs(bary[:, 1] - b).max() < 2e-4
    assert np.abs(hp - p).max() < 2e-4
    # the opposite rotation does not reproduce the hits
    _, _, bary_w, _ = oracle.hit_attributes(geom, nrm, None, _object(loc, -rot), o_w, d_w,
                                             np.arange(n, dtype=np.int32))
    bad = np.isnan(bary_w[:, 0]) | (np.abs(bary_w[:, 0] - a) > 1e-2)
    assert bad.mean() > 0.5
