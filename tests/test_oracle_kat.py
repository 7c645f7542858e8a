"""Pins the CPU oracle's primitives against closed forms and known answers.

Moller-Trumbore is checked against the textbook barycentric solution computed
in float64 by numpy's linear solver; the slab test against hand-derived
entry/exit distances (including the IEEE minNum behaviour for rays lying in a
slab plane); the rope face rule against its documented order; the Perlin
table generator against the C++ standard's mt19937 known answer.
"""
import numpy as np
import pytest

import oracle

TRI = dict(v0=[0, 0, 0], e1=[1, 0, 0], e2=[0, 1, 0])


def test_mt_front_face_hit():
    hit, t, uv = oracle.ray_tri([0.25, 0.25, 1.0], [0, 0, -1], **TRI)
    assert hit and t == 1.0 and np.allclose(uv, [0.25, 0.25])


def test_mt_back_face_culled_one_sided_hit_two_sided():
    hit, _, _ = oracle.ray_tri([0.25, 0.25, -1.0], [0, 0, 1], **TRI)
    assert not hit
    hit, t, _ = oracle.ray_tri([0.25, 0.25, -1.0], [0, 0, 1], two_sided=True, **TRI)
    assert hit and t == 1.0


def test_mt_edge_and_outside():
    assert oracle.ray_tri([0.5, 0.5, 1.0], [0, 0, -1], **TRI)[0]          # u + v == 1 is inside
    assert not oracle.ray_tri([0.6, 0.6, 1.0], [0, 0, -1], **TRI)[0]      # u + v > 1
    assert not oracle.ray_tri([-0.1, 0.5, 1.0], [0, 0, -1], **TRI)[0]     # u < 0
    assert oracle.ray_tri([0.0, 0.0, 1.0], [0, 0, -1], **TRI)[0]          # vertex v0


def test_mt_t_threshold_and_absolute_det_threshold():
    assert not oracle.ray_tri([0.25, 0.25, 5e-5], [0, 0, -1], **TRI)[0]   # t <= K_EPSILON
    assert oracle.ray_tri([0.25, 0.25, 2e-4], [0, 0, -1], **TRI)[0]
    # det = |e1||e2| = 2.5e-5 < 1e-4: the reference's absolute threshold rejects
    small = dict(v0=[0, 0, 0], e1=[0.005, 0, 0], e2=[0, 0.005, 0])
    assert not oracle.ray_tri([0.001, 0.001, 1.0], [0, 0, -1], **small)[0]


def test_mt_matches_textbook_barycentrics():
    rng = np.random.default_rng(7)
    agree = 0
    for _ in range(2000):
        v0 = rng.uniform(-1, 1, 3)
        e1 = rng.uniform(-1, 1, 3)
        e2 = rng.uniform(-1, 1, 3)
        o = rng.uniform(-3, 3, 3)
        target = v0 + rng.uniform(0, 0.7) * e1 + rng.uniform(0, 0.3) * e2
        d = target - o
        d /= np.linalg.norm(d)
        f32 = [np.float32(x) for x in (o, d, v0, e1, e2)]
        o32, d32, v032, e132, e232 = [np.asarray(x, np.float32) for x in f32]
        # textbook: o + t d = v0 + u e1 + v e2
        A = np.stack([-d32.astype(np.float64), e132.astype(np.float64), e232.astype(np.float64)], axis=1)
        t, u, v = np.linalg.solve(A, o32.astype(np.float64) - v032.astype(np.float64))
        det = np.dot(np.cross(e132.astype(np.float64), e232.astype(np.float64)), -d32.astype(np.float64))
        expect = det >= 2e-4 and u >= 1e-5 and v >= 1e-5 and u + v <= 1 - 1e-5 and t > 2e-4
        margin_ok = abs(det - 1e-4) > 1e-5 and min(abs(u), abs(v), abs(1 - u - v)) > 1e-5
        hit, tt, uv = oracle.ray_tri(o32, d32, v032, e132, e232)
        if not margin_ok:
            continue
        assert hit == expect
        if hit and det > 0.05:
            # FP32 rounding is amplified by 1/det; well-conditioned cases only
            assert abs(tt - t) <= 2e-5 * max(1.0, abs(t))
            assert np.allclose(uv, [u, v], atol=1e-4)
            agree += 1
    assert agree > 500


def test_box_closed_forms():
    lo, hi = [0, 0, 0], [1, 1, 1]
    hit, tn, tf = oracle.ray_box(lo, hi, [-1, 0.5, 0.5], [1, 0, 0])
    assert hit and tn == 1.0 and tf == 2.0
    hit, tn, tf = oracle.ray_box(lo, hi, [0.5, 0.5, 0.5], [0, 0, 1])
    assert hit and tn == 0.0 and tf == 0.5                  # inside: t_near clamped to 0
    assert not oracle.ray_box(lo, hi, [2, 0.5, 0.5], [1, 0, 0])[0]   # box behind
    assert not oracle.ray_box(lo, hi, [-1, 2, 0.5], [1, 0, 0])[0]    # passes above
    d = np.array([1, 1, 1], np.float32) / np.sqrt(np.float32(3))
    hit, tn, tf = oracle.ray_box(lo, hi, [-1, -1, -1], d)
    assert hit and abs(tn - np.sqrt(3)) < 1e-6 and abs(tf - 2 * np.sqrt(3)) < 1e-6


def test_box_ray_in_slab_plane_uses_minnum():
    lo, hi = [0, 0, 0], [1, 1, 1]
    # y = 0 with d.y = 0: (0-0)*inf = NaN, (1-0)*inf = inf -> minNum/maxNum give
    # [inf, inf] for the y slab, so tmin = inf > tmax: miss.
    assert not oracle.ray_box(lo, hi, [-1, 0, 0.5], [1, 0, 0])[0]
    # y = 1: (0-1)*inf = -inf, NaN -> [-inf, -inf]: tmax < 0: miss.
    assert not oracle.ray_box(lo, hi, [-1, 1, 0.5], [1, 0, 0])[0]
    # strictly inside the slab: hit
    assert oracle.ray_box(lo, hi, [-1, 0.5, 0.5], [1, 0, 0])[0]


def test_rope_face_order():
    lo, hi = [0, 0, 0], [1, 1, 1]
    assert oracle.neighbour_face(lo, hi, [0, 0.5, 1]) == 0       # LEFT beats FRONT
    assert oracle.neighbour_face(lo, hi, [0.5, 0.5, 1]) == 1     # FRONT
    assert oracle.neighbour_face(lo, hi, [1, 1, 0.5]) == 2       # RIGHT beats TOP
    assert oracle.neighbour_face(lo, hi, [0.5, 0.5, 0]) == 3     # BACK
    assert oracle.neighbour_face(lo, hi, [0.5, 1, 0.5]) == 4     # TOP
    assert oracle.neighbour_face(lo, hi, [0.5, 0, 0.5]) == 5     # BOTTOM
    assert oracle.neighbour_face(lo, hi, [0.5, 0.5, 0.5]) == -1
    assert oracle.neighbour_face(lo, hi, [0.5, 0.5, 1 - 0.9e-4]) == 1
    assert oracle.neighbour_face(lo, hi, [0.5, 0.5, 1 - 1.2e-4]) == -1


def test_mt19937_known_answers():
    assert oracle.mt19937_nth(5489, 10000) == 4123659995   # fixed by the C++ standard
    assert oracle.mt19937_nth(5489, 1) == 3499211612


def test_perlin_table_is_a_doubled_permutation():
    perm, grads = oracle.perlin_table(2016)
    assert sorted(perm[:256].tolist()) == list(range(256))
    assert np.array_equal(perm[:256], perm[256:])
    assert np.allclose(np.linalg.norm(grads, axis=1), 1.0, atol=1e-6)


def test_perlin_swaps_use_a_fresh_engine_copy():
    """std::bind copies the engine: the swap loop's draws are outputs 1..256 of a
    freshly seeded mt19937, not outputs 513..768."""
    perm, _ = oracle.perlin_table(2016)
    P = list(range(256))
    for i in range(256):
        j = oracle.mt19937_nth(2016, i + 1) & 255
        P[i], P[j] = P[j], P[i]
    assert perm[:256].tolist() == P
