"""Pins of the oracle's exact-field march, hybrid tracer and FP32 network.

orc_exact_march restates the library's exact-field tracer and
orc_hybrid_trace its per-ray surrogate/exact state machine (neither has a
reference counterpart: the reference's tracer is the sphere trace pinned in
test_terrain_oracle.py).  The checks fix them by what the mathematics gives:
a plane field's crossing in closed form, the FP64 first crossing of the
noise field, and a constant network whose value puts the state machine in
one phase (below tau the hybrid is the exact march, bit for bit; above it
the surrogate steps of s - margin run out of the box).  The FP32 network is
checked against the FP64 definition of the network (test_mlp_oracle.py pins
that one by finite differences).
"""
import numpy as np
import pytest

import oracle
import rayengine_amd as re

SP = 1.0 / 64


def _plane_grid(c, n=16):
    y = (np.arange(n, dtype=np.float32) * np.float32(1.0 / n))
    g = np.broadcast_to((y - np.float32(c))[None, :, None], (n, n, n))
    return np.ascontiguousarray(g, np.float32)


def _rays_down(rng, n):
    o = np.stack([rng.uniform(0.2, 0.8, n), np.full(n, 0.95), rng.uniform(0.2, 0.8, n)], 1).astype(np.float32)
    d = np.stack([rng.uniform(-0.3, 0.3, n), -np.ones(n), rng.uniform(-0.3, 0.3, n)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def test_exact_march_plane_closed_form():
    c, n_bisect, h_min = 0.37, 8, 1.0 / 64
    g = _plane_grid(c)
    rng = np.random.default_rng(1)
    o, d = _rays_down(rng, 500)
    t, hit, samples = oracle.exact_march(g, (1 / 16,) * 3, o, d, k_exact=0.5, h_min=h_min, n_bisect=n_bisect,
                                         max_iters=4096)
    assert hit.all()
    want = (o[:, 1].astype(np.float64) - c) / -d[:, 1].astype(np.float64)
    # the bracket of the last step (h_min once proportional steps stall) halved n_bisect times, midpoint taken
    assert np.abs(t - want).max() <= h_min / 2 ** (n_bisect + 1) + 2e-5
    # one sample at the entry, one per step, one per bisection
    assert (samples >= 2 + n_bisect).all()
    # without bisection the answer is only within the last step
    t0, _, _ = oracle.exact_march(g, (1 / 16,) * 3, o, d, k_exact=0.5, h_min=h_min, n_bisect=0, max_iters=4096)
    assert np.abs(t0 - want).max() > 4 * (h_min / 2 ** (n_bisect + 1) + 2e-5)


def test_exact_march_starting_inside_and_missing():
    g = _plane_grid(0.5)
    o = np.float32([[0.5, 0.2, 0.5], [0.5, 0.9, 0.5], [2.0, 2.0, 2.0]])
    d = np.float32([[0, 1, 0], [0, 1, 0], [1, 0, 0]])
    t, hit, samples = oracle.exact_march(g, (1 / 16,) * 3, o, d, 0.5, 1 / 64, 8, 4096)
    assert hit[0] == 1 and t[0] == 0.0 and samples[0] == 1      # inside at the entry
    assert hit[1] == 0 and t[1] == np.float32(1e20)           # leaves through the top
    assert hit[2] == 0 and samples[2] == 0                    # misses the box


@pytest.fixture(scope="module")
def noise64():
    perm = re.perlin_table(2016)
    g32 = oracle.noise_fill(perm, dim=(64, 64, 64), spacing=(SP,) * 3)
    g64 = oracle.noise_fill_f64(perm, dim=(64, 64, 64), spacing=(SP,) * 3)
    cam = re.Camera(pos=(0.5, 0.95, -0.35), view=(0.0, -0.55, 1.0))
    o, d = re.camera_rays(cam, 96, 64)
    return g32, g64, o, d


def test_exact_march_against_fp64_first_crossing(noise64):
    g32, g64, o, d = noise64
    t, hit, _ = oracle.exact_march(g32, (SP,) * 3, o, d, k_exact=0.05, h_min=0.5 * SP, n_bisect=8, max_iters=4096)
    tt, th = oracle.truth_trace(g64, (SP,) * 3, o, d, h_voxels=0.05)
    both = (hit == 1) & (th == 1)
    assert np.mean(hit == th) >= 0.995 and both.sum() > 3000
    err = np.abs(t[both].astype(np.float64) - tt[both]) / SP
    assert np.median(err) < 0.01 and np.percentile(err, 99) < 0.5


def test_hybrid_below_tau_is_the_exact_march(noise64):
    """A network whose value never exceeds tau hands every ray to the exact
    phase at the box entry and again after each rejoin at the same point:
    the steps, hence t and hit, are the exact march's bit for bit."""
    g32, _, o, d = noise64
    P = oracle.hybrid_world_params(SP, 4.0, 1.0, 0.05, 0.5, 4.0, 8, 4096)
    t, hit, ev, sm = oracle.hybrid_trace(g32, (SP,) * 3, None, o, d, const_s=P["tau"] * 0.5, **P)
    te, he, se = oracle.exact_march(g32, (SP,) * 3, o, d, P["k_exact"], P["h_min"], P["n_bisect"], P["max_iters"])
    assert np.array_equal(hit, he) and np.array_equal(t.view(np.uint32), te.view(np.uint32))
    inbox = se > 0
    assert (ev[inbox] >= 1).all() and (ev[~inbox] == 0).all()
    # every rejoin re-samples the point where the exact phase resumes
    assert (sm[inbox] == se[inbox] + ev[inbox] - 1).all()


def test_hybrid_above_tau_steps_s_minus_margin(noise64):
    g32, _, o, d = noise64
    P = oracle.hybrid_world_params(SP, 4.0, 3.0, 0.05, 0.5, 4.0, 8, 100000)
    s = np.float32(0.07)   # above tau (4 voxels = 0.0625); steps of s - margin = 0.07 - 3 voxels
    t, hit, ev, sm = oracle.hybrid_trace(g32, (SP,) * 3, None, o, d, const_s=float(s), **P)
    assert not hit.any() and (sm == 0).all()
    # the box entry and exit of each ray, and the steps of (s - margin) between them
    lo, hi = np.zeros(3, np.float32), np.float32([64 * SP] * 3)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = np.float32(1) / d
        t1, t2 = (lo - o) * inv, (hi - o) * inv
    tn = np.maximum(np.nanmax(np.minimum(t1, t2), axis=1), 0)
    tf = np.nanmin(np.maximum(t1, t2), axis=1)
    inbox = tf >= tn
    step = np.float64(s - np.float32(P["margin"]))
    want = np.ceil((tf[inbox] - tn[inbox]) / step)
    assert np.abs(ev[inbox] - want).max() <= 1
    # a step of s (margin dropped) would need far fewer evaluations
    assert np.mean(np.abs(ev[inbox] - np.ceil((tf[inbox] - tn[inbox]) / s))) > 5
    # the advance cap
    P3 = dict(P, max_iters=3)
    _, hit3, ev3, _ = oracle.hybrid_trace(g32, (SP,) * 3, None, o, d, const_s=float(s), **P3)
    assert not hit3.any() and ev3.max() == 3


def test_fp32_network_matches_fp64_definition():
    rng = np.random.default_rng(4)
    n_params = 18945
    p = np.zeros(n_params, np.float32)
    sizes = [(0, 6336, 99), (6400, 10496, 64), (10560, 14656, 64), (14720, 18816, 64), (18880, 18944, 64)]
    for a, b, fan_in in sizes:
        p[a:b] = rng.uniform(-1, 1, b - a) * np.sqrt(3.0 / fan_in)
    p[6336:6400] = rng.uniform(-0.1, 0.1, 64)
    p[18944] = 0.3
    pts = rng.uniform(0, 1, (200, 3)).astype(np.float32)
    f32 = oracle.mlp_eval_f32(p, pts)
    f64 = oracle.mlp_forward(p, pts)
    assert np.abs(f32 - f64).max() < 2e-4 * max(1.0, np.abs(f64).max())
