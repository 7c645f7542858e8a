"""GPU surrogate kernels against the FP64 CPU oracle.

The GPU computes in BF16 (weights, features, activations) with FP32
accumulation, so parity is to a stated tolerance, not bit-exact:
  forward: |gpu - fp64| <= 0.02 * max|fp64| + 2e-3 (weights pre-rounded to BF16
           for the oracle, so the tolerance covers activation/feature rounding);
  gradient: cosine similarity >= 0.995 per layer and relative L2 error <= 0.1;
  Adam: the FP32 update equals the FP64 closed form to 1e-6 relative.
The distance transform is exact for its seeds and is compared bit-for-bit in
structure (equal squared distances) with a brute-force numpy transform."""
import sys

import numpy as np
import pytest

import oracle
import rayengine_amd as re
from rayengine_amd import workloads as W

pytestmark = pytest.mark.gpu

NP = 18945
LAYERS = [("W0", 0, 6336), ("b0", 6336, 6400), ("W1", 6400, 10496), ("b1", 10496, 10560), ("W2", 10560, 14656),
          ("b2", 14656, 14720), ("W3", 14720, 18816), ("b3", 18816, 18880), ("W4", 18880, 18944), ("b4", 18944, 18945)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bf16_round(x):
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def weights_bf16(p):
    q = p.copy()
    for name, a, b in LAYERS:
        if name.startswith("W") and name != "W4":
            q[a:b] = bf16_round(p[a:b])
    return q


@pytest.fixture(scope="module")
def net(gpu_ctx):
    s = re.Surrogate(gpu_ctx, seed=3)
    yield s
    s.release()


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(NP, np.float32)
    for name, a, b in LAYERS:
        fan_in = 99 if name == "W0" else 64
        if name.startswith("W"):
            p[a:b] = rng.uniform(-1, 1, b - a) * np.sqrt(3.0 / fan_in)
        else:
            p[a:b] = rng.uniform(-0.1, 0.1, b - a)
    return p


def test_forward_matches_fp64_oracle(net):
    p = _params(1)
    net.set_params(p)
    pts = np.random.default_rng(2).uniform(0, 1, (4096, 3)).astype(np.float32)
    out = net.forward(pts)
    ref = oracle.mlp_forward(weights_bf16(p), pts)
    err = np.abs(out - ref)
    say(f"forward: max |err| {err.max():.3e}, max |ref| {np.abs(ref).max():.3e}")
    assert err.max() <= 0.02 * np.abs(ref).max() + 2e-3


def test_gradient_matches_fp64_oracle(net):
    p = _params(4)
    net.set_params(p)
    rng = np.random.default_rng(5)
    n = 1024
    s = np.concatenate([rng.uniform(0, 1, (n, 3)), rng.uniform(-0.05, 0.05, (n, 1))], 1).astype(np.float32)
    loss, g = net.batch_step(s, over_weight=4.0)
    lref, gref = oracle.mlp_loss_grad(weights_bf16(p), s[:, :3], s[:, 3], over_weight=4.0)
    say(f"loss gpu {loss:.6e} fp64 {lref:.6e}")
    assert abs(loss - lref) <= 0.05 * abs(lref)
    for name, a, b in LAYERS:
        x, y = g[a:b].astype(np.float64), gref[a:b]
        if np.linalg.norm(y) == 0:
            continue
        cos = x @ y / (np.linalg.norm(x) * np.linalg.norm(y) + 1e-30)
        rel = np.linalg.norm(x - y) / np.linalg.norm(y)
        say(f"{name}: cos {cos:.5f} rel {rel:.3e}")
        assert cos >= 0.995 and rel <= 0.1, name


def test_adam_step_closed_form(net):
    p = _params(6)
    net.set_params(p)
    rng = np.random.default_rng(7)
    s = np.concatenate([rng.uniform(0, 1, (512, 3)), rng.uniform(-0.05, 0.05, (512, 1))], 1).astype(np.float32)
    _, g = net.batch_step(s, adam=True, lr=1e-3)
    q, m, v, step = net.params()
    assert step == 1
    pe, me, ve = p.copy(), np.zeros(NP, np.float32), np.zeros(NP, np.float32)
    oracle.adam(pe, me, ve, g, 1, lr=1e-3)
    assert np.allclose(m, me, rtol=1e-5, atol=1e-12) and np.allclose(v, ve, rtol=1e-5, atol=1e-15)
    assert np.allclose(q, pe, rtol=1e-6, atol=1e-7)


def _brute_edt(field):
    """Squared distance (voxels) from every voxel to the seeds of the GPU rule."""
    f = field
    dz, dy, dx = f.shape
    init = np.full(f.shape, np.inf)
    pos = f > 0
    for ax in range(3):
        for sh in (1, -1):
            g = np.roll(f, sh, axis=ax)
            pg = np.roll(pos, sh, axis=ax)
            valid = np.ones(f.shape, bool)
            idx = [slice(None)] * 3
            idx[ax] = 0 if sh == 1 else -1
            valid[tuple(idx)] = False
            ch = valid & (pos != pg)
            fr = np.abs(f) / np.abs(f - g)
            init = np.where(ch, np.minimum(init, (fr * fr).astype(np.float32)), init)
    seeds = np.argwhere(np.isfinite(init))
    zz, yy, xx = np.meshgrid(np.arange(dz), np.arange(dy), np.arange(dx), indexing="ij")
    P = np.stack([zz, yy, xx], -1).reshape(-1, 3).astype(np.float64)
    best = np.full(P.shape[0], np.inf)
    for s in seeds:
        d2 = ((P - s) ** 2).sum(1) + init[tuple(s)]
        best = np.minimum(best, d2)
    return best.reshape(f.shape)


def test_distance_transform_matches_brute_force(gpu_ctx):
    rng = np.random.default_rng(8)
    g = (rng.standard_normal((14, 17, 20)) + 0.3).astype(np.float32)
    vol = gpu_ctx.upload_volume(g, (0.1, 0.1, 0.1))
    dv = re.distance_volume(vol)
    d = dv.read()
    ref = np.sqrt(_brute_edt(g)) * np.float32(0.1)
    ref = np.where(g > 0, ref, -ref)
    assert np.allclose(d, ref, rtol=1e-5, atol=1e-6)
    dv.release()
    vol.release()


def test_sampler_targets_and_mix(gpu_ctx):
    vol = gpu_ctx.noise_volume(dim=(64, 64, 64), spacing=(1 / 64,) * 3)
    dv = re.distance_volume(vol)
    s = re.sample_points(dv, 8192, seed=3, near_band=2 / 64)
    grid = dv.read()
    near = np.abs(s[1::2, 3]) < 2 / 64
    assert near.mean() > 0.95
    assert (s[:, :3] >= 0).all() and (s[:, :3] <= 63 / 64 + 1e-6).all()
    for i in range(0, 8192, 97):
        t = oracle.trilinear(grid, (1 / 64,) * 3, *s[i, :3])
        assert abs(t - s[i, 3]) <= 1e-6 + 1e-5 * abs(t)


def test_training_reduces_loss(gpu_ctx):
    vol = gpu_ctx.noise_volume(dim=(64, 64, 64), spacing=(1 / 64,) * 3)
    dv = re.distance_volume(vol)
    net = re.Surrogate(gpu_ctx, seed=11)
    hist = net.train(dv, n_samples=1 << 18, batch=8192, steps=400, lr=2e-3, seed=5)
    say(f"loss {hist[0]:.4e} -> {hist[-1]:.4e}")
    assert np.isfinite(hist).all()
    assert hist[-20:].mean() < 0.25 * hist[:5].mean()
    net.release()


# Hybrid settings of the terrain benchmark.
HYBRID = dict(tau_voxels=4.0, margin_voxels=1.0, k_exact=0.1, n_bisect=4)


def test_hybrid_accuracy_and_what_the_network_saves(gpu_ctx):
    """Against the FP64 first crossing on a 480x270 frame: the trained
    surrogate's hybrid trace stays within 0.5 voxel at the 99th percentile,
    and it takes fewer exact-field samples per ray than the exact march with
    the same near-surface settings (the network replaces the far-field
    march)."""
    vol = gpu_ctx.noise_volume()
    dv = re.distance_volume(vol)
    net = re.Surrogate(gpu_ctx, seed=1)
    hist = net.train(dv, n_samples=1 << 22, batch=65536, steps=4000, lr=1e-3, seed=1, over_weight=4.0)
    o, d = W.terrain_rays(480, 270)
    g64 = oracle.noise_fill_f64(re.perlin_table(2016))
    tt, th = oracle.truth_trace(g64, W.TERRAIN_SPACING, o, d)
    th_, hh, ev, sm = re.hybrid_trace(vol, net, o, d, **HYBRID)
    te, he, _, se = re.hybrid_trace(vol, None, o, d, k_exact=HYBRID["k_exact"], n_bisect=HYBRID["n_bisect"])
    both = (hh == 1) & (th == 1)
    e = np.abs(th_[both].astype(np.float64) - tt[both]) * 256
    say(f"loss {hist[-100:].mean():.2e}; hybrid agree {np.mean(hh == th):.5f} p99 {np.percentile(e, 99):.3f} voxel, "
        f"evals/ray {ev.mean():.2f}, samples/ray {sm.mean():.2f} vs exact march {se.mean():.2f}")
    assert np.mean(hh == th) >= 0.999
    assert np.percentile(e, 99) <= 0.5
    assert sm.mean() < 0.9 * se.mean()
    net.release()
    dv.release()
    vol.release()
