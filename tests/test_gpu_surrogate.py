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


def test_checkpoint_resume_is_bit_exact(gpu_ctx, tmp_path):
    """Training N steps, saving, loading into a new network and training N
    more reproduces 2N uninterrupted steps bit for bit: parameters, Adam
    moments, step and the losses of the second half.  Two calls of N steps on
    one network do the same (the sample pool's batch cursor continues)."""
    vol = gpu_ctx.noise_volume(dim=(48, 48, 48), spacing=(1 / 48,) * 3)
    dv = re.distance_volume(vol)
    kw = dict(n_samples=1 << 16, batch=4096, lr=2e-3, seed=9)
    say("uninterrupted 2N steps")
    a = re.Surrogate(gpu_ctx, seed=21)
    hist_a = a.train(dv, steps=24, **kw)   # 24 batches of 4096 wrap the 65536-sample pool
    pa = a.params()
    say("N steps, checkpoint, load, N steps")
    b = re.Surrogate(gpu_ctx, seed=21)
    b.train(dv, steps=12, **kw)
    ck = tmp_path / "net.ckpt"
    b.save(ck)
    b.release()
    c = re.Surrogate.load(gpu_ctx, ck)
    hist_c = c.train(dv, steps=12, **kw)
    pc = c.params()
    for x, y, name in zip(pa, pc, ("params", "adam_m", "adam_v", "step")):
        assert np.array_equal(np.asarray(x), np.asarray(y)), name
    assert np.array_equal(hist_a[12:], hist_c)
    say("two calls on one network")
    d = re.Surrogate(gpu_ctx, seed=21)
    d.train(dv, steps=12, **kw)
    d.train(dv, steps=12, **kw)
    pd = d.params()
    assert all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(pa, pd))
    # a new seed starts a new pool from its first batch
    e = re.Surrogate(gpu_ctx, seed=21)
    e.train(dv, steps=12, **kw)
    e.train(dv, steps=12, **dict(kw, seed=10))
    assert not np.array_equal(e.params()[0], pa[0])
    # files that are not checkpoints are refused
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(ck.read_bytes()[:-4])
    with pytest.raises(re.ReError):
        re.Surrogate.load(gpu_ctx, bad)
    bad.write_bytes(b"RESURR\x00\x02" + ck.read_bytes()[8:])
    with pytest.raises(re.ReError):
        re.Surrogate.load(gpu_ctx, bad)
    with pytest.raises(re.ReError):
        re.Surrogate.load(gpu_ctx, tmp_path / "missing.ckpt")
    for net in (a, c, d, e):
        net.release()
    dv.release()
    vol.release()


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


def _world(vol_spacing, **kw):
    p = re.hybrid_params(**kw)
    return oracle.hybrid_world_params(vol_spacing, p.tau_voxels, p.margin_voxels, p.k_exact, p.h_min_voxels,
                                      p.rejoin_voxels, p.n_bisect, p.max_iters)


def test_exact_march_bit_exact_with_oracle(gpu_ctx):
    """The exact-field march (re_hybrid_trace without a network) equals the
    oracle's restatement bit for bit: t, hit and samples per ray."""
    vol = gpu_ctx.noise_volume(dim=(128, 128, 128), spacing=(1 / 128,) * 3)
    g = vol.read()
    o, d = W.terrain_rays(480, 270)
    for kw in (dict(), dict(k_exact=0.1, n_bisect=4, h_min_voxels=1.0), dict(max_iters=7)):
        say(f"exact march {kw}")
        t, h, ev, sm = re.hybrid_trace(vol, None, o, d, **kw)
        P = _world(1 / 128, **kw)
        to, ho, so = oracle.exact_march(g, (1 / 128,) * 3, o, d, P["k_exact"], P["h_min"], P["n_bisect"],
                                        P["max_iters"])
        assert np.array_equal(h, ho) and np.array_equal(t.view(np.uint32), to.view(np.uint32))
        assert np.array_equal(sm, so) and not ev.any()
    vol.release()


def test_hybrid_state_machine_with_constant_networks(gpu_ctx):
    """With all weights zero the network's value is b4.  Below tau the hybrid
    is the exact march bit for bit; above tau every ray steps s - margin to
    the box exit; both equal the oracle's hybrid with that constant, network
    evaluations and field samples included."""
    vol = gpu_ctx.noise_volume(dim=(96, 96, 96), spacing=(1 / 96,) * 3)
    g = vol.read()
    o, d = W.terrain_rays(320, 180)
    net = re.Surrogate(gpu_ctx, seed=0)
    kw = dict(tau_voxels=4.0, margin_voxels=3.0, k_exact=0.05, n_bisect=6)
    P = _world(1 / 96, **kw)
    te, he, _, _ = re.hybrid_trace(vol, None, o, d, **kw)
    for s in (0.5 * P["tau"], 2.0 * P["tau"]):
        say(f"constant network {s:.4f} (tau {P['tau']:.4f})")
        p = np.zeros(NP, np.float32)
        p[-1] = s
        net.set_params(p)
        t, h, ev, sm = re.hybrid_trace(vol, net, o, d, **kw)
        to, ho, eo, so = oracle.hybrid_trace(g, (1 / 96,) * 3, None, o, d, const_s=float(np.float32(s)), **P)
        assert np.array_equal(h, ho) and np.array_equal(t.view(np.uint32), to.view(np.uint32))
        assert np.array_equal(ev, eo) and np.array_equal(sm, so)
        if s < P["tau"]:
            assert np.array_equal(h, he) and np.array_equal(t.view(np.uint32), te.view(np.uint32))
        else:
            assert not h.any() and ev.max() > 10
    net.release()
    vol.release()


def test_trained_hybrid_against_fp32_oracle_network(gpu_ctx):
    """The trained hybrid on the device (BF16 network) against the oracle's
    hybrid with the same parameters in FP32: the network values differ by
    the BF16 rounding, so the surrogate steps differ, but both stop at the
    exact field's crossing.  Bar: hit agreement >= 0.999 and the 99th
    percentile of |t_gpu - t_oracle| <= 0.25 voxel."""
    vol = gpu_ctx.noise_volume(dim=(128, 128, 128), spacing=(1 / 128,) * 3)
    g = vol.read()
    dv = re.distance_volume(vol)
    net = re.Surrogate(gpu_ctx, seed=1)
    net.train(dv, n_samples=1 << 20, batch=65536, steps=1500, lr=1e-3, seed=1, over_weight=4.0)
    o, d = W.terrain_rays(240, 135)
    P = _world(1 / 128, **HYBRID)
    t, h, ev, sm = re.hybrid_trace(vol, net, o, d, **HYBRID)
    to, ho, eo, so = oracle.hybrid_trace(g, (1 / 128,) * 3, net.params()[0], o, d, **P)
    both = (h == 1) & (ho == 1)
    err = np.abs(t[both].astype(np.float64) - to[both]) * 128
    say(f"trained hybrid vs FP32 oracle: agree {np.mean(h == ho):.5f}, p99 {np.percentile(err, 99):.4f} voxel, "
        f"evals/ray {ev.mean():.2f} vs {eo.mean():.2f}")
    assert np.mean(h == ho) >= 0.999 and both.sum() > 10000
    assert np.percentile(err, 99) <= 0.25
    assert abs(ev.mean() - eo.mean()) <= 0.1 * eo.mean()
    net.release()
    dv.release()
    vol.release()


def test_hybrid_is_network_bound(gpu_ctx):
    """The hybrid's throughput relation on the 256^3 terrain, 1920x1080 rays:
    its Mrays/s is bounded by the network's own evaluation rate divided by
    the network evaluations per ray (the per-ray cost floor), it reaches a
    fair share of that bound (the exact phase and the wave scheduling cost
    less than the network), and at 8+ evaluations per ray that bound lies
    below the exact-field march it is meant to beat."""
    vol = gpu_ctx.noise_volume()
    dv = re.distance_volume(vol)
    net = re.Surrogate(gpu_ctx, seed=1)
    net.train(dv, n_samples=1 << 22, batch=65536, steps=4000, lr=1e-3, seed=1, over_weight=4.0)
    o, d = W.terrain_rays(1920, 1080)
    n = len(o)

    def best_ms(fn, reps=4):
        fn()
        ts = []
        for _ in range(reps):
            fn()
            ts.append(gpu_ctx.last_kernel_ms())
        return float(np.median(ts))

    hy_ms = best_ms(lambda: re.hybrid_trace(vol, net, o, d, **HYBRID))
    _, _, ev, _ = re.hybrid_trace(vol, net, o, d, **HYBRID)
    ex_ms = best_ms(lambda: re.hybrid_trace(vol, None, o, d, k_exact=0.05, h_min_voxels=1.0, n_bisect=1))
    pts = np.random.default_rng(0).uniform(0, 1, (1 << 22, 3)).astype(np.float32)
    fw_ms = best_ms(lambda: net.forward(pts))
    evals_per_ray = float(ev.mean())
    net_gevals = (1 << 22) / (fw_ms * 1e-3) / 1e9
    bound = net_gevals * 1e3 / evals_per_ray          # Mrays/s the network alone allows
    hy, ex = n / (hy_ms * 1e-3) / 1e6, n / (ex_ms * 1e-3) / 1e6
    say(f"hybrid {hy:.0f} Mrays/s, exact march {ex:.0f}, network {net_gevals:.2f} Gevals/s x "
        f"{evals_per_ray:.2f} evals/ray -> bound {bound:.0f} Mrays/s ({hy / bound:.2f} of it)")
    assert evals_per_ray > 4
    assert hy <= 1.1 * bound
    assert hy >= 0.3 * bound
    assert bound < ex
    net.release()
    dv.release()
    vol.release()
