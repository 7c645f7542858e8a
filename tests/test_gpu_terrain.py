"""GPU parity of the noise-fill and sphere-trace kernels against the CPU
oracle: the volume bit-exact, the march bit-exact (t, hit, steps, normals);
the FP32 hit-depth error against the FP64 oracle is measured and bounded."""
import sys

import numpy as np
import pytest

import oracle
import rayengine_amd as re
from rayengine_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def perm():
    return re.perlin_table(2016)


@pytest.fixture(scope="module")
def terrain(gpu_ctx, perm):
    say("noise volume 256^3")
    vol = gpu_ctx.noise_volume()
    grid = vol.read()
    return vol, grid


def test_noise_volume_bit_exact(terrain, perm):
    vol, grid = terrain
    ref = oracle.noise_fill(perm)
    assert grid.shape == (256, 256, 256)
    assert np.array_equal(grid.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("dims,seed", [((20, 12, 36), 7), ((64, 64, 64), 2017), ((9, 130, 17), 1)])
def test_noise_volume_odd_dims_and_seeds(gpu_ctx, dims, seed):
    say(f"noise dims {dims} seed {seed}")
    sp = (0.05, 0.02, 0.03)
    vol = gpu_ctx.noise_volume(dim=dims, spacing=sp, seed=seed, octaves=4, base_freq=3.0)
    g = vol.read()
    ref = oracle.noise_fill(re.perlin_table(seed), dim=dims, spacing=sp, octaves=4, base_freq=3.0)
    assert np.array_equal(g.view(np.uint32), ref.view(np.uint32))
    vol.release()


def test_noise_volume_with_origin(gpu_ctx, perm):
    """The chunk-location origin (cuda_update_chunk_gen, renderer.cu:982-1000)."""
    for origin in ((0.25, -0.5, 3.0), (-17.125, 0.0, 40.5)):
        say(f"noise origin {origin}")
        vol = gpu_ctx.noise_volume(dim=(40, 24, 33), spacing=(0.05, 0.02, 0.03), origin=origin)
        ref = oracle.noise_fill(perm, dim=(40, 24, 33), spacing=(0.05, 0.02, 0.03), origin=origin)
        assert np.array_equal(vol.read().view(np.uint32), ref.view(np.uint32))
        vol.release()


def test_volume_upload_pairs_keeps_material(gpu_ctx):
    rng = np.random.default_rng(1)
    g = rng.standard_normal((9, 14, 11)).astype(np.float32)
    m = rng.integers(0, 5, g.shape).astype(np.float32)
    vol = gpu_ctx.upload_volume_pairs(g, m, (0.1, 0.2, 0.3))
    assert np.array_equal(vol.read(), g) and np.array_equal(vol.read_material(), m)
    assert list(vol.dim) == [11, 14, 9]
    vol.release()


def test_volume_upload_roundtrip(gpu_ctx):
    rng = np.random.default_rng(0)
    g = rng.standard_normal((13, 27, 10)).astype(np.float32)
    vol = gpu_ctx.upload_volume(g, (0.1, 0.2, 0.3))
    assert np.array_equal(vol.read(), g)
    assert list(vol.dim) == [10, 27, 13]
    vol.release()


def test_sphere_trace_bit_exact_full_frame(terrain):
    vol, grid = terrain
    o, d = W.terrain_rays(1920, 1080)
    say("sphere trace 1920x1080")
    t, hit, steps, nrm = vol.trace(o, d, normals=True)
    to, ho, so, no = oracle.sdf_trace(grid, W.TERRAIN_SPACING, o, d, normals=True)
    assert np.array_equal(hit, ho) and np.array_equal(steps, so)
    assert np.array_equal(t.view(np.uint32), to.view(np.uint32))
    assert np.array_equal(nrm.view(np.uint32), no.view(np.uint32))
    assert hit.mean() > 0.5
    # the terrain depth bound: within 1e-4 of the FP32 reference march
    assert np.abs(t[hit == 1] - to[hit == 1]).max() <= 1e-4


def test_sphere_trace_reference_step_rule_and_caps(terrain):
    vol, grid = terrain
    o, d = W.terrain_rays(320, 180)
    for kw in (dict(step_rule=1), dict(max_steps=1), dict(max_steps=3, eps=1e-2, step_scale=0.5)):
        say(f"sphere trace {kw}")
        t, hit, steps = vol.trace(o, d, **kw)
        to, ho, so = oracle.sdf_trace(grid, W.TERRAIN_SPACING, o, d, **kw)
        assert np.array_equal(hit, ho) and np.array_equal(steps, so)
        assert np.array_equal(t.view(np.uint32), to.view(np.uint32))


def test_sphere_trace_plane_closed_form(gpu_ctx, perm):
    vol = gpu_ctx.noise_volume(dim=(32, 32, 32), spacing=(1 / 32,) * 3, octaves=0)
    rng = np.random.default_rng(6)
    n = 4096
    o = np.stack([rng.uniform(0.2, 0.8, n), np.full(n, 0.95), rng.uniform(0.2, 0.8, n)], 1).astype(np.float32)
    d = np.stack([rng.uniform(-0.3, 0.3, n), rng.uniform(-1, -0.4, n), rng.uniform(-0.3, 0.3, n)], 1)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    t, hit, steps = vol.trace(o, d)
    tstar = (0.4 - o[:, 1].astype(np.float64)) / d[:, 1]
    ok = (np.abs(o[:, 0] + tstar * d[:, 0] - 0.5) < 0.49) & (np.abs(o[:, 2] + tstar * d[:, 2] - 0.5) < 0.49)
    assert hit[ok].all()
    gap = tstar[ok] - t[ok]
    assert (gap >= -1e-6).all() and (gap <= 1e-3 * t[ok] * (1 - 0.8 * np.abs(d[ok, 1])) / np.abs(d[ok, 1]) + 1e-6).all()
    vol.release()


def test_fp32_hit_depth_error_vs_fp64(terrain, perm):
    """Reports the FP32 march's hit-depth error against the FP64 oracle on a
    480x270 frame (the FP64 chain: field, sampling and march in double)."""
    vol, grid = terrain
    o, d = W.terrain_rays(480, 270)
    t, hit, steps = vol.trace(o, d)
    g64 = oracle.noise_fill_f64(perm)
    t64, h64, s64 = oracle.sdf_trace_f64(g64, W.TERRAIN_SPACING, o, d)
    both = (hit == 1) & (h64 == 1)
    err = np.abs(t[both].astype(np.float64) - t64[both])
    say(f"FP32 vs FP64 hit depth: agree {np.mean(hit == h64):.6f}, median {np.median(err):.3e}, "
        f"p99 {np.percentile(err, 99):.3e}, max {err.max():.3e} (voxel = {1 / 256:.3e})")
    assert np.mean(hit == h64) > 0.999
    assert np.percentile(err, 99) < 1e-3


# The depth reference of the terrain benchmark: the exact-field march with
# small steps and many bisections, pinned here against the FP64 first
# crossing of the oracle.
REFERENCE_DEPTH = dict(k_exact=0.02, h_min_voxels=0.25, n_bisect=20, max_iters=1 << 16)


def test_fine_exact_march_is_a_depth_reference(terrain, perm):
    vol, grid = terrain
    o, d = W.terrain_rays(480, 270)
    t, h, _, samples = re.hybrid_trace(vol, None, o, d, **REFERENCE_DEPTH)
    g64 = oracle.noise_fill_f64(perm)
    tt, th = oracle.truth_trace(g64, W.TERRAIN_SPACING, o, d)
    both = (h == 1) & (th == 1)
    e = np.abs(t[both].astype(np.float64) - tt[both]) * 256
    say(f"fine exact march vs FP64 crossing: agree {np.mean(h == th):.5f}, p50 {np.median(e):.4f}, "
        f"p99 {np.percentile(e, 99):.4f} voxel, samples/ray {samples.mean():.1f}")
    assert np.mean(h == th) >= 0.999
    assert np.percentile(e, 99) <= 0.02
