"""Pins the terrain part of the CPU oracle: gradient table, noise primitive,
fBm octave structure, trilinear sampling and the sphere-trace march, against
closed forms; and the library's Perlin table against the standard library's
own engine and distributions."""
import numpy as np
import pytest

import oracle
import rayengine_amd as re

# Improved-Perlin gradient set in the order of the reference switch
# (RayEngine/src/cuda-src/sdf_functions.cuh:30-46).
GRADS = [(1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1),
         (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1), (1, 1, 0), (-1, 1, 0), (0, -1, 1), (0, -1, -1)]


@pytest.fixture(scope="module")
def perm():
    return re.perlin_table(2016)


def test_library_perlin_table_equals_std_library(perm):
    for seed in (2016, 0, 1, 2017, 123456789):
        po, _ = oracle.perlin_table(seed)
        assert np.array_equal(re.perlin_table(seed), po)


def test_gradient_table(perm):
    rng = np.random.default_rng(1)
    for h in range(256):
        x, y, z = rng.uniform(-1, 1, 3).astype(np.float32)
        g = GRADS[h & 15]
        assert oracle.grad_dot(h, x, y, z) == np.float32(g[0] * x + g[1] * y + g[2] * z) or \
            abs(oracle.grad_dot(h, x, y, z) - (g[0] * x + g[1] * y + g[2] * z)) < 1e-6


def test_noise_zero_on_lattice_and_bounded(perm):
    rng = np.random.default_rng(2)
    for _ in range(200):
        p = rng.integers(-300, 300, 3).astype(np.float32)
        assert oracle.noise_eval(perm, *p) == 0.0
    vals = np.array([oracle.noise_eval(perm, *rng.uniform(-50, 50, 3).astype(np.float32)) for _ in range(5000)])
    assert np.abs(vals).max() < 1.1 and vals.std() > 0.1


def test_noise_continuous_across_cells(perm):
    rng = np.random.default_rng(3)
    for _ in range(200):
        p = rng.uniform(-20, 20, 3)
        ax = rng.integers(0, 3)
        p[ax] = np.round(p[ax])
        a, b = p.copy(), p.copy()
        a[ax] -= 1e-4
        b[ax] += 1e-4
        assert abs(oracle.noise_eval_f64(perm, *a) - oracle.noise_eval_f64(perm, *b)) < 2e-3


def _textbook_noise(P, p):
    """Improved Perlin noise in its lerp form (Perlin 2002): nested linear
    interpolation of the 8 corner gradient dots with quintic fades."""
    fl = np.floor(p).astype(np.int64)
    i = fl & 255
    t = p - fl
    fade = t * t * t * (t * (t * 6 - 15) + 10)

    def g(dx, dy, dz):
        h = P[P[P[(i[0] + dx) & 255] + ((i[1] + dy) & 255)] + ((i[2] + dz) & 255)]
        gx, gy, gz = GRADS[h & 15]
        return gx * (t[0] - dx) + gy * (t[1] - dy) + gz * (t[2] - dz)

    lerp = lambda a, b, s: a + s * (b - a)
    u, v, w = fade
    x00 = lerp(g(0, 0, 0), g(1, 0, 0), u)
    x10 = lerp(g(0, 1, 0), g(1, 1, 0), u)
    x01 = lerp(g(0, 0, 1), g(1, 0, 1), u)
    x11 = lerp(g(0, 1, 1), g(1, 1, 1), u)
    return lerp(lerp(x00, x10, v), lerp(x01, x11, v), w)


def test_noise_equals_textbook_lerp_form(perm):
    """The reference's k0..k7 polynomial expansion (sdf_functions.cuh:121-134)
    equals the nested-lerp definition; checked in FP64."""
    rng = np.random.default_rng(4)
    P = perm.astype(np.int64)
    for _ in range(2000):
        p = rng.uniform(-300, 300, 3)
        assert abs(oracle.noise_eval_f64(perm, *p) - _textbook_noise(P, p)) < 1e-12


def test_fbm_is_the_octave_sum(perm):
    grid = oracle.noise_fill(perm, dim=(8, 8, 8), spacing=(0.13, 0.11, 0.17))
    sp = np.float32([0.13, 0.11, 0.17])
    for (z, y, x) in [(0, 0, 0), (3, 5, 7), (7, 7, 7), (2, 6, 1)]:
        p = np.float32([x, y, z]) * sp
        s = np.float32(0)
        amp, freq = np.float32(1), np.float32(4)
        for _ in range(5):
            s = np.float32(s + amp * oracle.noise_eval(perm, *(p * freq)))
            freq = np.float32(freq * 2)
            amp = np.float32(amp * 0.5)
        assert grid[z, y, x] == np.float32((p[1] - np.float32(0.4)) - s)


def test_trilinear_reproduces_linear_fields_and_corners():
    rng = np.random.default_rng(5)
    dims = (9, 7, 11)
    sp = (0.25, 0.5, 0.125)
    zz, yy, xx = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    a, b, c, d = 0.3, -1.2, 0.7, 2.0
    grid = (a * xx * sp[0] + b * yy * sp[1] + c * zz * sp[2] + d).astype(np.float32)
    for _ in range(300):
        p = rng.uniform(0, 1, 3) * (np.array(dims) - 1) * sp
        v = oracle.trilinear(grid, sp, *p.astype(np.float32))
        assert abs(v - (a * p[0] + b * p[1] + c * p[2] + d)) < 2e-5
    for (x, y, z) in [(0, 0, 0), (3, 2, 5), (7, 5, 9)]:
        v = oracle.trilinear(grid, sp, x * sp[0], y * sp[1], z * sp[2])
        assert v == grid[z, y, x]


def test_sphere_trace_plane_closed_form():
    """Field y - 0.4 (no octaves): the march converges geometrically onto the
    plane; the returned post-step t satisfies 0 <= t* - t <= eps t (1 - s|d_y|) / |d_y|."""
    perm = re.perlin_table(2016)
    grid = oracle.noise_fill(perm, dim=(32, 32, 32), spacing=(1 / 32,) * 3, octaves=0)
    rng = np.random.default_rng(6)
    n = 2000
    o = np.stack([rng.uniform(0.2, 0.8, n), np.full(n, 0.95), rng.uniform(0.2, 0.8, n)], 1).astype(np.float32)
    d = np.stack([rng.uniform(-0.3, 0.3, n), rng.uniform(-1, -0.4, n), rng.uniform(-0.3, 0.3, n)], 1)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    t, hit, steps = oracle.sdf_trace(grid, (1 / 32,) * 3, o, d)
    tstar = (0.4 - o[:, 1].astype(np.float64)) / d[:, 1]
    inside = (o[:, 0] + tstar * d[:, 0] > 0) & (o[:, 0] + tstar * d[:, 0] < 1) & \
             (o[:, 2] + tstar * d[:, 2] > 0) & (o[:, 2] + tstar * d[:, 2] < 1)
    assert hit[inside].all()
    gap = tstar[inside] - t[inside]
    bound = 1e-3 * t[inside] * (1 - 0.8 * np.abs(d[inside, 1])) / np.abs(d[inside, 1])
    assert (gap >= -1e-6).all() and (gap <= bound + 1e-6).all()
    assert steps[inside].max() < 64


def test_sphere_trace_step_cap_and_miss():
    grid = np.full((8, 8, 8), 0.5, np.float32)            # never reaches the surface
    o = np.float32([[0.5, 0.5, -1.0]])
    d = np.float32([[0, 0, 1]])
    t, hit, steps = oracle.sdf_trace(grid, (1 / 8,) * 3, o, d, max_steps=3)
    assert hit[0] == 0 and steps[0] == 3 and t[0] == np.float32(1e20)
    t, hit, steps = oracle.sdf_trace(grid, (1 / 8,) * 3, o, d)
    assert hit[0] == 0 and steps[0] == 3                   # 0.4 per step across a unit box
    o2 = np.float32([[0.5, 2.0, 0.5]])
    t, hit, steps = oracle.sdf_trace(grid, (1 / 8,) * 3, o2, np.float32([[0, 1, 0]]))
    assert hit[0] == 0 and steps[0] == 0                   # box behind the ray


@pytest.mark.slow
def test_fp32_march_against_fp64(perm):
    from rayengine_amd import workloads as W
    kw = dict(dim=(128, 128, 128), spacing=(1 / 128,) * 3)
    g = oracle.noise_fill(perm, **kw)
    g64 = oracle.noise_fill_f64(perm, **kw)
    assert np.abs(g - g64).max() < 5e-6
    o, d = W.terrain_rays(240, 135)
    t, h, s = oracle.sdf_trace(g, kw["spacing"], o, d)
    t64, h64, s64 = oracle.sdf_trace_f64(g64, kw["spacing"], o, d)
    assert (h == h64).mean() > 0.999
    both = (h == 1) & (h64 == 1)
    err = np.abs(t[both] - t64[both])
    assert np.percentile(err, 99) < 1e-3


def test_noise_origin_shifts_the_sampled_window(perm):
    """A chunk generated at origin k * spacing is the window of the grid at
    origin 0 that starts at voxel k (power-of-two spacing keeps both sample
    positions exact), and the origin moves y against the y_offset plane."""
    sp = (1.0 / 64, 1.0 / 64, 1.0 / 64)
    big = oracle.noise_fill(perm, dim=(24, 20, 28), spacing=sp)
    k = (4, 8, 12)
    chunk = oracle.noise_fill(perm, dim=(16, 12, 16), spacing=sp, origin=tuple(np.float32(ki / 64) for ki in k))
    assert np.array_equal(chunk, big[k[2]:k[2] + 16, k[1]:k[1] + 12, k[0]:k[0] + 16])
    # origin 0 is the default grid bit for bit
    assert np.array_equal(oracle.noise_fill(perm, dim=(8, 8, 8), spacing=sp, origin=(0, 0, 0)),
                          oracle.noise_fill(perm, dim=(8, 8, 8), spacing=sp))
