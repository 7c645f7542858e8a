"""GPU frame output against the oracle: a traced terrain frame (sky for the
misses, simple_shade for the hits) packed to RGBA8, and the PPM file."""
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


@pytest.mark.parametrize("opts", [dict(), dict(gamma=1, vignetting=1, base_color=(0.2, 0.3, 0.1),
                                                sun_dir=(0.0, 0.8, 0.6))])
def test_terrain_frame_matches_oracle(gpu_ctx, opts):
    vol = gpu_ctx.noise_volume(dim=(96, 96, 96), spacing=(1 / 96,) * 3)
    w, h = 320, 180
    o, d = W.terrain_rays(w, h)
    t, hit, _, nrm = vol.trace(o, d, normals=True)
    say(f"frame {w}x{h}: {int(hit.sum())} hits, options {opts}")
    px = re.frame_shade(gpu_ctx, d, hit, nrm, w, h, **opts)
    ref = oracle.frame_shade(oracle.frame_params(**opts), d, hit, nrm, w, h)
    a = px.view(np.uint8).reshape(-1, 4).astype(int)
    b = ref.view(np.uint8).reshape(-1, 4).astype(int)
    hits = hit.astype(bool)
    assert np.array_equal(a[hits], b[hits])                       # hit pixels: exact arithmetic, identical
    diff = np.abs(a - b).max(1)
    say(f"sky: largest difference {diff.max()} step(s), {(diff > 0).mean():.5f} of the pixels differ")
    assert diff.max() <= 1 and (diff > 0).mean() < 0.01           # sky: libm vs hardware exp/rcp/sqrt, one step
    assert (a[:, 3] == 255).all() and 0 < hits.mean() < 1
    vol.release()


def test_ppm_file(gpu_ctx, tmp_path):
    rng = np.random.default_rng(1)
    px = rng.integers(0, 2 ** 32, (5, 7), dtype=np.uint64).astype(np.uint32)
    path = tmp_path / "f.ppm"
    re.write_ppm(path, px)
    raw = path.read_bytes()
    head = b"P6\n7 5\n255\n"
    assert raw[:len(head)] == head and len(raw) == len(head) + 5 * 7 * 3
    rgb = np.frombuffer(raw[len(head):], np.uint8).reshape(5, 7, 3)
    assert np.array_equal(rgb, px.view(np.uint8).reshape(5, 7, 4)[..., :3])


def test_frame_rejects_bad_parameters(gpu_ctx):
    d = np.zeros((4, 3), np.float32)
    with pytest.raises(re.ReError):
        re.frame_shade(gpu_ctx, d, np.zeros(4, np.uint8), d, 2, 2, atmosphere_radius=1.0)
    with pytest.raises(re.ReError):
        re.frame_shade(gpu_ctx, d, np.zeros(4, np.uint8), d, 2, 2, hm=0.0)
