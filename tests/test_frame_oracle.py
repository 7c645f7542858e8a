"""Pins the oracle's frame tail: the sky integral against its closed form on
a radial ray (geometric sums over the 32 view and 8 light samples), the tone
curve's two branches meeting at 1.413, and the simple_shade / vignetting /
pack arithmetic on hand-checked pixels."""
import numpy as np
import pytest

import oracle


params = oracle.frame_params


def _radial_closed_form(er, ar, hr, hm, h0, beta_r, beta_m):
    """View ray and sun straight up from height h0: every sample height and
    segment length is closed form."""
    tmax = ar - (er + h0)
    seg = tmax / 32
    g = 0.76
    pr = 3 / (16 * np.pi) * 2
    pm = 3 / (8 * np.pi) * ((1 - g * g) * 2) / ((2 + g * g) * (1 + g * g - 2 * g) ** 1.5)
    h = h0 + (np.arange(32) + 0.5) * seg
    hr_i, hm_i = np.exp(-h / hr) * seg, np.exp(-h / hm) * seg
    odr, odm = np.cumsum(hr_i), np.cumsum(hm_i)
    segl = (tmax - (np.arange(32) + 0.5) * seg) / 8
    j = np.arange(8) + 0.5
    olr = np.array([np.sum(np.exp(-(hh + j * s) / hr) * s) for hh, s in zip(h, segl)])
    olm = np.array([np.sum(np.exp(-(hh + j * s) / hm) * s) for hh, s in zip(h, segl)])
    out = []
    for b in beta_r:
        att = np.exp(-(b * (odr + olr) + beta_m * 1.1 * (odm + olm)))
        out.append((np.sum(att * hr_i) * b * pr + np.sum(att * hm_i) + 100 * beta_m * pm) * 20)
    return np.array(out)


@pytest.mark.parametrize("h0", [100.0, 10000.0, 40000.0])
def test_sky_radial_ray_closed_form(h0):
    p = params(sky_origin=(0.0, 6360e3 + h0, 0.0))
    ref = _radial_closed_form(6360e3, 6420e3, 79945.0, 1200.0, np.float32(6360e3 + h0) - 6360e3,
                              oracle.ATM_BETA_R.astype(np.float64), float(oracle.ATM_BETA_M))
    got64 = oracle.sky(p, [[0, 1, 0]], f64=True)[0]
    got32 = oracle.sky(p, [[0, 1, 0]])[0]
    assert np.allclose(got64, ref, rtol=1e-9), (got64, ref)
    assert np.allclose(got32, ref, rtol=2e-3), (got32, ref)


def test_sky_fp32_close_to_fp64_over_directions():
    p = params(sun_dir=(0.0, 0.8, 0.6))
    rng = np.random.default_rng(0)
    d = rng.normal(size=(400, 3))
    d[:, 1] = np.abs(d[:, 1]) + 0.05        # upward rays (downward ones pass through the planet)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    a, b = oracle.sky(p, d), oracle.sky(p, d, f64=True)
    assert np.all(np.abs(a - b) <= 2e-3 * np.abs(b) + 1e-7)
    assert (b > 0).all()


def test_tone_curve():
    lo = float(oracle.sky_tone(np.nextafter(np.float32(1.413), np.float32(0))))
    hi = float(oracle.sky_tone(1.413))
    assert abs(lo - hi) < 2e-4                         # the two branches meet
    assert oracle.sky_tone(0.0) == 0.0
    xs = np.linspace(0, 8, 400, dtype=np.float32)
    ys = np.array([oracle.sky_tone(x) for x in xs])
    assert (np.diff(ys) >= 0).all() and ys[-1] < 1.0
    assert abs(oracle.sky_tone(1.0) - (0.38317 ** (1 / 2.2))) < 1e-6


def test_hit_pixels_pack_and_vignetting():
    d = np.float32([[0, 0, -1]] * 4)
    n = np.float32([[0, 0, 1], [0, 0, 1], [1, 0, 0], [0, 0, 1]])
    hit = np.uint8([1, 1, 1, 0])
    p = params(base_color=(0.5, 0.0, 0.25))
    px = oracle.frame_shade(p, d, hit, n, 2, 2).reshape(-1)
    # facing ratio 1/2, mixed in at 1/2: +0.25 on every channel; R in the low byte
    assert px[0] == px[1] == (0xFF << 24) | (int(0.5 * 255) << 16) | (int(0.25 * 255) << 8) | int(0.75 * 255)
    assert px[2] == (0xFF << 24) | (int(0.25 * 255) << 16) | int(0.5 * 255)     # normal across the ray: base only
    assert px[3] >> 24 == 0xFF
    v = oracle.frame_shade(params(base_color=(0.5, 0.0, 0.25), vignetting=1), d, hit, n, 2, 2).reshape(-1)
    assert v[0] == (0xFF << 24) | (int(0.25 * 255) << 16) | (int(0.125 * 255) << 8) | int(0.375 * 255)  # u = v = 0
    g = oracle.frame_shade(params(base_color=(0.5, 0.0, 0.25), gamma=1), d, hit, n, 2, 2).reshape(-1)
    assert g[0] & 0xFF == int(np.float32(0.75) ** np.float32(1 / 2.2) * 255)
