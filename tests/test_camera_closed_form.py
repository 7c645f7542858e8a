"""Closed-form pins of the primary-ray generator (re_gen_camera_rays).

generate_ray (RayEngine/src/cuda-src/ray_functions.cuh:13-55) maps pixel
(x, y) of a width x height frame to sx = x / (width - 1), sy = 1 - y /
(height - 1) and the direction

    normalize(view + (2 sx - 1) tan(fov.x / 2) h + (2 sy - 1) tan(fov.y / 2) v)

with h = normalize(view x camdown), v = normalize(h x view), fov in degrees
as the full angle; the focal distance scales the vector and drops out after
normalisation; every origin is campos.  The bases below are worked out by
hand for the cameras used, so a swapped basis, a flipped cross product, a
full-angle tangent in place of the half angle, or a swapped sx / sy fails.
"""
import numpy as np
import pytest

import rayengine_amd as re


def _expected(view, h, v, fov, w, hgt):
    view = np.asarray(view, np.float64)
    view = view / np.linalg.norm(view)
    tx, ty = np.tan(np.radians(fov[0]) / 2), np.tan(np.radians(fov[1]) / 2)
    y, x = np.mgrid[0:hgt, 0:w]
    sx = x / (w - 1.0)
    sy = 1.0 - y / (hgt - 1.0)
    d = (view[None, None, :] + ((2 * sx - 1) * tx)[..., None] * np.asarray(h, np.float64)
         + ((2 * sy - 1) * ty)[..., None] * np.asarray(v, np.float64))
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    return d.reshape(-1, 3)


# (view, camdown, h = view x camdown, v = h x view), by hand
CAMERAS = [
    ((0.0, 0.0, -1.0), (0.0, -1.0, 0.0), (-1.0, 0.0, 0.0), (0.0, -1.0, 0.0)),
    ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0)),
    ((0.0, -2.0, 0.0), (0.0, 0.0, 3.0), (-1.0, 0.0, 0.0), (0.0, 0.0, 1.0)),
]


@pytest.mark.parametrize("view,down,h,v", CAMERAS)
@pytest.mark.parametrize("fov", [(60.0, 60.0), (90.0, 40.0)])
@pytest.mark.parametrize("focal", [4.0, 0.5])
def test_directions_match_closed_form(view, down, h, v, fov, focal):
    w, hgt = 33, 21
    cam = re.Camera(pos=(0.3, -1.2, 4.0), view=view, camdown=down, fov=fov, focal_distance=focal)
    o, d = re.camera_rays(cam, w, hgt)
    assert np.array_equal(o, np.tile(np.float32([0.3, -1.2, 4.0]), (w * hgt, 1)))
    want = _expected(view, h, v, fov, w, hgt)
    assert np.abs(d - want).max() < 2e-6
    # the centre pixel looks along the normalised view
    c = (hgt // 2) * w + w // 2
    vn = np.asarray(view) / np.linalg.norm(view)
    assert np.abs(d[c] - vn).max() < 1e-6


def test_full_hd_corners_and_wrong_forms_fail():
    """1920 x 1080 with fov (60, 60), as the reference frame is traced: the
    four corners are normalize(view -+ tan30 h +- tan30 v)."""
    w, hgt = 1920, 1080
    view, h, v = np.array([0, 0, -1.0]), np.array([-1.0, 0, 0]), np.array([0, -1.0, 0])
    o, d = re.camera_rays(re.Camera(), w, hgt)
    t = np.tan(np.radians(30.0))
    corners = {0: view - t * h + t * v, w - 1: view + t * h + t * v,
               (hgt - 1) * w: view - t * h - t * v, hgt * w - 1: view + t * h - t * v}
    for i, want in corners.items():
        assert np.abs(d[i] - want / np.linalg.norm(want)).max() < 1e-6
    full = _expected(view, h, v, (60.0, 60.0), w, hgt)
    assert np.abs(d - full).max() < 2e-6
    # the plausible mistakes are far outside the tolerance
    wrong = {
        "full-angle tangent": _expected(view, h, v, (120.0, 120.0), w, hgt),
        "swapped basis": _expected(view, v, h, (60.0, 60.0), w, hgt),
        "flipped cross product": _expected(view, -h, -v, (60.0, 60.0), w, hgt),
    }
    for name, alt in wrong.items():
        assert np.abs(d - alt).max() > 0.1, name
