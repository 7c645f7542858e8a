"""Pins the oracle's SDF brush edit (insert_sphere_to_texture with sdf_smin,
sdf_fmaxf, componentwise fminf, sphere and box distances) against closed
forms evaluated in FP64, and checks the smooth minimum's defining
properties."""
import numpy as np

import oracle

SP = (0.25, 0.25, 0.25)
DIM = (20, 24, 28)  # x, y, z


def _points():
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in DIM[::-1]], indexing="ij")
    return x * SP[0], y * SP[1], z * SP[2]


def _field(seed=0):
    rng = np.random.default_rng(seed)
    return (rng.uniform(-2, 2, DIM[::-1]).astype(np.float32), rng.integers(0, 4, DIM[::-1]).astype(np.float32))


def _box64(px, py, pz, s):
    d = [np.abs(px) - s[0], np.abs(py) - s[1], np.abs(pz) - s[2]]
    out = np.sqrt(sum(np.maximum(v, 0.0) ** 2 for v in d))
    return out + np.minimum(np.maximum(d[0], np.maximum(d[1], d[2])), 0.0)


C = (2.1, 3.05, 3.3)


def test_sphere_add_on_empty_field_is_the_sphere():
    far = np.full(DIM[::-1], 1e3, np.float32)
    d, m = oracle.volume_edit(far, np.zeros_like(far), SP, "sphere_add", C, (1.2, 0, 0), material=5)
    x, y, z = _points()
    ref = np.sqrt((x - C[0]) ** 2 + (y - C[1]) ** 2 + (z - C[2]) ** 2) - 1.2
    assert np.abs(d - ref).max() < 2e-6
    assert (m == 5).all()


def test_smooth_min_properties():
    a, am = _field(1)
    d, m = oracle.volume_edit(a, am, SP, "sphere_add", C, (1.2, 0, 0), material=7, smooth_k=0.5)
    x, y, z = _points()
    b = np.sqrt((x - C[0]) ** 2 + (y - C[1]) ** 2 + (z - C[2]) ** 2) - 1.2
    a64 = a.astype(np.float64)
    h = np.maximum(0.5 - np.abs(a64 - b), 0.0) / 0.5
    ref = np.minimum(a64, b) - h * h * 0.5 / 4
    assert np.abs(d - ref).max() < 1e-5
    assert (d <= np.minimum(a64, b) + 1e-6).all()                  # never above the minimum
    far = np.abs(a64 - b) > 0.5 + 1e-5
    assert np.abs(d[far] - np.minimum(a64, b)[far]).max() < 1e-6   # exact minimum outside the band
    assert (m == np.where(a64 < b, am, 7)).mean() > 0.999           # material of the smaller distance


def test_subtract_is_the_larger_of_field_and_negated_brush():
    a, am = _field(2)
    x, y, z = _points()
    for kind, brush in (("sphere_subtract", np.sqrt((x - C[0]) ** 2 + (y - C[1]) ** 2 + (z - C[2]) ** 2) - 1.2),
                        ("cube_subtract", _box64(x - C[0], y - C[1], z - C[2], (1.0, 0.6, 0.8)))):
        d, m = oracle.volume_edit(a, am, SP, kind, C, (1.2 if kind == "sphere_subtract" else 1.0, 0.6, 0.8),
                                  material=9)
        ref = np.maximum(a.astype(np.float64), -brush)
        assert np.abs(d - ref).max() < 2e-6
        win = a.astype(np.float64) > -brush + 1e-5
        assert (m[win] == am[win]).all() and (m[-brush > a + 1e-5] == 9).all()


def test_cube_add_is_componentwise_min():
    a, am = _field(3)
    x, y, z = _points()
    d, m = oracle.volume_edit(a, am, SP, "cube_add", C, (1.0, 0.6, 0.8), material=1.5)
    ref = np.minimum(a.astype(np.float64), _box64(x - C[0], y - C[1], z - C[2], (1.0, 0.6, 0.8)))
    assert np.abs(d - ref).max() < 2e-6
    assert np.array_equal(m, np.minimum(am, np.float32(1.5)))     # material min, not the winner's material


def test_box_distance_known_values():
    far = np.full((3, 3, 3), 1e3, np.float32)
    d, _ = oracle.volume_edit(far, np.zeros_like(far), (1, 1, 1), "cube_add", (1, 1, 1), (0.5, 0.5, 0.5))
    assert d[1, 1, 1] == -0.5                                      # centre: -half extent
    assert d[1, 1, 2] == 0.5                                       # one voxel off along x: face distance
    assert abs(d[2, 2, 2] - np.sqrt(3) * 0.5) < 1e-6               # corner direction


def test_snap_is_identity_on_lattice_offsets():
    a, am = _field(4)
    sp = (1.0, 1.0, 1.0)
    d0, m0 = oracle.volume_edit(a, am, sp, "sphere_add", (5.0, 4.0, 6.0), (1.0, 0, 0), material=2)
    d1, m1 = oracle.volume_edit(a, am, sp, "sphere_add", (5.0, 4.0, 6.0), (1.0, 0, 0), material=2, snap=True)
    assert np.array_equal(d0, d1) and np.array_equal(m0, m1)
    # off the lattice the snapped x/z offsets are integers at or below the true offset
    d2, _ = oracle.volume_edit(np.full_like(a, 1e3), np.zeros_like(am), sp, "sphere_add", (5.3, 4.0, 6.6),
                               (1.0, 0, 0), snap=True)
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in DIM[::-1]], indexing="ij")  # spacing 1
    px, pz = x - 5.3, z - 6.6
    sx = np.floor(px - np.abs(np.fmod(px, 1.0)))
    sz = np.floor(pz - np.abs(np.fmod(pz, 1.0)))
    assert (sx <= px).all() and (sz <= pz).all()
    ref = np.sqrt(sx ** 2 + (y - 4.0) ** 2 + sz ** 2) - 1.0
    assert np.abs(d2 - ref).max() < 2e-5
