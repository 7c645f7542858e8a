"""What the object-rotation seam changes (ray_functions.cuh:238-239).

The reference evaluates sinf/cosf of the object's Y rotation on the device
for every ray; this library takes the sine and cosine as object fields
(re_object.sin_rot_y / cos_rot_y), and the reference-shaped header fills
them with the correctly rounded float of the double-precision value.  A
device sinf/cosf may differ from that by an ulp or two.  These oracle runs
show what such a difference does to a rotated object's hits: t moves by a
few parts in 1e6 (most hit t's change bits), and a prim id changes only for
rays grazing a silhouette.  A caller that needs the reference's exact
values passes its own sine and cosine in the object record."""
import numpy as np
import pytest

import oracle
import rayengine_amd as re


def _step(x, k):
    for _ in range(abs(k)):
        x = np.nextafter(x, np.float32(np.inf if k > 0 else -np.inf), dtype=np.float32)
    return x


@pytest.mark.parametrize("soup", ["sphere", "random"])
def test_sin_cos_ulps_move_t_not_prims(soup):
    tri_v = re.icosphere(4, 1.0) if soup == "sphere" else re.random_tris(5, 8000, -1.5, 1.5, 0.05, 0.3)
    tree = re.KDTree(tri_v)
    nodes, leaves, tris = tree.export()
    o, d = re.camera_rays(re.Camera(pos=(0.3, 0.5, 6.0), view=(0.0, -0.1, -1.0)), 192, 144)
    flips, bit_equal, worst = 0, [], 0.0
    for ang in (0.7, -1.3, 2.1):
        ob = tree.identity_object()
        ob["loc"] = (0.2, -0.1, 0.05)
        s, c = np.float32(np.sin(ang)), np.float32(np.cos(ang))
        ob["sin_y"], ob["cos_y"] = s, c
        p0, t0 = oracle.trace(nodes, leaves, tris, ob, o, d, secondary=True)
        assert (p0 >= 0).sum() > 1500
        for k in (1, 2, -1, -2):
            ob2 = ob.copy()
            ob2["sin_y"], ob2["cos_y"] = _step(s, k), _step(c, -k)
            p1, t1 = oracle.trace(nodes, leaves, tris, ob2, o, d, secondary=True)
            flips += int((p0 != p1).sum())
            both = (p0 >= 0) & (p0 == p1)
            rel = np.abs(t1[both].astype(np.float64) - t0[both]) / t0[both]
            worst = max(worst, float(rel.max()))
            bit_equal.append(float((t0[both].view(np.uint32) == t1[both].view(np.uint32)).mean()))
    n = 3 * 4 * len(o)
    assert flips <= n * 5e-4, f"{flips} prim ids of {n} changed"
    assert worst < 1e-4, worst
    assert np.mean(bit_equal) < 0.9          # the seam is visible in t's bits

