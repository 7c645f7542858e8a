"""GPU brush edits and .rsdf files through the C ABI: edits bit-identical to
the oracle on distance and material; files in the reference's layout (int32
dims, float spacing, then (distance, material) float pairs, x fastest)."""
import struct
import sys

import numpy as np
import pytest

import oracle
import rayengine_amd as re

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


DIM = (44, 37, 29)     # not multiples of the 8^3 brick
SP = (1 / 32, 1 / 32, 1 / 32)
BRUSHES = [
    ("sphere_add", dict(center=(0.7, 0.45, 0.4), extent=(0.2, 0, 0), material=3.0)),
    ("sphere_add", dict(center=(0.3, 0.6, 0.5), extent=(0.15, 0, 0), material=2.0, smooth_k=0.05)),
    ("sphere_subtract", dict(center=(0.5, 0.5, 0.45), extent=(0.25, 0, 0), material=1.0)),
    ("cube_add", dict(center=(0.9, 0.3, 0.2), extent=(0.1, 0.2, 0.15), material=0.5)),
    ("cube_subtract", dict(center=(0.2, 0.2, 0.6), extent=(0.12, 0.1, 0.2), material=4.0)),
    ("sphere_add", dict(center=(0.61, 0.35, 0.43), extent=(0.125, 0, 0), material=6.0, snap=True)),
]


def _noise(ctx):
    return ctx.noise_volume(dim=DIM, spacing=SP, seed=7)


def test_edits_match_oracle(gpu_ctx):
    vol = _noise(gpu_ctx)
    d = vol.read()
    m = vol.read_material()
    assert (m == 0).all()
    for kind, kw in BRUSHES:
        say(f"brush {kind} {kw}")
        vol.edit(kind, **kw)
        d, m = oracle.volume_edit(d, m, SP, kind, kw["center"], kw["extent"], kw["material"], kw.get("snap", False),
                                  kw.get("smooth_k", 0.1))
        gd, gm = vol.read(), vol.read_material()
        assert np.array_equal(gd.view(np.uint32), d.view(np.uint32)), f"{kind}: {(gd != d).sum()} distances differ"
        assert np.array_equal(gm.view(np.uint32), m.view(np.uint32)), f"{kind}: materials differ"
    assert len(np.unique(m)) >= 4
    vol.release()


def test_edit_rejects_bad_brushes(gpu_ctx):
    vol = _noise(gpu_ctx)
    with pytest.raises(re.ReError):
        vol.edit(7)
    with pytest.raises(re.ReError):
        vol.edit("sphere_add", smooth_k=0.0)
    with pytest.raises(re.ReError):
        vol.edit("cube_add", center=(np.nan, 0, 0))
    with pytest.raises(re.ReError):
        vol.edit("sphere_add", extent=(0, 1, 1), snap=True)
    vol.release()


def test_rsdf_write_layout_and_round_trip(gpu_ctx, tmp_path):
    vol = _noise(gpu_ctx)
    vol.edit("sphere_add", center=(0.5, 0.5, 0.5), extent=(0.2, 0, 0), material=2.0)
    d, m = vol.read(), vol.read_material()
    path = tmp_path / "a.rsdf"
    vol.save_rsdf(path)
    raw = path.read_bytes()
    n = DIM[0] * DIM[1] * DIM[2]
    assert len(raw) == 24 + 8 * n
    assert struct.unpack("<3i3f", raw[:24]) == (*DIM, *[float(np.float32(s)) for s in SP])
    pairs = np.frombuffer(raw[24:], np.float32).reshape(-1, 2)
    assert np.array_equal(pairs[:, 0], d.reshape(-1)) and np.array_equal(pairs[:, 1], m.reshape(-1))
    back = gpu_ctx.load_rsdf(path)
    assert tuple(back.dim) == DIM and np.array_equal(back.spacing, np.float32(SP))
    assert np.array_equal(back.read(), d) and np.array_equal(back.read_material(), m)
    # a second round trip writes the same bytes
    path2 = tmp_path / "b.rsdf"
    back.save_rsdf(path2)
    assert path2.read_bytes() == raw
    back.release()
    vol.release()


def test_rsdf_reads_a_hand_built_file(gpu_ctx, tmp_path):
    dim = (3, 2, 2)
    vals = np.arange(12, dtype=np.float32) * 0.5 - 2.0     # x fastest
    mats = (np.arange(12) % 3).astype(np.float32)
    raw = struct.pack("<3i3f", *dim, 0.5, 0.25, 1.0) + np.stack([vals, mats], 1).astype("<f4").tobytes()
    path = tmp_path / "hand.rsdf"
    path.write_bytes(raw)
    v = gpu_ctx.load_rsdf(path)
    assert np.array_equal(v.read().reshape(-1), vals) and np.array_equal(v.read_material().reshape(-1), mats)
    assert v.read()[1, 0, 2] == vals[2 + 3 * (0 + 2 * 1)]
    # the loaded volume is traceable: a ray along +x near y = z = 0 enters where the field is about -2
    t, h, _ = v.trace(np.float32([[-1.0, 0.05, 0.05]]), np.float32([[1.0, 0.0, 0.0]]))
    assert h[0] == 1
    v.release()


def test_rsdf_errors(gpu_ctx, tmp_path):
    with pytest.raises(re.ReError):
        gpu_ctx.load_rsdf(tmp_path / "missing.rsdf")
    good = struct.pack("<3i3f", 2, 2, 2, 1.0, 1.0, 1.0) + np.zeros(16, "<f4").tobytes()
    for name, raw in (("short.rsdf", good[:-4]), ("long.rsdf", good + b"\0"), ("tiny.rsdf", good[:10]),
                      ("dims.rsdf", struct.pack("<3i3f", 1, 2, 2, 1.0, 1.0, 1.0) + np.zeros(8, "<f4").tobytes()),
                      ("spacing.rsdf", struct.pack("<3i3f", 2, 2, 2, 0.0, 1.0, 1.0) + good[24:])):
        p = tmp_path / name
        p.write_bytes(raw)
        with pytest.raises(re.ReError):
            gpu_ctx.load_rsdf(p)
    p = tmp_path / "ok.rsdf"
    p.write_bytes(good)
    gpu_ctx.load_rsdf(p).release()
