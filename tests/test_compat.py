"""The reference-named C++ API (include/rayengine_compat.hpp) compiles and
links against librayengine.so, builds trees on the CPU, and on the GPU traces
the same hits as the Python host mirror."""
import json
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

import rayengine_amd as re

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def demo(tmp_path_factory):
    out = tmp_path_factory.mktemp("compat") / "compat_demo"
    lib_dir = ROOT / "rayengine_amd"
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "compat_demo.cpp"),
           "-o", str(out), f"-L{lib_dir}", "-l:librayengine.so", f"-Wl,-rpath,{lib_dir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_compat_builds_tree_on_cpu(demo):
    r = subprocess.run([str(demo), "cpu"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.splitlines()[0])
    tree = re.KDTree(re.icosphere(2, 1.0))
    assert info["n_tris"] == 320 and info["n_leaves"] == tree.info["n_leaves"]
    assert info["n_trirecs"] == tree.info["n_trirecs"]


@pytest.mark.gpu
def test_compat_trace_matches_python_api(demo, gpu_ctx, tmp_path):
    out = tmp_path / "hits.bin"
    r = subprocess.run([str(demo), "gpu", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rec = np.fromfile(out, dtype=np.uint32).reshape(-1, 2)
    tree = re.KDTree(re.icosphere(2, 1.0))
    rot = float(np.float32(0.3))   # the demo's float rotation.y
    scene = gpu_ctx.upload_scene([re.SceneObject(tree, location=(0.25, -0.1, 0.0), rotation_y=rot)])
    o, d = re.camera_rays(re.Camera(pos=(0.0, 0.0, 5.0)), 64, 48)
    p, t = scene.trace_closest(o, d)
    assert np.array_equal(rec[:, 0].view(np.int32), p)
    assert np.array_equal(rec[:, 1], t.view(np.uint32))
    assert (p >= 0).sum() > 100


@pytest.mark.gpu
def test_compat_shadow_brush_files_and_frame(demo, gpu_ctx, tmp_path):
    """trace_shadow_rays, insert_sphere_to_texture, save_map / load_map and
    cuda_render_frame through the compat header agree with the C-ABI calls
    made from Python."""
    out = tmp_path / "hits.bin"
    r = subprocess.run([str(demo), "gpu", str(out), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    occ = [int(x) for x in [l for l in lines if l.startswith("shadow")][0].split()[1:]]
    tree = re.KDTree(re.icosphere(2, 1.0))
    rot = float(np.float32(0.3))
    scene = gpu_ctx.upload_scene([re.SceneObject(tree, location=(0.25, -0.1, 0.0), rotation_y=rot)])
    so = np.float32([[0.1 * i, 0.05 * i, 4.0] for i in range(8)])
    sd = np.float32([[0, 0, -1 if i < 4 else 1] for i in range(8)])
    assert occ == list(scene.occluded(so, sd))
    assert occ[:4] == [1, 1, 1, 1] and occ[4:] == [0, 0, 0, 0]
    scene.release()
    assert "brush 1" in lines
    # the reloaded map is the edited volume; the frame is the Python pipeline's
    vol = gpu_ctx.load_rsdf(tmp_path / "map.rsdf")
    assert tuple(vol.dim) == (64, 64, 64) and (vol.read_material() == 3).any()
    o, d = re.camera_rays(re.Camera(pos=(0.5, 0.95, -0.35), view=(0.0, -0.55, 1.0)), 96, 64)
    t, h, _, nrm = vol.trace(o, d, normals=True)
    px = re.frame_shade(gpu_ctx, d, h, nrm, 96, 64)
    demo_px = np.fromfile(tmp_path / "frame.bin", np.uint32).reshape(64, 96)
    assert np.array_equal(demo_px, px)
    assert (tmp_path / "frame.ppm").read_bytes()[:12] == b"P6\n96 64\n255"
    vol.release()
