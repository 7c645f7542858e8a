import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")
    # torch first: it loads its bundled HIP runtime as libamdhip64.so, and
    # librayengine.so's libamdhip64.so.7 then resolves to that same copy by
    # soname.  Loaded the other way round the process holds two runtimes and
    # the library finds no device.
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def _gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_ctx():
    """One library context for the whole GPU session (device 0)."""
    if not _gpu_available():
        pytest.skip("no GPU")
    import rayengine_amd as re
    ctx = re.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def small():
    """Small icosphere workload: (tri_v, tree, ray_o, ray_d)."""
    from rayengine_amd import workloads as W
    return W.small_scene()


@pytest.fixture(scope="session")
def cluster():
    """The 2^20-triangle clustered soup and its tree: (tri_v, tree)."""
    import rayengine_amd as re
    from rayengine_amd import workloads as W
    print("building the 2^20-triangle cluster soup", flush=True)
    tri_v = W.cluster_soup_default()
    return tri_v, re.KDTree(tri_v)
