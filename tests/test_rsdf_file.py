""".rsdf file checks that need no device (re_rsdf_info): the reference's header
(int32 dims, float spacing; renderer.cu:873-890, Grid.cpp:23-64) followed by
exactly dim.x * dim.y * dim.z (distance, material) float pairs.  A header
that claims more than the file holds is rejected before anything of its size
is allocated."""
import numpy as np
import pytest

import rayengine_amd as re


def _write(path, dims, spacing, n_pairs=None, extra=b""):
    dims = np.asarray(dims, np.int32)
    n = int(np.prod(dims.astype(np.int64))) if n_pairs is None else n_pairs
    with open(path, "wb") as f:
        f.write(dims.tobytes())
        f.write(np.asarray(spacing, np.float32).tobytes())
        f.write(np.arange(2 * n, dtype=np.float32).tobytes())
        f.write(extra)


def test_well_formed_file(tmp_path):
    p = tmp_path / "ok.rsdf"
    _write(p, (3, 4, 5), (0.5, 0.25, 0.125))
    dim, sp = re.rsdf_info(p)
    assert dim == (3, 4, 5) and sp == (0.5, 0.25, 0.125)


def test_header_claiming_4096_cubed_on_a_tiny_file(tmp_path):
    p = tmp_path / "huge.rsdf"
    _write(p, (4096, 4096, 4096), (1.0, 1.0, 1.0), n_pairs=4)     # 24 + 32 bytes
    with pytest.raises(re.ReError, match="shorter"):
        re.rsdf_info(p)


@pytest.mark.parametrize("dims,spacing,n_pairs,extra,what", [
    ((3, 4, 5), (1, 1, 1), 59, b"", "shorter"),
    ((3, 4, 5), (1, 1, 1), None, b"\0", "longer"),
    ((1, 4, 5), (1, 1, 1), None, b"", "dims"),
    ((4097, 2, 2), (1, 1, 1), 0, b"", "dims"),
    ((-3, 4, 5), (1, 1, 1), 0, b"", "dims"),
    ((3, 4, 5), (1, 0, 1), None, b"", "spacing"),
    ((3, 4, 5), (1, float("nan"), 1), None, b"", "spacing"),
])
def test_malformed_files(tmp_path, dims, spacing, n_pairs, extra, what):
    p = tmp_path / "bad.rsdf"
    _write(p, dims, spacing, n_pairs, extra)
    with pytest.raises(re.ReError, match=what):
        re.rsdf_info(p)


def test_short_header_and_missing_file(tmp_path):
    p = tmp_path / "short.rsdf"
    p.write_bytes(np.int32([4, 4, 4]).tobytes() + b"\0\0")
    with pytest.raises(re.ReError, match="short header"):
        re.rsdf_info(p)
    with pytest.raises(re.ReError, match="cannot open"):
        re.rsdf_info(tmp_path / "absent.rsdf")
