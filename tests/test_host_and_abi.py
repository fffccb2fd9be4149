"""CPU tests: host-side input producers, scene generators, and the C ABI
library (loads, exports every declared symbol; no compute without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from sparrowpy_amd import geometry, samplings, scenes, testing
from sparrowpy_amd.coordinates import Coordinates

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sparrow_hip.h")


# ---- subdivision (geometry.py:290-410) ------------------------------------------
def test_patch_counts():
    """tests/test_DRadiosityFast.py:10-16, 207-211."""
    pts = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]])
    assert geometry.total_number_of_patches(pts, 0.2) == 25
    walls = testing.shoebox_room_stub(1, 1, 1)
    wp = np.array([w.pts for w in walls])
    wn = np.array([w.normal for w in walls])
    p, n, P, ids = geometry.process_patches(wp, wn, .5)
    assert P == 24 and p.shape == (24, 4, 3) and n.shape == (24, 3)
    np.testing.assert_array_equal(ids, np.repeat(np.arange(6), 4))
    np.testing.assert_allclose(geometry.calculate_area(p), .25)
    # every patch lies on its wall's plane and inside its bounding box
    for k in range(P):
        w = wp[ids[k]]
        assert np.all(p[k] >= w.min(0) - 1e-12)
        assert np.all(p[k] <= w.max(0) + 1e-12)


def test_config_sizes():
    """Patch counts of the benchmark scenes."""
    def count(sc):
        wp = np.array([w.pts for w in sc.walls])
        wn = np.array([w.normal for w in sc.walls])
        return geometry.process_patches(wp, wn, sc.patch_size)[2]
    assert count(scenes.small_shoebox()) == 148
    sc2 = scenes.directional_shoebox()
    assert count(sc2) == 1072 and sc2.directions.csize == 64
    assert sc2.n_samples == 6000
    sc3 = scenes.street_canyon()
    assert 5000 <= count(sc3) <= 5200 and sc3.directions.csize == 32


def test_polygon_validation():
    with pytest.raises(ValueError):
        geometry.Polygon([[0, 0, 0], [1, 0, 0], [1, 1, 1], [0, 1, 0]],
                         [1, 0, 0], [0, 0, 1])
    with pytest.raises(ValueError):
        geometry.Polygon([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]],
                         [1, 0, 0], [1, 0, 0])


def test_brdf_frames_point_into_room():
    """tests/test_DRadiosityFast.py:153-197: rotated BRDF directions lie in
    each wall's positive half space; rotation keeps the up vector."""
    c = samplings.hemisphere(samplings.sph_gaussian(sh_order=3))
    for w in testing.shoebox_room_stub(2, 3, 4):
        d = geometry.rotate_coords_to_normal(w.normal, w.up_vector,
                                             c.cartesian)
        assert np.all(d @ w.normal > 0)
        np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1, atol=1e-14)
        z = geometry.rotate_coords_to_normal(w.normal, w.up_vector,
                                             [[0, 0, 1.], [1, 0, 0]])
        np.testing.assert_allclose(z[0], w.normal, atol=1e-14)
        np.testing.assert_allclose(z[1], w.up_vector, atol=1e-14)


def test_sampling_weights_and_order():
    g = samplings.sph_gaussian(sh_order=3)
    assert g.csize == 32
    np.testing.assert_allclose(g.weights.sum(), 1.0)
    h = samplings.hemisphere(g)
    assert h.csize == 16 and np.all(h.z > 0)
    np.testing.assert_allclose(np.linalg.norm(g.cartesian, axis=1), 1)


def test_coordinates_find_nearest_and_sub():
    c = Coordinates([0, 1, 0], [0, 0, 1], [1, 0, 0])
    (idx,), d = c.find_nearest(Coordinates([0.9], [0.1], [0.0]))
    assert idx[0] == 1
    diff = c - Coordinates(0, 0, 1)
    np.testing.assert_allclose(diff.radius, [0, np.sqrt(2), np.sqrt(2)])


# ---- C ABI ----------------------------------------------------------------------
def _declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from sparrowpy_amd.build import LIB
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    lib = ctypes.CDLL(LIB)
    names = _declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_binding_covers_header():
    from sparrowpy_amd import _hip
    assert set(_declared_symbols()) == set(_hip.SYMBOLS)


def test_library_pure_queries_without_gpu():
    """Size queries run on the host (no device call)."""
    from sparrowpy_amd import _hip
    lib = _hip.lib(require_gpu=False)
    assert lib.spr_padded_directions(32) == 32
    assert lib.spr_padded_directions(33) == 64
    assert lib.spr_padded_directions(65) == 0
    assert lib.spr_visibility_workspace(10) > 0
    assert lib.spr_energy_exchange_workspace(10, 4, 2, 100, 20) > 10 * 4 * 2
    assert b"gfx950" in lib.spr_version()


def test_ops_fail_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from sparrowpy_amd import _hip, ops
    with pytest.raises(_hip.HipUnavailable):
        ops.check_patch2patch_visibility(np.zeros((2, 3)),
                                         np.zeros((2, 3)),
                                         np.zeros((2, 4, 3)))


def test_state_pool_reuses_and_grows():
    """StatePool (the batched run's chunk buffers) on CPU tensors: takes
    are views of the k-th buffer, a rewind reuses the same memory, growth
    replaces every buffer at the larger size, and over-asking raises."""
    import torch
    from sparrowpy_amd import ops
    pool = ops.StatePool(torch.device("cpu"))
    pool.begin([10, 10, 4])
    a, b, c = pool.take(10), pool.take(7), pool.take(4)
    assert a.data_ptr() != b.data_ptr() != c.data_ptr()
    assert [t.numel() for t in (a, b, c)] == [10, 7, 4]
    with pytest.raises(RuntimeError):
        pool.take(1)                 # no fourth buffer
    first = [t.data_ptr() for t in pool.bufs]
    pool.begin([8, 10, 2])           # fits: same memory, rewound
    assert [t.data_ptr() for t in pool.bufs] == first
    assert pool.take(8).data_ptr() == first[0]
    with pytest.raises(RuntimeError):
        pool.take(11)                # beyond buffer 1
    pool.begin([12, 3])              # grows: old sizes kept where larger
    assert [t.numel() for t in pool.bufs] == [12, 10, 4]


def test_bench_refuses_several_gpus():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says
    why, before importing torch or touching a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
