"""Host logic of the destination tilings, the directional-scattering canyon
and shoebox scenes and the profile-name parsing of bench.py, on the CPU (no
kernels)."""
import os
import sys

import numpy as np
import pytest

from sparrowpy_amd import DirectionalRadiosityFast, ops, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _small_canyon(**kw):
    return scenes.street_canyon(seed=5, length=12., width=6., height=5.,
                                n_recesses=3, **kw)


@pytest.mark.parametrize("rows,block", [(16, (4, 4)), (8, (2, 4)),
                                        (4, (2, 2))])
def test_auto_tile_block_follows_tile_rows(rows, block):
    """TILE_BLOCK = "auto" orders the destinations in the patch blocks of
    ops.block_order for the tile's row count."""
    sc = _small_canyon()
    r = DirectionalRadiosityFast.from_polygon(sc.walls, sc.patch_size)
    assert r.TILE_BLOCK == "auto"
    r.TILE_ROWS = rows
    want = ops.block_order(r._patches_points, r._patch_to_wall_ids, block,
                           pad=r.TILE_PAD)
    np.testing.assert_array_equal(r._tile_order(), want)
    r.TILE_PAD = False
    want = ops.block_order(r._patches_points, r._patch_to_wall_ids, block,
                           pad=False)
    np.testing.assert_array_equal(r._tile_order(), want)


def test_tile_block_choices():
    sc = _small_canyon()
    r = DirectionalRadiosityFast.from_polygon(sc.walls, sc.patch_size)
    r.TILE_ROWS = 2               # no automatic block for 2 rows
    assert r._tile_order() is None
    r.TILE_BLOCK = None
    r.TILE_ROWS = 16
    assert r._tile_order() is None
    r.TILE_BLOCK = (2, 2)         # 4 patches for 16 rows
    with pytest.raises(ValueError, match="TILE_ROWS"):
        r._tile_order()
    r.TILE_BLOCK = "square"
    with pytest.raises(ValueError, match="auto"):
        r._tile_order()


def test_batched_tiling_resolution():
    """BatchedExchange's (rows, block, bins per lane): "auto" blocks as the
    class's, explicit blocks checked against the row count."""
    from sparrowpy_amd.batch import BatchedExchange
    t = BatchedExchange._tiling
    assert t(16, "auto", 8) == (16, (4, 4), 8)
    assert t(8, "auto", 8) == (8, (2, 4), 8)
    assert t(4, "auto", 16) == (4, (2, 2), 16)
    assert t(0, None, 8) == (0, None, 8)
    assert t(16, [2, 8], 16) == (16, (2, 8), 16)
    with pytest.raises(ValueError, match="tile_rows"):
        t(16, (2, 2), 8)
    with pytest.raises(ValueError, match="auto"):
        t(16, "square", 8)
    # every candidate the tuner tries resolves
    for cands in BatchedExchange.TILINGS.values():
        for c in cands:
            rows, block, r = t(*c)
            assert block is None or block[0] * block[1] == rows


def test_directional_canyon_scene():
    """street_canyon(brdf="directional"): the same geometry and sources as
    the scattering canyon; every wall's BRDF from a directional scattering
    coefficient whose rows sum to 1 over the outgoing directions, so that
    sum_o rho cos(theta_o) w_o = 1 - alpha for every incoming direction
    (brdf.py:133-224), with alpha drawn in [0.05, 0.3]."""
    a = _small_canyon()
    b = _small_canyon(brdf="directional")
    assert b.name == a.name + "-dirscat"
    assert len(a.walls) == len(b.walls)
    for wa, wb in zip(a.walls, b.walls):
        np.testing.assert_array_equal(wa.pts, wb.pts)
    np.testing.assert_array_equal(a.sources, b.sources)
    dirs = b.directions
    cos_o = dirs.cartesian[:, 2] / np.linalg.norm(dirs.cartesian, axis=1)
    w = np.asarray(dirs.weights, float)
    w = w * 2 * np.pi / w.sum()
    for (ia, bra), (ib, brb) in zip(a.brdf_groups, b.brdf_groups):
        assert ia == ib
        rho = np.asarray(brb.freq)             # [in, out, band]
        assert rho.shape == (dirs.csize, dirs.csize, len(b.frequencies))
        assert (rho > 0).all()
        energy = np.einsum("iob,o->ib", rho, cos_o * w)
        np.testing.assert_allclose(
            energy, np.broadcast_to(energy[:1], energy.shape), rtol=1e-12)
        assert ((energy > 1 - 0.3) & (energy < 1 - 0.05)).all()
        # rows differ from a constant plus one specular entry
        assert np.ptp(rho[0, :, 0]) > 0
    with pytest.raises(ValueError, match="brdf"):
        scenes.street_canyon(brdf="lambert")


def test_directional_shoebox_scene():
    """directional_shoebox(brdf="directional"): the six walls unsplit (the
    same 1,072 patches once subdivided), the same source and receivers as
    the scattering form, and per wall a BRDF whose rows carry the wall's
    1 - alpha as energy (brdf.py:133-224) and are not constant plus one
    specular entry."""
    from sparrowpy_amd import DirectionalRadiosityFast
    a = scenes.directional_shoebox()
    b = scenes.directional_shoebox(brdf="directional")
    assert b.name == a.name + "-dirscat" and len(b.walls) == 6
    np.testing.assert_array_equal(a.sources, b.sources)
    np.testing.assert_array_equal(a.receivers, b.receivers)
    ra = DirectionalRadiosityFast.from_polygon(a.walls, a.patch_size)
    rb = DirectionalRadiosityFast.from_polygon(b.walls, b.patch_size)
    assert ra.n_patches == rb.n_patches == 1072
    dirs = b.directions
    cos_o = dirs.cartesian[:, 2] / np.linalg.norm(dirs.cartesian, axis=1)
    w = np.asarray(dirs.weights, float)
    w = w * 2 * np.pi / w.sum()
    alpha, sd = b.materials
    assert alpha.shape == (6, 6) and sd.shape == (6, 64, 64, 6)
    for n, (idx, br) in enumerate(b.brdf_groups):
        assert idx == [n]
        rho = np.asarray(br.freq)
        energy = np.einsum("iob,o->ib", rho, cos_o * w)
        np.testing.assert_allclose(
            energy, np.broadcast_to(1 - alpha[n], energy.shape), rtol=1e-12)
        row = rho[0, :, 0]
        assert np.unique(np.round(row / row.max(), 12)).size > 2
    with pytest.raises(ValueError, match="brdf"):
        scenes.directional_shoebox(brdf="lambert")


def test_profile_kernel_names():
    """bench._kernel_name: rocprofv3's demangled names to the short names
    the counter profile is keyed by (template arguments kept)."""
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import bench
    k = bench._kernel_name
    assert k("void spr::(anonymous namespace)::tile_lds_gather_kernel<8, 16,"
             " 64, false>(long, int, int)") == \
        "tile_lds_gather_kernel<8, 16, 64, false>"
    assert k("spr::scan_exclusive_kernel(long*, long)") == \
        "scan_exclusive_kernel"
    assert k("spr::(anonymous namespace)::init_kernel(long, float*)") == \
        "init_kernel"
    assert k("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"


def test_profile_stamp_ignores_comments():
    """The counter profile is stamped with a hash of the kernels' code:
    comments and layout do not change it, code does."""
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import bench
    c = bench._code_only
    a = "int f(int x) {  // twice\n  return 2 * x; /* doubled */\n}\n"
    b = "/* header */\nint f(int x) {\n    return 2 * x;\n}"
    assert c(a) == c(b) == "int f(int x) { return 2 * x; }"
    assert c(a) != c(a.replace("2 * x", "3 * x"))
    assert len(bench.kernel_sources_sha()) == 16
