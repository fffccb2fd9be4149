"""BRDFs from directional scattering coefficients
(``create_from_directional_scattering``, reference ``brdf.py:133-224``):
rows that are neither a constant plus one specular entry nor shared with
other walls' rows.  The class path picks the grouped plan (the rows are
shared by every source of a wall with the same incoming direction); every
plan kind and the reference-shaped dense entry point must meet the FP64
oracle at the exchange tolerance."""
import numpy as np
import pytest

import oracle as orc
from conftest import make_oracle, rel_err_within_db
from sparrowpy_amd import brdf as B
from sparrowpy_amd import samplings, testing
from sparrowpy_amd.coordinates import Coordinates, FrequencyData

pytestmark = pytest.mark.gpu

EXCHANGE_RTOL = 1e-5
FREQS = np.array([500., 1000.])
C, DT, T, K = 343., 1 / 4000, 600, 6
SOURCE = np.array([1.1, .8, .6])


@pytest.fixture(scope="module")
def torch():
    import torch as _torch
    if not _torch.cuda.is_available():
        pytest.skip("no GPU")
    return _torch


def _brdfs(n_walls, dirs, seed=4):
    rng = np.random.default_rng(seed)
    S = dirs.cartesian.shape[0]
    out = []
    for _ in range(n_walls):
        ds = rng.uniform(.2, 1., (S, S, len(FREQS)))
        ds /= ds.sum(axis=1, keepdims=True)
        out.append(B.create_from_directional_scattering(
            dirs, dirs, FrequencyData(ds, FREQS),
            FrequencyData(rng.uniform(.05, .3, (1, len(FREQS))), FREQS)))
    return out


@pytest.fixture(scope="module")
def scene(torch):
    walls = testing.shoebox_room_stub(3, 2.5, 2)
    dirs = samplings.hemisphere(samplings.sph_gaussian(sh_order=3))
    brdfs = _brdfs(len(walls), dirs)
    o = make_oracle(walls, .5)
    for w, br in enumerate(brdfs):
        o.set_wall_brdf([w], br.freq, dirs.cartesian, dirs.cartesian, FREQS)
    o.set_air_attenuation(np.array([1e-3, 3e-3]), FREQS)
    o.bake_geometry()
    o.init_source_energy(SOURCE)
    ref = orc.exchange(T, o.energy_init_source, o.distance_patches_to_source,
                       o.pair_distances(), o.fft, o.p2o, C, DT, K,
                       o.visible_patches, nthreads=8)
    return walls, dirs, brdfs, o, ref


def _gpu(walls, dirs, brdfs, tile_rows=4):
    from sparrowpy_amd import DirectionalRadiosityFast
    r = DirectionalRadiosityFast.from_polygon(walls, .5)
    for w, br in enumerate(brdfs):
        r.set_wall_brdf([w], br, dirs, dirs)
    r.set_air_attenuation(FrequencyData(np.array([1e-3, 3e-3]), FREQS))
    r.TILE_ROWS = tile_rows
    r.bake_geometry()
    r.init_source_energy(Coordinates(*SOURCE))
    return r


@pytest.mark.parametrize("kind,tile_rows", [(None, 4), ("grouped", 16),
                                            ("generic", 0)])
def test_directional_scattering_exchange(scene, kind, tile_rows):
    from sparrowpy_amd import ops
    walls, dirs, brdfs, o, ref = scene
    r = _gpu(walls, dirs, brdfs, tile_rows)
    plan = r._exchange_plan(C, DT, kind)
    if kind is None:
        # not diffuse + specular; shared rows per wall: the grouped plan
        assert isinstance(plan, ops.GroupedPlan)
        assert not plan.diffuse_specular
    r.calculate_energy_exchange(C, DT, T * DT, K)
    err = rel_err_within_db(r._energy_exchange_etc, ref)
    print(f"[dirscat] plan {kind} tiles {tile_rows}: max rel err {err:.2e}")
    assert err <= EXCHANGE_RTOL


def test_directional_scattering_dense_entry(scene):
    """spr_energy_exchange on the reference's dense F~ (the rebinding shim
    of the integration notes)."""
    from sparrowpy_amd import ops
    walls, dirs, brdfs, o, ref = scene
    out = ops.energy_exchange(T, o.energy_init_source,
                              o.distance_patches_to_source,
                              o.pair_distances(), o.fft, o.p2o, C, DT, K,
                              o.visible_patches).cpu().numpy()
    err = rel_err_within_db(out, ref)
    print(f"[dirscat] dense entry: max rel err {err:.2e}")
    assert err <= EXCHANGE_RTOL
