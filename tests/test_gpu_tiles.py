"""Destination tiles of the grouped plan (spr_plan_tile_*,
spr_exchange_gather_tiled).

Bar: the merged plan holds exactly the grouped plan's entries, slot by
slot and in the same order; the tiled gather's group sums are bit-identical
to the row-by-row gather's (each slot sums its groups in the same order);
the exchange on tiles meets the oracle at the exchange tolerance.
"""
import numpy as np
import pytest

import oracle as orc
from conftest import make_oracle, rel_err_within_db
from sparrowpy_amd import testing

from test_gpu_parity import _directional, _grouped_and_generic

pytestmark = pytest.mark.gpu

EXCHANGE_RTOL = 1e-5
LAG_MASK = 0x3fffffff


@pytest.fixture(scope="module")
def ops():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from sparrowpy_amd import ops as _ops
    return _ops


def _np(t):
    return t.detach().cpu().numpy()


# 148 patches (a partial tile of 8) and 62 (partial tiles of 4 and 8)
ROOMS = {"3x2.5x2": (3, 2.5, 2), "2.5x1.5x1": (2.5, 1.5, 1)}


def _room(name, per_wall=True, seed=3):
    r = make_oracle(testing.shoebox_room_stub(*ROOMS[name]), .5)
    _directional(r, per_wall=per_wall, seed=seed)
    r.bake_geometry()
    return r


@pytest.mark.parametrize("jt", [4, 8, 16])
@pytest.mark.parametrize("name", list(ROOMS))
def test_tile_plan_contents(ops, name, jt):
    r = _room(name)
    gp, _ = _grouped_and_generic(ops, r, 343., 1 / 2000)
    tp = ops.tile_plan(gp, jt, min_merge=0)
    P, B = gp.P, gp.B
    flag = ops.GroupedPlan.LAST_IN_GROUP
    rp, gs, gptr = _np(gp.row_ptr), _np(gp.grp_start), _np(gp.grp_ptr)
    rowid, lag, scal = _np(gp.rowid), _np(gp.lag), _np(gp.scal)
    tptr, mrow = _np(tp.tptr), _np(tp.mrow)
    mlag, mw = _np(tp.mlag), _np(tp.mw)
    n_tiles = -(-P // jt)
    assert tptr.shape == (n_tiles + 1,) and tptr[0] == 0
    assert tptr[-1] == tp.M and np.all(np.diff(tptr) >= 0)
    # M = number of distinct (tile, group key, source row)
    g_of_e = np.searchsorted(gs, np.arange(gp.Vd), side="right") - 1
    j_of_e = np.searchsorted(rp, np.arange(gp.Vd), side="right") - 1
    key = _np(gp.grp_key)[g_of_e].astype(np.int64)
    distinct = np.unique(((j_of_e // jt) << 44) | (key << 22) |
                         rowid.astype(np.int64))
    assert distinct.size == tp.M
    assert tp.M < gp.Vd
    for t in range(n_tiles):
        ms = slice(tptr[t], tptr[t + 1])
        for k in range(jt):
            j = t * jt + k
            words = mlag[ms, k]
            present = (words & LAG_MASK) != LAG_MASK
            if j >= P:
                assert not present.any() and not (words & flag).any()
                assert not mw[:, ms, k].any()
                continue
            es = slice(rp[j], rp[j + 1])
            # the slot's entries, in merged order, are the row's, in order
            np.testing.assert_array_equal(mrow[ms][present], rowid[es])
            np.testing.assert_array_equal(words[present] & ~flag,
                                          lag[es] & LAG_MASK)
            np.testing.assert_array_equal(mw[:, ms, k][:, present],
                                          scal[:, es])
            assert not mw[:, ms, k][:, ~present].any()
            # a group ends where the slot has seen exactly its entries
            seen = np.cumsum(present)[(words & flag) != 0]
            ends = gs[gptr[j] + 1:gptr[j + 1] + 1] - rp[j]
            np.testing.assert_array_equal(seen, ends)
    if jt == 16:
        # each merged entry's span: its row and its slots' delay range
        span = _np(tp.mspan)
        assert span.shape == (tp.M, 4)
        lags = np.where((mlag & LAG_MASK) != LAG_MASK, mlag & LAG_MASK,
                        -1).astype(np.int64)
        np.testing.assert_array_equal(span[:, 0], mrow)
        np.testing.assert_array_equal(span[:, 1], lags.max(axis=1))
        big = np.where(lags >= 0, lags, np.iinfo(np.int64).max)
        np.testing.assert_array_equal(span[:, 2], big.min(axis=1))
        assert (span[:, 3] == 0).all()
    else:
        assert tp.mspan is None


def _random_state(ops, P, D, B, T, seed):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    st = ops.ExchangeState(P, D, B, T)
    st.prev.copy_(torch.rand(st.prev.shape, generator=g, device="cuda"))
    return st


@pytest.mark.parametrize("jt", [4, 8, 16])
@pytest.mark.parametrize("per_wall", [True, False])
@pytest.mark.parametrize("name", list(ROOMS))
def test_tiled_gather_bitexact(ops, name, per_wall, jt):
    """Group sums on tiles equal the row-by-row gather's bit for bit, for
    histograms spanning several time blocks (T = 1,300 > 512)."""
    import torch
    r = _room(name, per_wall=per_wall)
    gp, _ = _grouped_and_generic(ops, r, 343., 1 / 2000)
    tp = ops.tile_plan(gp, jt, min_merge=0)
    P, D, B, T = gp.P, gp.D, gp.B, 1300
    st = _random_state(ops, P, D, B, T, seed=jt)
    st.gather(gp)
    ref = st.gsum[:gp.n_groups].clone()
    st.gsum.fill_(-1.0)
    gp.tiles = tp
    st.gather(gp)
    torch.cuda.synchronize()
    assert torch.equal(st.gsum[:gp.n_groups], ref)


@pytest.mark.parametrize("jt", [4, 8, 16])
def test_tiled_gather_bitexact_wide_delays(ops, jt):
    """At 48 kHz delays exceed 256 bins, so a row's first bins fall inside
    the workgroups' bin ranges, behind lanes whose bins precede the row
    (t < delay): the row-by-row sums match a CPU sum and the tiled ones
    match them bit for bit."""
    import torch
    r = _room("3x2.5x2")
    gp, _ = _grouped_and_generic(ops, r, 343., 1 / 48000)
    tp = ops.tile_plan(gp, jt, min_merge=0)
    P, D, B, T = gp.P, gp.D, gp.B, 1500
    st = _random_state(ops, P, D, B, T, seed=11)
    st.gather(gp)
    ref = st.gsum[:gp.n_groups].clone()
    # delays past 256 bins put a row's first bins at the ends of the
    # workgroups' bin ranges: check the row-by-row sums against a CPU sum
    cpu = _gsum_cpu(gp, _np(st.prev), T)
    np.testing.assert_allclose(_np(ref), cpu, rtol=1e-6, atol=1e-30)
    st.gsum.fill_(-1.0)
    gp.tiles = tp
    st.gather(gp)
    torch.cuda.synchronize()
    assert torch.equal(st.gsum[:gp.n_groups], ref)


@pytest.mark.parametrize("jt,block,R,pad", [
    (16, (4, 4), 16, True), (16, (4, 4), 8, True), (16, None, 8, True),
    (16, (4, 4), 16, False), (8, (2, 4), 16, True), (4, (2, 2), 16, True)])
def test_block_tiles_bitexact(ops, jt, block, R, pad):
    """Tiles as blocks of neighbouring patches (rows permuted, walls'
    edges padded with empty slots): the same group sums, bit for bit, in
    the plan's own group order (each slot stores at its destination's
    groups)."""
    import torch
    r = _room("3x2.5x2")
    gp, _ = _grouped_and_generic(ops, r, 343., 1 / 48000)
    order = None if block is None else ops.block_order(
        r.patches_points, r.patch_to_wall_ids, block, pad=pad)
    if block is not None:
        assert (order < 0).any() == pad   # padded blocks have empty slots
    tp = ops.tile_plan(gp, jt, min_merge=0, order=order)
    tp.R = R
    n = gp.P if order is None else len(order)
    assert tp.P == n + (-n) % jt
    P, D, B, T = gp.P, gp.D, gp.B, 1500
    st = _random_state(ops, P, D, B, T, seed=5)
    st.gather(gp)
    ref = st.gsum[:gp.n_groups].clone()
    st.gsum.fill_(-1.0)
    gp.tiles = tp
    st.gather(gp)
    torch.cuda.synchronize()
    assert torch.equal(st.gsum[:gp.n_groups], ref)


@pytest.mark.parametrize("block", [None, (4, 4)])
@pytest.mark.parametrize("per_wall", [True, False])
def test_ds_tiles_bitexact(ops, per_wall, block):
    """Diffuse + specular plan on 16-row tiles (merged by each entry's
    specular direction, one wave per row, LDS-staged): the specular group
    sums and the rows' diffuse sums equal ds_gather_kernel's bit for bit,
    at 48 kHz (delays of hundreds of bins) and with a bin count that is
    not a multiple of the block."""
    import torch
    from test_gpu_parity import _ds_plan
    r = _room("3x2.5x2", per_wall=per_wall)
    gp = _ds_plan(ops, r, 343., 1 / 48000)
    order = None if block is None else ops.block_order(
        r.patches_points, r.patch_to_wall_ids, block)
    tp = ops.tile_plan(gp, 16, min_merge=0, order=order)
    P, D, B, T = gp.P, gp.D, gp.B, 1500
    st = _random_state(ops, P, D, B, T, seed=7)
    st.gather(gp)
    ref = st.gsum[:gp.n_groups].clone()
    st.gsum.fill_(-1.0)
    gp.tiles = tp
    st.gather(gp)
    torch.cuda.synchronize()
    assert torch.equal(st.gsum[:gp.n_groups], ref)


def _gsum_cpu(gp, prev, T):
    """gsum[g, b, t] = sum_{e in g} scal[b, e] prev[rowid[e], b, t - lag[e]]
    in float64."""
    P, D, B = gp.P, gp.D, gp.B
    rows = prev.reshape(P * D, B, T).astype(np.float64)
    gs, rowid = _np(gp.grp_start), _np(gp.rowid)
    lag, scal = _np(gp.lag) & LAG_MASK, _np(gp.scal)
    out = np.zeros((gp.n_groups, B, T))
    g_of_e = np.searchsorted(gs, np.arange(gp.Vd), side="right") - 1
    for e in range(gp.Vd):
        n = int(lag[e])
        if n < T:
            out[g_of_e[e], :, n:] += scal[:, e, None] * rows[rowid[e], :, :T - n]
    return out


@pytest.mark.parametrize("jt", [4, 8, 16])
def test_tiled_exchange_matches_oracle(ops, jt):
    """Whole exchange on tiles: equal to the untiled grouped plan and
    within the exchange tolerance of the oracle's FP64 loop."""
    import torch
    r = _room("2.5x1.5x1", seed=5)
    r.init_source_energy(np.array([1.1, .7, .4]))
    T, c, dt, K = 400, 343., 1 / 2000, 8
    gp, _ = _grouped_and_generic(ops, r, c, dt)
    P, D, B = r.energy_init_source.shape
    outs = []
    for tiles in (None, ops.tile_plan(gp, jt, min_merge=0)):
        gp.tiles = tiles
        st = ops.ExchangeState(P, D, B, T)
        st.init(r.energy_init_source, r.distance_patches_to_source, c, dt)
        for _ in range(K):
            st.step(gp)
        st.check()
        outs.append(st.finalize())
    assert torch.equal(outs[0], outs[1])
    ref = orc.exchange(T, r.energy_init_source, r.distance_patches_to_source,
                       r.pair_distances(), r.fft, r.p2o, c, dt, K,
                       r.visible_patches, nthreads=8)
    assert rel_err_within_db(_np(outs[1]), ref) <= EXCHANGE_RTOL


def test_class_path_uses_tiles(ops):
    """DirectionalRadiosityFast builds tiles for its grouped plan when they
    merge entries, and its histograms do not depend on them."""
    import torch
    from sparrowpy_amd import DirectionalRadiosityFast
    from sparrowpy_amd import brdf as B
    from sparrowpy_amd import samplings
    from sparrowpy_amd.coordinates import Coordinates, FrequencyData
    freqs = np.array([500., 1000.])
    dirs = samplings.hemisphere(samplings.sph_gaussian(sh_order=3))
    br = B.create_from_scattering(
        dirs, dirs, FrequencyData(np.array([.3, .6]), freqs),
        FrequencyData(np.array([.1, .2]), freqs))
    hists = []
    for rows in (0, 4):
        r = DirectionalRadiosityFast.from_polygon(
            testing.shoebox_room_stub(3, 2.5, 2), .5)
        r.set_wall_brdf(np.arange(6), br, dirs, dirs)
        r.set_air_attenuation(FrequencyData(np.array([1e-3, 2e-3]), freqs))
        r.TILE_ROWS, r.TILE_MIN_MERGE = rows, 0.
        r.bake_geometry()
        r.init_source_energy(Coordinates(1.2, .9, .7))
        r.calculate_energy_exchange(343., 1 / 2000, .2, 6)
        plan = r._exchange_plan(343., 1 / 2000)
        assert isinstance(plan, ops.GroupedPlan)
        assert (plan.tiles is not None) == bool(rows)
        hists.append(r.collect_energy_receiver_mono(
            Coordinates(1.5, 1.2, 1.)).time)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(hists[0], hists[1])


@pytest.mark.parametrize("kind", ["tiled", "grouped", "ds"])
def test_exchange_long_delays_matches_oracle(ops, kind):
    """Whole exchange at 48 kHz (delays of up to ~600 bins, past the
    kernels' 256-bin lane strides) on the tiled, row-by-row grouped and
    diffuse + specular plans, against the oracle's FP64 loop."""
    from test_gpu_parity import _ds_plan
    r = _room("3x2.5x2", seed=6)
    r.init_source_energy(np.array([1.3, .8, .9]))
    T, c, dt, K = 1500, 343., 1 / 48000, 4
    if kind == "ds":
        plan = _ds_plan(ops, r, c, dt)
    else:
        plan, _ = _grouped_and_generic(ops, r, c, dt)
        if kind == "tiled":
            plan.tiles = ops.tile_plan(plan, 4, min_merge=0)
    P, D, B = r.energy_init_source.shape
    st = ops.ExchangeState(P, D, B, T)
    st.init(r.energy_init_source, r.distance_patches_to_source, c, dt)
    for _ in range(K):
        st.step(plan)
    st.check()
    ref = orc.exchange(T, r.energy_init_source, r.distance_patches_to_source,
                       r.pair_distances(), r.fft, r.p2o, c, dt, K,
                       r.visible_patches, nthreads=8)
    assert rel_err_within_db(_np(st.finalize()), ref) <= EXCHANGE_RTOL


@pytest.mark.parametrize("K", [1, 2, 3, 6, 7])
def test_totals_added_in_pairs(ops, K):
    """Grouped steps add orders to the running total two at a time (an odd
    last order is added by finalize): totals match the generic plan's,
    which adds every order, for even and odd order counts."""
    import torch
    r = _room("2.5x1.5x1", seed=7)
    r.init_source_energy(np.array([.9, .6, .5]))
    T, c, dt = 300, 343., 1 / 2000
    gp, gen = _grouped_and_generic(ops, r, c, dt)
    gp.tiles = ops.tile_plan(gp, 4, min_merge=0)
    P, D, B = r.energy_init_source.shape
    outs = []
    for pl in (gen, gp):
        st = ops.ExchangeState(P, D, B, T)
        st.init(r.energy_init_source, r.distance_patches_to_source, c, dt)
        for k in range(K):
            st.step(pl)
            assert st.pending == (isinstance(pl, ops.GroupedPlan) and
                                  k % 2 == 0)
        outs.append(st.finalize())
        assert not st.pending
        # a second finalize sees the same total
        assert torch.equal(st.finalize(), outs[-1])
    a, b = _np(outs[0]).astype(np.float64), _np(outs[1]).astype(np.float64)
    assert np.abs(a - b).max() <= 1e-6 * np.abs(a).max()
    assert rel_err_within_db(b, a) <= 1e-6
