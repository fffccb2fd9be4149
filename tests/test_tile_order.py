"""Destination orders for tiles (ops.block_order) and the grouped plan
with its rows permuted (ops._rows_in_order), on the CPU with torch CPU
tensors: host logic only, no kernels."""
import numpy as np
import pytest
import torch

from sparrowpy_amd import geometry, ops, scenes, testing


@pytest.mark.parametrize("block", [(4, 4), (2, 8), (1, 16)])
def test_block_order_covers_every_patch_in_compact_blocks(block):
    sc = scenes.street_canyon(seed=5, length=12., width=6., height=5.,
                              n_recesses=3)
    wp = np.array([w.pts for w in sc.walls])
    wn = np.array([w.normal for w in sc.walls])
    pts, _, P, ids = geometry.process_patches(wp, wn, 1.)
    order = ops.block_order(pts, ids, block)
    n = block[0] * block[1]
    assert len(order) % n == 0
    real = order[order >= 0]
    np.testing.assert_array_equal(np.sort(real), np.arange(P))
    ctr = geometry.calculate_center(pts)
    for t in range(len(order) // n):
        sel = order[t * n:(t + 1) * n]
        sel = sel[sel >= 0]
        assert len(set(ids[sel].tolist())) == 1      # one wall per block
        ext = ctr[sel].max(0) - ctr[sel].min(0)
        # the block spans at most block[0] x block[1] patches of 1 m
        assert sorted(ext)[-1] <= max(block) - 1 + 1e-9
        assert sorted(ext)[-2] <= min(block) - 1 + 1e-9


def _fake_plan(seed=0, P=9, B=2, D=4):
    """A GroupedPlan-shaped set of CPU tensors: random rows, each with
    groups of ascending keys."""
    rng = np.random.default_rng(seed)
    row_ptr, grp_ptr, grp_start, grp_key = [0], [0], [], []
    rowid, lag = [], []
    for j in range(P):
        n_g = int(rng.integers(0, 4))
        for g in range(n_g):
            grp_start.append(len(rowid))
            grp_key.append(10 * g + j % 3)
            for _ in range(int(rng.integers(1, 4))):
                rowid.append(int(rng.integers(0, P * D)))
                lag.append(int(rng.integers(0, 50)))
            lag[-1] |= ops.GroupedPlan.LAST_IN_GROUP
        row_ptr.append(len(rowid))
        grp_ptr.append(len(grp_key))
    Vd = len(rowid)
    grp_start.append(Vd)
    t = lambda a, d=torch.int64: torch.tensor(a, dtype=d)  # noqa: E731
    scal = torch.tensor(rng.uniform(size=(B, Vd)))
    return ops.GroupedPlan(P, D, B, t(row_ptr), t(grp_ptr), t(grp_start),
                           t(grp_key, torch.int32), t(rowid, torch.int32),
                           t(lag, torch.int32), scal, torch.zeros(1))


def test_rows_in_order_permutes_rows_and_keeps_groups():
    plan = _fake_plan()
    order = np.array([4, -1, 0, 8, 2, -1, 1, 3, 5, 7, 6, -1])
    q = ops._rows_in_order(plan, order)
    rp, gs, gk, rowid, lag, scal, gstart = (
        q[k] for k in ("row_ptr", "grp_start", "grp_key", "rowid", "lag",
                       "scal", "gstart"))
    o_rp, o_gp = plan.row_ptr.numpy(), plan.grp_ptr.numpy()
    o_gs = plan.grp_start.numpy()
    for r, j in enumerate(order):
        a, b = rp[r].item(), rp[r + 1].item()
        if j < 0:
            assert a == b and gstart[r] == 0
            continue
        e0, e1 = o_rp[j], o_rp[j + 1]
        np.testing.assert_array_equal(rowid[a:b], plan.rowid[e0:e1])
        np.testing.assert_array_equal(lag[a:b], plan.lag[e0:e1])
        np.testing.assert_array_equal(scal[:, a:b], plan.scal[:, e0:e1])
        assert gstart[r] == o_gp[j] and q["gend"][r] == o_gp[j + 1]
        # the row's groups, in order, with their keys and entry spans
        g_new = [g for g in range(len(gk)) if a <= gs[g] < b]
        g_old = list(range(o_gp[j], o_gp[j + 1]))
        assert len(g_new) == len(g_old)
        for gn, go in zip(g_new, g_old):
            assert gk[gn] == plan.grp_key[go]
            assert gs[gn] - a == o_gs[go] - e0
    assert gs[-1] == rowid.numel()
