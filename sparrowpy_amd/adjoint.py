"""Gradients of receiver histograms with respect to the BRDF tables.

The reference computes no gradients; this module differentiates its
forward path -- injection (``RadiosityFast.py:988-1034``), the energy
exchange (``:1073-1145``) and the mono receiver collection (``:570-602``,
``:1148-1185``).  Two forms:

* BRDF rows that are a diffuse constant plus one specular entry (what
  ``create_from_scattering``, ``brdf.py:8-130``, produces): such a row is
  ``A[k, b] * 1 + C[k, b] * e_m(k)``; the forward runs on the diffuse +
  specular exchange plan with weights rebuilt from ``(A, C)``, and the
  backward is the hand-written adjoint (``spr_exchange_adjoint_ds``).
  :class:`ExchangeAdjoint` holds one scene/source/receiver set;
  :func:`receiver_histograms` is its ``torch.autograd.Function`` entry and
  :func:`scattering_tables` builds ``(A, C)`` from absorption and
  scattering coefficients differentiably.
* BRDF rows of any form (``create_from_directional_scattering``,
  ``brdf.py:133-224``, measured tables): :class:`BRDFExchangeAdjoint` on
  the grouped plan, gradients with respect to the whole stored table;
  :func:`receiver_histograms_brdf` is its entry and
  :func:`directional_scattering_tables` builds the table from absorption
  and directional scattering coefficients differentiably.

Both run the transposed exchange order by order from the last, with the
forward order histograms recomputed from checkpoints kept every
``checkpoint_every`` orders.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _hip, ops
from ._hip import call, ptr, stream_ptr

__all__ = ["ExchangeAdjoint", "receiver_histograms", "scattering_tables",
           "specular_factors", "log_energy_loss", "BRDFExchangeAdjoint",
           "receiver_histograms_brdf", "directional_scattering_tables"]


def _i32(n, dev):
    return torch.empty(max(int(n), 1), dtype=torch.int32, device=dev)


def _source_and_receivers(adj, r, source, receivers):
    """Set on ``adj`` (with P, D, B set) what both adjoints share: the
    source position, its energy per patch before the BRDF (``energy0``,
    ``d0``; reference ``RadiosityFast.py:436-522``), the geometry, the BRDF
    row maps, and the receivers' collection terms."""
    b, g = r._bake, r._geometry()
    pos = np.asarray(source.cartesian if hasattr(source, "cartesian")
                     else source, dtype=float).reshape(3)
    adj.source = ops.as_dev(pos, torch.float64)
    vis = ops.check_point2patch_visibility(
        pos, g["centers"], g["walls_normal"], g["walls_points"])[0]
    e0, _, d0 = ops.source_energy(
        pos, g["centers"], g["points"], vis, b["air"], adj.B,
        g["wall_ids"], b["sources"], b["brdf"], b["brdf_index"], adj.D)
    adj.energy0, adj.d0 = e0, d0
    adj._geom = g
    adj.brdf_index = ops.as_dev(b["brdf_index"], torch.int64)
    adj.in_dirs = ops.as_dev(b["sources"], torch.float64)
    pts, _, adj.rfactor, adj.ridx = r._receiver_terms(receivers)
    adj.rpts = ops.as_dev(pts, torch.float64)
    adj.air = ops.as_dev(b["air"] if b["air"] is not None
                         else np.zeros(adj.B), torch.float64)
    adj.R = adj.rpts.shape[0]


class ExchangeAdjoint:
    """Forward and adjoint of ``tables -> receiver histograms`` for one
    baked scene, one source and a set of receivers.

    ``radiosity``: a baked :class:`~sparrowpy_amd.radiosity.
    DirectionalRadiosityFast` with BRDFs (of the diffuse + specular form)
    and air attenuation set.  The histogram settings are those of
    ``calculate_energy_exchange``.
    """

    def __init__(self, radiosity, source, receivers, speed_of_sound,
                 time_resolution, duration, max_order, checkpoint_every=5):
        r = radiosity
        if r._bake is None:
            raise RuntimeError("bake_geometry() must run first")
        b, g = r._bake, r._geometry()
        if b["brdf"] is None:
            raise ValueError("the adjoint needs BRDF tables")
        self.c, self.dt = float(speed_of_sound), float(time_resolution)
        self.T = int(duration / time_resolution)
        self.K = int(max_order)
        self.every = max(1, int(checkpoint_every))
        self.last = None   # ForwardRecord of the last forward()
        plan = ops.diffuse_specular_plan(
            b["vis"], g["centers"], g["areas"], b["ff"], b["p2o"], b["air"],
            g["wall_ids"], b["sources"], b["brdf"], b["brdf_index"], b["D"],
            b["n_bins"], self.c, self.dt, keep_adjoint=True)
        if plan is None:
            raise ValueError("BRDF rows are not diffuse + specular")
        self.plan, adj = plan, plan.adjoint
        self.P, self.D, self.B = plan.P, plan.D, plan.B
        self.NK = int(np.shape(b["brdf"])[0]) * adj["S"]
        dev = g["centers"].device
        self.dev = dev
        tables = ops.diffuse_specular_tables(b["brdf"], self.D, self.B,
                                             self.NK)
        self.A0, self.C0, self.spec_dir = tables
        self._transpose()
        _source_and_receivers(self, r, source, receivers)
        self.k0 = _i32(self.P, dev)
        self.e0dir = torch.empty((self.P, self.D, self.B), dtype=torch.float64,
                                 device=dev)
        self.S = adj["S"]

    # -- plan views ----------------------------------------------------------
    def _transpose(self):
        p, adj, dev = self.plan, self.plan.adjoint, self.dev
        P, D, Vd = self.P, self.D, self.plan.Vd
        self.trow_ptr = torch.empty(P + 1, dtype=torch.int64, device=dev)
        call("spr_plan_ds_transpose_rows", P, D, ptr(p.row_ptr),
             ptr(p.rowid), ptr(self.trow_ptr), stream_ptr())
        cnt = self.trow_ptr[1:] - self.trow_ptr[:-1]
        max_trow = int(cnt.max().item()) if P else 0
        cursor = torch.empty(P, dtype=torch.int64, device=dev)
        self.te, self.tkey = _i32(Vd, dev), _i32(Vd, dev)
        self.tj, self.tlag, self.tm = _i32(Vd, dev), _i32(Vd, dev), \
            _i32(Vd, dev)
        self.tgptr = torch.empty(P * D + 1, dtype=torch.int64, device=dev)
        call("spr_plan_ds_transpose_fill", P, D, Vd, max_trow,
             ptr(p.row_ptr), ptr(p.rowid), ptr(p.lag), ptr(adj["key"]),
             ptr(self.trow_ptr), ptr(cursor), ptr(self.te), ptr(self.tkey),
             ptr(self.tj), ptr(self.tlag), ptr(self.tm), ptr(self.tgptr),
             stream_ptr())

    def _set_tables(self, A, C):
        p, adj = self.plan, self.plan.adjoint
        A = ops.as_dev(A, torch.float64)
        C = ops.as_dev(C, torch.float64)
        if A.shape != (self.NK, self.B) or C.shape != (self.NK, self.B):
            raise ValueError(f"tables must be ({self.NK}, {self.B})")
        call("spr_ds_weights", self.B, p.Vd, ptr(adj["krow"]),
             ptr(adj["wgt"]), ptr(A), ptr(C), ptr(p.scal), ptr(p.dscal),
             stream_ptr())
        g = self._geom
        call("spr_ds_injection", self.P, self.D, self.B, ptr(self.source),
             ptr(g["centers"]), ptr(g["wall_ids"]), self.S,
             ptr(self.brdf_index), ptr(self.in_dirs), ptr(self.energy0),
             ptr(A), ptr(C), ptr(self.spec_dir), ptr(self.k0),
             ptr(self.e0dir), stream_ptr())

    # -- forward ----------------------------------------------------------------
    def _order(self, state_prev, out, gsum):
        """One order without the running total: out = A state_prev."""
        p = self.plan
        call("spr_exchange_gather_ds", self.P, self.B, self.T, p.Vd,
             ptr(p.row_ptr), ptr(p.grp_ptr), ptr(p.rowid), ptr(p.lag),
             ptr(p.scal), ptr(p.dscal), ptr(state_prev), ptr(gsum),
             stream_ptr())
        call("spr_exchange_spread_ds", self.P, self.D, self.B, self.T,
             ptr(p.grp_ptr), ptr(p.grp_key), ptr(gsum), None, ptr(out),
             None, None, stream_ptr())

    def forward(self, A, C):
        """Receiver histograms (float32 [R, B, T]) for tables (A, C).

        The call's order checkpoints and tables are kept as ``self.last``
        (a :class:`ForwardRecord`) for :meth:`backward`; a later forward
        replaces it, so the autograd entry takes each call's record into
        its own context instead."""
        A = ops.as_dev(A, torch.float64).detach().clone()
        C = ops.as_dev(C, torch.float64).detach().clone()
        self._set_tables(A, C)
        st = ops.ExchangeState(self.P, self.D, self.B, self.T, self.dev)
        st.init(self.e0dir, self.d0, self.c, self.dt)
        checkpoints = {0: st.prev.clone()}
        for k in range(self.K):
            st.step(self.plan)
            if (k + 1) % self.every == 0 and k + 1 < self.K:
                checkpoints[k + 1] = st.prev.clone()
        st.check()
        self.last = ForwardRecord(A, C, checkpoints)
        total = st.finalize(out=st.hi)
        self._gsum = st.gsum if st.gsum is not None else torch.empty(
            (max(self.plan.n_groups, 1), self.B, self.T), dtype=torch.float32,
            device=self.dev)
        H = ops.collect(total, self.rpts, self._geom["centers"], self.ridx,
                        self.rfactor, self.air, self.c, self.dt)
        del st, total
        return H

    # -- backward ---------------------------------------------------------------
    def backward(self, grad_hist, record=None):
        """(dL/dA, dL/dC) (float64 [NK, B]) for dL/dH = ``grad_hist``
        ([R, B, T]) at the tables of ``record`` (default: the last
        :meth:`forward`).  The record's tables are set again first, so
        the recomputed orders and the transposed weights are those of the
        forward the record comes from."""
        record = self.last if record is None else record
        if record is None or not record.checkpoints:
            raise RuntimeError("ExchangeAdjoint.backward: no forward record "
                               "(forward() was not called, or its record "
                               "was released)")
        self._set_tables(record.A, record.C)
        checkpoints = record.checkpoints
        P, D, B, T, Vd = self.P, self.D, self.B, self.T, self.plan.Vd
        dev, p, adj = self.dev, self.plan, self.plan.adjoint
        gH = ops.as_dev(grad_hist, torch.float32).contiguous()
        shape = (P, D, B, T)
        lam = torch.empty(shape, dtype=torch.float32, device=dev)
        call("spr_collect_adjoint", self.R, P, D, B, T, ptr(gH),
             ptr(self.rpts), ptr(self._geom["centers"]), ptr(self.ridx),
             ptr(self.rfactor), ptr(self.air), self.c, self.dt, ptr(lam),
             stream_ptr())
        twa = torch.empty((B, max(Vd, 1)), dtype=torch.float64, device=dev)
        twc = torch.empty_like(twa)
        call("spr_ds_transposed_weights", B, Vd, ptr(self.te), ptr(p.scal),
             ptr(p.dscal), ptr(twa), ptr(twc), stream_ptr())
        ga = torch.zeros((2, B, max(Vd, 1)), dtype=torch.float64, device=dev)
        musum = torch.empty((P, B, T), dtype=torch.float32, device=dev)
        mu = lam.clone()                     # mu_K
        mu_out = torch.empty_like(lam)
        starts = sorted(checkpoints)
        if starts != list(range(0, max(self.K, 1), self.every)):
            raise RuntimeError("ExchangeAdjoint.backward: the record's "
                               "checkpoints do not cover the orders")
        for s in reversed(starts):
            end = min(s + self.every, self.K)
            states = [checkpoints[s]]   # E_s .. E_{end-1}
            for k in range(s, end - 1):
                nxt = torch.empty_like(lam)
                self._order(states[-1], nxt, self._gsum)
                states.append(nxt)
            for k in range(end - 1, s - 1, -1):
                call("spr_direction_sum", P, D, B, T, ptr(mu), ptr(musum),
                     stream_ptr())
                call("spr_exchange_adjoint_ds", P, D, B, T, Vd,
                     ptr(self.tgptr), ptr(self.tkey), ptr(self.tlag),
                     ptr(self.tm), ptr(twa), ptr(twc), ptr(lam), ptr(musum),
                     ptr(mu), ptr(states[k - s]), ptr(mu_out), ptr(ga),
                     stream_ptr())
                mu, mu_out = mu_out, mu
            del states
        gA = torch.zeros((self.NK, B), dtype=torch.float64, device=dev)
        gC = torch.zeros_like(gA)
        call("spr_ds_injection_grad", P, D, B, T, ptr(self.energy0),
             ptr(self.d0), self.c, self.dt, ptr(self.k0), ptr(self.spec_dir),
             ptr(mu), ptr(gA), ptr(gC), stream_ptr())
        call("spr_ds_table_grad", B, Vd, ptr(self.te), ptr(adj["krow"]),
             ptr(adj["wgt"]), ptr(ga), ptr(gA), ptr(gC), stream_ptr())
        return gA, gC

    def release(self):
        """Drop the record of the last forward."""
        self.last = None


class ForwardRecord:
    """What the adjoint of one forward call needs: its tables (A, C) and
    the order histograms kept every ``checkpoint_every`` orders."""

    def __init__(self, A, C, checkpoints):
        self.A, self.C, self.checkpoints = A, C, checkpoints


class _Histograms(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, C, adj):
        ctx.adj = adj
        H = adj.forward(A.detach(), C.detach()).to(torch.float64)
        # each call keeps its own record (freed with the graph), so several
        # calls in one graph, and backward with retain_graph, stay correct
        ctx.record, adj.last = adj.last, None
        return H

    @staticmethod
    def backward(ctx, grad):
        gA, gC = ctx.adj.backward(grad, ctx.record)
        return gA, gC, None


def receiver_histograms(adj, A, C):
    """Receiver histograms [R, B, T] (float64 tensor) as a differentiable
    function of the tables ``A``, ``C`` ([NK, B] float64 device tensors)."""
    return _Histograms.apply(A, C, adj)


class BRDFExchangeAdjoint:
    """Forward and adjoint of ``BRDF table -> receiver histograms`` for BRDF
    rows of any form (``create_from_directional_scattering`` among them,
    reference ``brdf.py:133-224``), on the grouped plan.

    The table is ``table[k, d, b]`` with row ``k = brdf * S + s`` as
    ``set_wall_brdf`` stores it (``brdf.freq * pi``; :attr:`table0` is the
    radiosity object's own); one order is ``gsum = gather(E_k)`` per group
    of sources sharing a BRDF row, ``E_{k+1}[j, d] = sum_g table[key_g, d]
    gsum[g]``, and injection puts ``energy0[p, b] table[k0_p, :, b]`` at
    each patch's delay bin (``RadiosityFast.py:988-1145``).  The backward
    runs the transposed orders from the last (``spr_grouped_spread_adjoint``,
    ``spr_grouped_gather_adjoint``), the forward orders recomputed from
    checkpoints kept every ``checkpoint_every`` orders, and sums the table
    gradient per group and order (``spr_grouped_table_grad``), then per
    BRDF row.  Arguments as for :class:`ExchangeAdjoint`.
    """

    def __init__(self, radiosity, source, receivers, speed_of_sound,
                 time_resolution, duration, max_order, checkpoint_every=5):
        r = radiosity
        if r._bake is None:
            raise RuntimeError("bake_geometry() must run first")
        b, g = r._bake, r._geometry()
        if b["brdf"] is None:
            raise ValueError("the adjoint needs BRDF tables")
        self.c, self.dt = float(speed_of_sound), float(time_resolution)
        self.T = int(duration / time_resolution)
        self.K = int(max_order)
        self.every = max(1, int(checkpoint_every))
        self.last = None   # TableRecord of the last forward()
        self.plan = ops.GroupedPlan.from_geometry(
            b["vis"], g["centers"], g["areas"], b["ff"], b["p2o"], b["air"],
            g["wall_ids"], b["sources"], b["brdf"], b["brdf_index"], b["D"],
            b["n_bins"], self.c, self.dt)
        p = self.plan
        if r.TILE_ROWS:
            # the class's destination tiles: the same group sums, faster
            p.tiles = ops.tile_plan(p, r.TILE_ROWS, r.TILE_MIN_MERGE,
                                    r._tile_order())
            if p.tiles is not None:
                p.tiles.R = r.TILE_R
        self.P, self.D, self.B = p.P, p.D, p.B
        self.dev = dev = g["centers"].device
        brdf = ops.as_dev(b["brdf"], torch.float64)
        self.S = int(brdf.shape[1])
        self.NK = int(brdf.shape[0]) * self.S
        self.table0 = brdf.reshape(self.NK, self.D, self.B).clone()
        self._transpose()
        _source_and_receivers(self, r, source, receivers)
        # each patch's BRDF row for the direction of the source
        # (spr_ds_injection with zero tables fills k0 only)
        self.k0 = _i32(self.P, dev)
        zero = torch.zeros((self.NK, self.B), dtype=torch.float64, device=dev)
        spec = torch.zeros(self.NK, dtype=torch.int32, device=dev)
        scratch = torch.empty((self.P, self.D, self.B), dtype=torch.float64,
                              device=dev)
        call("spr_ds_injection", self.P, self.D, self.B, ptr(self.source),
             ptr(g["centers"]), ptr(g["wall_ids"]), self.S,
             ptr(self.brdf_index), ptr(self.in_dirs), ptr(self.energy0),
             ptr(zero), ptr(zero), ptr(spec), ptr(self.k0), ptr(scratch),
             stream_ptr())
        self._k0 = self.k0.to(torch.int64)

    def _transpose(self):
        """Entries by source row r = i * D + d, in entry order: trow_ptr
        [P*D+1], each entry's group, delay (flag cleared) and weights."""
        p, dev = self.plan, self.dev
        Vd, G, PD = p.Vd, p.n_groups, self.P * self.D
        rowid = p.rowid.to(torch.int64)
        order = torch.sort(rowid, stable=True).indices
        self.trow_ptr = torch.zeros(PD + 1, dtype=torch.int64, device=dev)
        if Vd:
            self.trow_ptr[1:] = torch.cumsum(
                torch.bincount(rowid, minlength=PD), 0)
        gid = torch.repeat_interleave(
            torch.arange(G, dtype=torch.int64, device=dev),
            p.grp_start[1:G + 1] - p.grp_start[:G])
        self.tgrp = _i32(Vd, dev)
        self.tlag = _i32(Vd, dev)
        self.tscal = torch.empty((self.B, max(Vd, 1)), dtype=torch.float64,
                                 device=dev)
        if Vd:
            self.tgrp[:Vd] = gid[order].to(torch.int32)
            self.tlag[:Vd] = p.lag[order] & 0x3fffffff
            self.tscal[:, :Vd] = p.scal[:, order]

    def _set_table(self, table):
        t = ops.as_dev(table, torch.float64)
        if tuple(t.shape) != (self.NK, self.D, self.B):
            raise ValueError(f"table must be ({self.NK}, {self.D}, {self.B})")
        self.plan.rtab[:, :, :self.D] = t.permute(0, 2, 1)
        self.e0dir = (self.energy0[:, None, :] * t[self._k0]).contiguous()

    def _gather(self, state, gsum):
        ops.gather_groups(self.plan, (self.P, self.D, self.B, self.T), state,
                          gsum)

    def _order(self, state_prev, out, gsum):
        """One order without the running total: out = A state_prev."""
        p = self.plan
        self._gather(state_prev, gsum)
        call("spr_exchange_spread_grouped", self.P, self.D, self.B, self.T,
             ptr(p.grp_ptr), ptr(p.grp_key), ptr(p.rtab), ptr(gsum),
             ptr(out), None, None, stream_ptr())

    def forward(self, table):
        """Receiver histograms (float32 [R, B, T]) for ``table``; the call's
        checkpoints and table are kept as ``self.last`` (see
        :meth:`ExchangeAdjoint.forward`)."""
        table = ops.as_dev(table, torch.float64).detach().clone()
        self._set_table(table)
        st = ops.ExchangeState(self.P, self.D, self.B, self.T, self.dev)
        st.init(self.e0dir, self.d0, self.c, self.dt)
        checkpoints = {0: st.prev.clone()}
        for k in range(self.K):
            st.step(self.plan)
            if (k + 1) % self.every == 0 and k + 1 < self.K:
                checkpoints[k + 1] = st.prev.clone()
        st.check()
        self.last = TableRecord(table, checkpoints)
        total = st.finalize(out=st.hi)
        H = ops.collect(total, self.rpts, self._geom["centers"], self.ridx,
                        self.rfactor, self.air, self.c, self.dt)
        del st, total
        return H

    def backward(self, grad_hist, record=None):
        """dL/dtable (float64 [NK, D, B]) for dL/dH = ``grad_hist`` ([R, B,
        T]) at the table of ``record`` (default: the last :meth:`forward`).
        """
        record = self.last if record is None else record
        if record is None or not record.checkpoints:
            raise RuntimeError("BRDFExchangeAdjoint.backward: no forward "
                               "record (forward() was not called, or its "
                               "record was released)")
        self._set_table(record.table)
        checkpoints = record.checkpoints
        starts = sorted(checkpoints)
        if starts != list(range(0, max(self.K, 1), self.every)):
            raise RuntimeError("BRDFExchangeAdjoint.backward: the record's "
                               "checkpoints do not cover the orders")
        P, D, B, T = self.P, self.D, self.B, self.T
        p, dev = self.plan, self.dev
        G = p.n_groups
        gH = ops.as_dev(grad_hist, torch.float32).contiguous()
        shape = (P, D, B, T)
        lam = torch.empty(shape, dtype=torch.float32, device=dev)
        call("spr_collect_adjoint", self.R, P, D, B, T, ptr(gH),
             ptr(self.rpts), ptr(self._geom["centers"]), ptr(self.ridx),
             ptr(self.rfactor), ptr(self.air), self.c, self.dt, ptr(lam),
             stream_ptr())
        DP = p.rtab.shape[2]
        gg = torch.zeros((max(G, 1), B, DP), dtype=torch.float64, device=dev)
        gsum = torch.empty((max(G, 1), B, T), dtype=torch.float32, device=dev)
        nu = torch.empty_like(gsum)
        mu = lam.clone()                     # mu_K
        mu_out = torch.empty_like(lam)
        for s in reversed(starts):
            end = min(s + self.every, self.K)
            states = [checkpoints[s]]        # E_s .. E_{end-1}
            # the group sums of orders s .. end-2 come out of recomputing
            # the segment; kept (when there is room) instead of gathered
            # again for the backward orders
            keep = _fits(gsum, end - s - 1)
            sums = {}
            for k in range(s, end - 1):
                nxt = torch.empty_like(lam)
                gk = torch.empty_like(gsum) if keep else gsum
                self._order(states[-1], nxt, gk)
                if keep:
                    sums[k] = gk
                states.append(nxt)
            for k in range(end - 1, s - 1, -1):
                # mu = dL/dE_{k+1}; the group sums of order k -> k+1
                gk = sums.pop(k, None)
                if gk is None:
                    gk = gsum
                    self._gather(states[k - s], gk)
                call("spr_grouped_table_grad", P, D, B, T, ptr(p.grp_ptr),
                     ptr(mu), ptr(gk), ptr(gg), stream_ptr())
                del gk
                call("spr_grouped_spread_adjoint", P, D, B, T,
                     ptr(p.grp_ptr), ptr(p.grp_key), ptr(p.rtab), ptr(mu),
                     ptr(nu), stream_ptr())
                call("spr_grouped_gather_adjoint", P, D, B, T, p.Vd,
                     ptr(self.trow_ptr), ptr(self.tgrp), ptr(self.tlag),
                     ptr(self.tscal), ptr(nu), ptr(lam), ptr(mu_out),
                     stream_ptr())
                mu, mu_out = mu_out, mu
            del states
        gt = torch.zeros((self.NK, D, B), dtype=torch.float64, device=dev)
        call("spr_grouped_key_grad", G, D, B, ptr(p.grp_key), ptr(gg),
             ptr(gt), stream_ptr())
        call("spr_injection_table_grad", P, D, B, T, ptr(self.energy0),
             ptr(self.d0), self.c, self.dt, ptr(self.k0), ptr(mu), ptr(gt),
             stream_ptr())
        return gt

    def release(self):
        """Drop the record of the last forward."""
        self.last = None


def _fits(like, n):
    """Whether ``n`` more tensors like ``like`` fit in half of the device
    memory still available (free plus the allocator's unused cache)."""
    if n <= 0:
        return False
    free, _ = torch.cuda.mem_get_info(like.device)
    idle = torch.cuda.memory_reserved(like.device) - \
        torch.cuda.memory_allocated(like.device)
    return n * like.numel() * like.element_size() < 0.5 * (free + idle)


class TableRecord:
    """What the adjoint of one BRDFExchangeAdjoint forward needs: its table
    and the order histograms kept every ``checkpoint_every`` orders."""

    def __init__(self, table, checkpoints):
        self.table, self.checkpoints = table, checkpoints


class _TableHistograms(torch.autograd.Function):
    @staticmethod
    def forward(ctx, table, adj):
        ctx.adj = adj
        H = adj.forward(table.detach()).to(torch.float64)
        ctx.record, adj.last = adj.last, None
        return H

    @staticmethod
    def backward(ctx, grad):
        return ctx.adj.backward(grad, ctx.record), None


def receiver_histograms_brdf(adj, table):
    """Receiver histograms [R, B, T] (float64 tensor) as a differentiable
    function of the BRDF table ``table`` ([NK, D, B] float64 device tensor,
    rows as :attr:`BRDFExchangeAdjoint.table0`)."""
    return _TableHistograms.apply(table, adj)


def directional_scattering_tables(alpha, directional_scattering,
                                  receiver_directions):
    """BRDF tables of ``create_from_directional_scattering`` (reference
    ``brdf.py:133-224``) as the radiosity class stores them (times pi):
    ``table[n*S + s, d, b] = pi (1 - alpha[n, b]) sd[n, s, d, b] /
    (w_d cos(theta_d))``, with the outgoing weights normalised to 2 pi.
    Differentiable in ``alpha`` ([N, B]) and ``directional_scattering``
    ([N, S, D, B]); returns [N*S, D, B]."""
    from .brdf import _colatitude
    cart = np.asarray(receiver_directions.cartesian, float)
    w = np.array(receiver_directions.weights, dtype=float)
    w = w * 2 * np.pi / w.sum()
    scale = np.pi / (w * np.cos(_colatitude(cart)))         # [D]
    sd = directional_scattering
    f = torch.as_tensor(scale, dtype=sd.dtype, device=sd.device)
    t = (1.0 - alpha)[:, None, None, :] * sd * f[None, None, :, None]
    return t.reshape(-1, t.shape[2], t.shape[3])


def specular_factors(directions):
    """Per incoming direction s of a BRDF direction set, the specular value
    1 / (cos(theta_s) w_m) that ``create_from_scattering`` puts at the
    specular direction m(s) (``brdf.py:104-118``), and m(s)."""
    from .brdf import create_from_scattering
    from .coordinates import FrequencyData
    zero = create_from_scattering(
        directions, directions, FrequencyData(np.zeros((1, 1)), [1000.]),
        FrequencyData(np.zeros((1, 1)), [1000.])).freq[..., 0]
    m = np.argmax(zero, axis=1)
    return zero[np.arange(zero.shape[0]), m], m


def scattering_tables(alpha, scattering, spec_value):
    """Diffuse and specular tables of ``create_from_scattering`` BRDFs:
    for BRDF n with absorption alpha[n, b] and scattering s[n, b] and
    incoming direction s_in, row k = n * S + s_in holds, as the radiosity
    class stores BRDFs (times pi), A = (1 - alpha) s and
    C = (1 - alpha)(1 - s) pi spec_value[s_in].
    Differentiable in alpha and scattering ([NBRDF, B] tensors)."""
    sv = torch.as_tensor(np.asarray(spec_value), dtype=alpha.dtype,
                         device=alpha.device)
    keep = 1.0 - alpha                                  # [N, B]
    A = (keep * scattering)[:, None, :].expand(-1, sv.shape[0], -1)
    C = (keep * (1.0 - scattering))[:, None, :] * sv[None, :, None] * np.pi
    return A.reshape(-1, alpha.shape[1]), C.reshape(-1, alpha.shape[1])


def log_energy_loss(H, target, dynamic_range_db=60.0):
    """Sum of squared log10-energy differences over the bins of each
    target histogram within ``dynamic_range_db`` of its peak."""
    t = torch.as_tensor(target, dtype=H.dtype, device=H.device)
    peak = t.amax(dim=-1, keepdim=True)
    mask = (t >= peak * 10 ** (-dynamic_range_db / 10)) & (t > 0)
    Hm = torch.where(mask, H, torch.ones_like(H))
    tm = torch.where(mask, t, torch.ones_like(t))
    d = torch.log10(Hm) - torch.log10(tm)
    return (d * d * mask).sum()


def recover_materials(adj, target, alpha0, scattering0, spec_value,
                      steps=200, lr=0.02, bounds=(0.01, 0.99), log=None):
    """Fit per-BRDF, per-band absorption and scattering ([NBRDF, B]) to
    target receiver histograms: Adam on ``log_energy_loss`` through the
    adjoint exchange, the parameters clipped to ``bounds`` after each step.
    Returns (alpha, scattering, losses per step)."""
    kw = dict(dtype=torch.float64, device=adj.dev)
    a = torch.tensor(np.asarray(alpha0), requires_grad=True, **kw)
    s = torch.tensor(np.asarray(scattering0), requires_grad=True, **kw)
    opt = torch.optim.Adam([a, s], lr=lr)
    losses = []
    for step in range(steps):
        opt.zero_grad()
        A, C = scattering_tables(a, s, spec_value)
        loss = log_energy_loss(receiver_histograms(adj, A, C), target)
        loss.backward()
        opt.step()
        with torch.no_grad():
            a.clamp_(*bounds)
            s.clamp_(*bounds)
        losses.append(float(loss))
        if log is not None and (step % 20 == 0 or step == steps - 1):
            log(f"step {step}: loss {losses[-1]:.4e}")
    return a.detach(), s.detach(), losses

