"""Many scenes and sources in one exchange.

The reference runs one source of one scene per ``calculate_energy_exchange``
call (``RadiosityFast.py:524-568``; a multi-source study loops in Python,
``tests/test_multisource.py:82-132``).  Here every (scene, source) pair is
an *instance*: instances whose scenes share the number of directions and
bands and the plan kind form a class, and a class is run as one exchange
over the concatenation of its instances' patches -- a block-diagonal plan
(each instance's rows only reference its own patches), so one launch pair
per order covers the whole class and results are bit-for-bit those of the
per-instance runs (same entries, same per-row order).  The K orders of a
class are captured once into a HIP graph and replayed; per class the plan
kind (grouped or diffuse + specular) is picked by timing one order of each
candidate on the class's largest scene (``autotune``; only that scene's
plans are built for it, the others once, for the kind chosen); a grouped
class gathers on destination tiles (``ops.tile_plan``), whose shape --
rows per tile, patch block, bins per lane of the 16-row gather -- is timed
against the candidates of ``TILINGS`` on the class with the most work and
the fastest used for every class (``tune_tiles``).
Classes larger than the memory budget are split into chunks.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np
import torch

from . import ops
from .radiosity import DirectionalRadiosityFast

__all__ = ["BatchItem", "BatchedExchange", "run_per_instance"]


@dataclasses.dataclass
class BatchItem:
    """One baked scene (a DirectionalRadiosityFast with BRDFs and air set),
    its sources [n, 3] and receivers (Coordinates or [R, 3])."""

    radiosity: object
    sources: np.ndarray
    receivers: object


def _cat(ts, dim=0):
    return torch.cat([t.reshape(t.shape) for t in ts], dim=dim).contiguous()


def _concat_plans(plans, offsets):
    """Block-diagonal concatenation of per-instance plans (patch offsets)."""
    p0 = plans[0]
    D = p0.D
    dev = p0.row_ptr.device
    vd = [p.Vd for p in plans]
    voff = np.concatenate([[0], np.cumsum(vd)]).astype(np.int64)
    rows = [p.row_ptr[:-1] + int(voff[n]) for n, p in enumerate(plans)]
    rows.append(torch.tensor([voff[-1]], dtype=torch.int64, device=dev))
    row_ptr = torch.cat(rows)
    P = int(sum(p.P for p in plans))
    lag = _cat([p.lag for p in plans])
    if isinstance(p0, ops.GroupedPlan):
        gs = [p.n_groups for p in plans]
        goff = np.concatenate([[0], np.cumsum(gs)]).astype(np.int64)
        ks = [p.rtab.shape[0] for p in plans]
        koff = np.concatenate([[0], np.cumsum(ks)]).astype(np.int64)
        if p0.diffuse_specular:
            # the diffuse + specular table depends on (D, B) only: shared,
            # keys kept (the diffuse + specular spread reads them as
            # directions)
            koff[:] = 0
        grp = [p.grp_ptr[:-1] + int(goff[n]) for n, p in enumerate(plans)]
        grp.append(torch.tensor([goff[-1]], dtype=torch.int64, device=dev))
        starts = [p.grp_start[:-1] + int(voff[n]) for n, p in
                  enumerate(plans)]
        starts.append(torch.tensor([voff[-1]], dtype=torch.int64,
                                   device=dev))
        out = ops.GroupedPlan(
            P, D, p0.B, row_ptr, torch.cat(grp), torch.cat(starts),
            _cat([p.grp_key + int(koff[n]) for n, p in enumerate(plans)]),
            _cat([p.rowid + int(offsets[n]) * D for n, p in
                  enumerate(plans)]),
            lag, _cat([p.scal for p in plans], 1),
            p0.rtab if p0.diffuse_specular else _cat([p.rtab for p in plans]),
            _cat([p.dscal for p in plans], 1) if p0.diffuse_specular
            else None)
        if p0.diffuse_specular:
            # each entry's specular direction (D: none), the merge key of
            # the plan's 16-row tiles
            out.entry_key = _cat([p.entry_key for p in plans])
        return out
    return ops.ExchangePlan(
        P, D, p0.B, row_ptr,
        _cat([p.src + int(offsets[n]) for n, p in enumerate(plans)]),
        lag, _cat([p.dir for p in plans]), _cat([p.coef for p in plans]))


def _kind(plan):
    if isinstance(plan, ops.GroupedPlan):
        return "ds" if plan.diffuse_specular else "grouped"
    return "generic"


class BatchedExchange:
    """Exchange and receiver collection of many instances at once.

    ``items``: list of :class:`BatchItem`; all scenes share the histogram
    settings (speed of sound, bin width, duration, order).
    """

    # Destination tilings the autotuner tries per plan kind: (rows per
    # tile, block of patches, bins per lane of the 16-row gather).
    TILINGS = {"grouped": [(16, "auto", 8), (16, "auto", 16),
                           (4, "auto", 8), (0, None, 8)],
               "ds": [(16, "auto", 8), (0, None, 8)]}

    def __init__(self, items, speed_of_sound, time_resolution, duration,
                 max_order, autotune=True, memory_budget=None, graphs=True,
                 tile_rows=16, tile_block="auto", tile_r=8, tune_tiles=True):
        self.items = list(items)
        # tiles as blocks of each room's patches: (a, b) with a * b =
        # tile_rows, None (consecutive patch ids) or "auto" (the block
        # DirectionalRadiosityFast uses for tile_rows)
        self.tiling = self._tiling(tile_rows, tile_block, tile_r)
        # with autotune: the tiling is timed against TILINGS on the class
        # with the most work and the fastest used for every class
        self.tune_tiles = bool(tune_tiles)
        self.tiling_times = {}
        self.c, self.dt = float(speed_of_sound), float(time_resolution)
        self.T = int(duration / time_resolution)
        self.K = int(max_order)
        self.graphs = graphs
        dev = ops.device()
        if memory_budget is None:
            memory_budget = torch.cuda.get_device_properties(
                dev).total_memory // 2
        self.budget = int(memory_budget)
        self.autotune = autotune
        # one set of state buffers for all chunks (allocating each chunk's
        # anew cost more than the batched exchange saved)
        self.pool = ops.StatePool(dev)
        self.choice = {}          # class key -> plan kind
        self._receivers = {}
        for x, it in enumerate(self.items):
            r = it.radiosity
            pts, _, factor, idx = r._receiver_terms(it.receivers)
            g = r._geometry()
            self._receivers[x] = (ops.as_dev(pts, torch.float64), factor,
                                  idx, g["centers"],
                                  ops.as_dev(r._air_attenuation,
                                             torch.float64))

    @staticmethod
    def _tiling(rows, block, r):
        """(rows, block, R) with the block resolved and checked."""
        rows = int(rows)
        if isinstance(block, str):
            if block != "auto":
                raise ValueError('tile_block must be None, "auto" or (a, b)')
            block = DirectionalRadiosityFast._AUTO_BLOCKS.get(rows)
        if block is not None:
            block = (int(block[0]), int(block[1]))
            if block[0] * block[1] != rows:
                raise ValueError("tile_block must hold tile_rows patches")
        return rows, block, int(r)

    @property
    def tile_rows(self):
        return self.tiling[0]

    # -- planning ------------------------------------------------------------
    def _plan(self, x, kind):
        """Scene x's plan of ``kind``, without destination tiles: the
        batch tiles the concatenation of its instances' plans."""
        r = self.items[x].radiosity
        rows = r.TILE_ROWS
        r.TILE_ROWS = 0
        try:
            return r._exchange_plan(self.c, self.dt, kind)
        except ValueError:
            return None
        finally:
            r.TILE_ROWS = rows

    def _classes(self):
        cls = {}
        for x, it in enumerate(self.items):
            b = it.radiosity._bake
            cls.setdefault((b["D"], b["n_bins"]), []).append(x)
        return cls

    def _instance_bytes(self, x, plan):
        r = self.items[x].radiosity
        P = r.n_patches
        D, B = r._bake["D"], r._bake["n_bins"]
        G = plan.n_groups if isinstance(plan, ops.GroupedPlan) else 0
        return 4 * self.T * B * (4 * P * D + G)

    def _chunks(self, xs, plans):
        """(scene, source) instances of a class, packed into chunks."""
        chunks, cur, used = [], [], 0
        for x in xs:
            nb = self._instance_bytes(x, plans[x])
            for s in range(len(self.items[x].sources)):
                if cur and used + nb > self.budget:
                    chunks.append(cur)
                    cur, used = [], 0
                cur.append((x, s))
                used += nb
        if cur:
            chunks.append(cur)
        return chunks

    def _kinds(self, xs):
        """Plan kinds to try for a class, best first: the grouped and
        diffuse + specular plans ordered by the time of one order on the
        class's largest scene (only that scene's plans are built for it),
        then the generic plan, which every scene supports."""
        x_big = max(xs, key=lambda x: self.items[x].radiosity.n_patches)
        times = {}
        for k in ("grouped", "ds"):
            plan = self._plan(x_big, k)
            if plan is None:
                continue
            if not self.autotune:
                times[k] = len(times)
                continue
            times[k] = self._time_order(x_big, plan, self.tiling)
        return sorted(times, key=times.get) + ["generic"]

    def _time_order(self, x, plan, tiling):
        """Seconds of one exchange order of scene x's first instance on
        ``plan`` with ``tiling`` (after one untimed order)."""
        # its own buffers: the pool is sized for the chunks
        run = self._build([(x, 0)], {x: plan}, pooled=False, tiling=tiling)
        st = run["state"]
        st.step(run["plan"])             # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st.step(run["plan"])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        del run, st
        return dt

    def _tune_tiling(self, classes):
        """Time the candidate tilings (TILINGS, for the plan kind of the
        class with the most work) on that class's largest scene; keep the
        fastest for every class."""
        def work(xs):
            return sum(self.items[x].radiosity.n_patches *
                       self.items[x].radiosity._bake["D"] *
                       self.items[x].radiosity._bake["n_bins"] *
                       len(self.items[x].sources) for x in xs)
        xs, kind, plans = max(classes, key=lambda c: work(c[0]))
        if kind not in self.TILINGS:
            return
        x_big = max(xs, key=lambda x: self.items[x].radiosity.n_patches)
        cands = [self.tiling]
        for t in self.TILINGS[kind]:
            t = self._tiling(*t)
            if t not in cands:
                cands.append(t)
        self.tiling_times = {t: self._time_order(x_big, plans[x_big], t)
                             for t in cands}
        self.tiling = min(self.tiling_times, key=self.tiling_times.get)

    def _class_plans(self, xs):
        """(kind, {scene: plan}) for the first kind every scene supports."""
        for kind in self._kinds(xs):
            plans = {}
            for x in xs:
                plans[x] = self._plan(x, kind)
                if plans[x] is None:
                    break
            else:
                return kind, plans
        raise RuntimeError("no exchange plan for a class")  # unreachable

    def _pool_sizes(self, PDB, G, B):
        """Pool buffers one chunk takes, in elements: four [P,D,B,T]
        histograms and the [max(G, 1),B,T] group sums."""
        return [PDB * self.T] * 4 + [max(G, 1) * B * self.T]

    def _chunk_sizes(self, chunk, plans):
        PDB = G = B = 0
        for x, _ in chunk:
            b = self.items[x].radiosity._bake
            B = b["n_bins"]          # one value for the class
            PDB += self.items[x].radiosity.n_patches * b["D"] * B
            p = plans[x]
            G += p.n_groups if isinstance(p, ops.GroupedPlan) else 0
        return self._pool_sizes(PDB, G, B)

    def _build(self, chunk, plans, pooled=True, tiling=None):
        xs = [x for x, _ in chunk]
        offsets = np.concatenate(
            [[0], np.cumsum([self.items[x].radiosity.n_patches
                             for x in xs])]).astype(np.int64)
        plan = _concat_plans([plans[x] for x in xs], offsets[:-1])
        rows, block, tile_r = tiling or self.tiling
        if rows and isinstance(plan, ops.GroupedPlan) and (
                not plan.diffuse_specular or rows == 16):
            # diffuse + specular plans tile at 16 rows only; with a tile
            # block, each room's patches in blocks of that shape, the rooms
            # one after the other
            order = None
            if block is not None:
                parts = []
                for x, off in zip(xs, offsets[:-1]):
                    r = self.items[x].radiosity
                    o = ops.block_order(r._patches_points,
                                        r._patch_to_wall_ids, block)
                    parts.append(np.where(o >= 0, o + off, -1))
                order = np.concatenate(parts)
            plan.tiles = ops.tile_plan(plan, rows, order=order)
            if plan.tiles is not None:
                plan.tiles.R = tile_r
        e0d, d0 = [], []
        for x, s in chunk:
            r = self.items[x].radiosity
            r.init_source_energy(self.items[x].sources[s])
            e0d.append(r._dev["energy_init_source"])
            d0.append(r._dev["distance_patches_to_source"])
        e0d, d0 = _cat(e0d), _cat(d0)
        P, D, B = e0d.shape
        pool = None
        if pooled:
            G = plan.n_groups if isinstance(plan, ops.GroupedPlan) else 0
            self.pool.begin(self._pool_sizes(P * D * B, G, B))
            pool = self.pool
        st = ops.ExchangeState(P, D, B, self.T, pool=pool)
        st.init(e0d, d0, self.c, self.dt)
        if isinstance(plan, ops.GroupedPlan):
            st._gsum(plan)
        return dict(plan=plan, state=st, e0d=e0d, d0=d0, offsets=offsets,
                    chunk=chunk)

    # -- running ---------------------------------------------------------------
    def _exchange(self, run):
        st, plan = run["state"], run["plan"]
        if not self.graphs or self.K == 0:
            for _ in range(self.K):
                st.step(plan)
            return
        a, b = st.prev, st.cur
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(graph, stream=side):
                for _ in range(self.K):
                    st.step(plan)
        torch.cuda.current_stream().wait_stream(side)
        # capture recorded the launches without running them: reset the
        # state to order 0 and replay (after which an odd K leaves its last
        # order out of the running total, as the captured steps did)
        pending = st.pending
        st.prev, st.cur = a, b
        st.init(run["e0d"], run["d0"], self.c, self.dt)
        graph.replay()
        st.pending = pending
        if self.K % 2:
            st.prev, st.cur = b, a
        run["graph"] = graph

    def run(self):
        """{(scene, source): receiver histograms float32 [R, B, T]}."""
        out = {}
        self.updates = 0
        work = []
        classes = []
        for key, xs in self._classes().items():
            kind, plans = self._class_plans(xs)
            self.choice[key] = kind
            classes.append((xs, kind, plans))
        if self.autotune and self.tune_tiles:
            self._tune_tiling(classes)
        for xs, kind, plans in classes:
            work.append((plans, self._chunks(xs, plans)))
        # the pool is sized once, for the largest chunk of every class: a
        # freed buffer of tens of GiB is soon pinned by small allocations
        # placed in its tail, so growing it later would need the memory twice
        sizes = [self._chunk_sizes(c, plans) for plans, chunks in work
                 for c in chunks]
        self.pool.begin([max(z) for z in zip(*sizes)])
        for plans, chunks in work:
            for chunk in chunks:
                run = self._build(chunk, plans)
                self._exchange(run)
                st = run["state"]
                st.check()
                total = st.finalize(out=st.hi)
                off = run["offsets"]
                for n, (x, s) in enumerate(chunk):
                    pts, factor, idx, centers, air = self._receivers[x]
                    slab = total[int(off[n]):int(off[n + 1])]
                    out[(x, s)] = ops.collect(slab, pts, centers, idx,
                                              factor, air, self.c, self.dt)
                    del slab      # a view of the pool
                plan = run["plan"]
                lag = plan.lag.to(torch.int64) & (
                    ops.GroupedPlan.LAST_IN_GROUP - 1)
                self.updates += int(
                    (st.shape[2] * torch.clamp(self.T - lag, min=0)).sum()
                    .item()) * self.K
                del run, st, total
        return out


def run_per_instance(items, speed_of_sound, time_resolution, duration,
                     max_order):
    """The reference's loop: one ``init_source_energy`` /
    ``calculate_energy_exchange`` / ``collect_energy_receiver_mono`` per
    (scene, source).  {(scene, source): float32 [R, B, T] device tensor}."""
    out = {}
    for x, it in enumerate(items):
        r = it.radiosity
        pts, _, factor, idx = r._receiver_terms(it.receivers)
        g = r._geometry()
        air = ops.as_dev(r._air_attenuation, torch.float64)
        for s, src in enumerate(it.sources):
            r.init_source_energy(src)
            r.calculate_energy_exchange(speed_of_sound, time_resolution,
                                        duration, max_order,
                                        recalculate=True)
            out[(x, s)] = ops.collect(r._etc, pts, g["centers"], idx, factor,
                                      air, speed_of_sound, time_resolution)
            r._etc = None
    return out
