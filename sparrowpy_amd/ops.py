"""Operator layer: one function per kernel function of the reference path.

Each function takes numpy arrays or device tensors, runs the HIP kernel on
the current stream and returns device tensors.  Names and argument meaning
follow the reference's module-level functions (file:line in each
docstring); see :mod:`sparrowpy_amd.radiosity` for the class that chains
them like ``DirectionalRadiosityFast``.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from . import _hip
from ._hip import call, ptr, stream_ptr

__all__ = [
    "device", "as_dev", "check_patch2patch_visibility",
    "check_point2patch_visibility", "visible_pairs",
    "patch2patch_ff_universal", "direction_map", "point_direction_index",
    "form_factors_with_directivity_dim", "source_energy",
    "receiver_energy", "collect", "energy_exchange", "ExchangePlan",
    "GroupedPlan", "ExchangeState",
]


def device():
    _hip.lib()  # raises HipUnavailable without a GPU or library
    return torch.device("cuda", torch.cuda.current_device())


def as_dev(x, dtype):
    """Contiguous device tensor of ``dtype`` (copies only if needed)."""
    if x is None:
        return None
    if isinstance(x, torch.Tensor):
        t = x.to(device=device(), dtype=dtype)
    else:
        t = torch.as_tensor(np.ascontiguousarray(x), dtype=dtype,
                            device=device())
    return t.contiguous()


def _f64(x):
    return as_dev(x, torch.float64)


def _i64(x):
    return as_dev(x, torch.int64)


def _u8(x):
    return as_dev(x, torch.uint8)


def _vis_workspace(ns):
    n = int(_hip.lib().spr_visibility_workspace(int(ns)))
    return torch.empty(n, dtype=torch.uint8, device=device())


def check_patch2patch_visibility(patches_center, surf_normal, surf_points):
    """geometry.py:750-797 ``_check_patch2patch_visibility`` -> uint8 [P,P]
    (upper triangle)."""
    c = _f64(patches_center)
    n = _f64(surf_normal)
    p = _f64(surf_points)
    P, NS = c.shape[0], n.shape[0]
    if p.shape[1:] != (4, 3):
        raise ValueError("surfaces must have 4 corners")
    vis = torch.empty((P, P), dtype=torch.uint8, device=c.device)
    ws = _vis_workspace(NS)
    call("spr_p2p_visibility", P, ptr(c), NS, ptr(n), ptr(p), ptr(vis),
         ptr(ws), ws.numel(), stream_ptr())
    return vis


def check_point2patch_visibility(eval_points, patches_center, surf_normal,
                                 surf_points):
    """geometry.py:799-839 ``_check_point2patch_visibility``, batched over
    query points [NQ,3] -> uint8 [NQ,P]."""
    q = _f64(np.atleast_2d(eval_points) if not isinstance(
        eval_points, torch.Tensor) else eval_points.reshape(-1, 3))
    c = _f64(patches_center)
    n = _f64(surf_normal)
    p = _f64(surf_points)
    NQ, P, NS = q.shape[0], c.shape[0], n.shape[0]
    vis = torch.empty((NQ, P), dtype=torch.uint8, device=c.device)
    ws = _vis_workspace(NS)
    call("spr_point_visibility", NQ, ptr(q), P, ptr(c), NS, ptr(n), ptr(p),
         ptr(vis), ptr(ws), ws.numel(), stream_ptr())
    return vis


def visible_pairs(vis):
    """RadiosityFast.py:377-387 -> int32 [V,2], row-major i<j."""
    vis = _u8(vis)
    P = vis.shape[0]
    offs = torch.empty(P + 1, dtype=torch.int64, device=vis.device)
    call("spr_visible_pairs", P, ptr(vis), ptr(offs), None, stream_ptr())
    V = int(offs[P].item())
    pairs = torch.empty((V, 2), dtype=torch.int32, device=vis.device)
    if V:
        call("spr_visible_pairs", P, ptr(vis), ptr(offs), ptr(pairs),
             stream_ptr())
    return pairs


def patch2patch_ff_universal(patches_points, patches_normals, patches_areas,
                             visible_patches):
    """universal.py:12-52 ``patch2patch_ff_universal`` -> float64 [P,P]."""
    pts = _f64(patches_points)
    nrm = _f64(patches_normals)
    area = _f64(patches_areas)
    pairs = as_dev(visible_patches, torch.int32).reshape(-1, 2)
    P = pts.shape[0]
    ff = torch.empty((P, P), dtype=torch.float64, device=pts.device)
    call("spr_form_factors", P, ptr(pts), ptr(nrm), ptr(area),
         pairs.shape[0], ptr(pairs), ptr(ff), stream_ptr())
    return ff


def direction_map(vis, patches_center, out_dirs, wall_ids):
    """RadiosityFast.py:403-414 (``get_scattering_data_receiver_index``
    over visible pairs) -> int64 [P,P], sentinel D."""
    vis = _u8(vis)
    c = _f64(patches_center)
    od = _f64(out_dirs)
    w = _i64(wall_ids)
    P, D = c.shape[0], od.shape[1]
    p2o = torch.empty((P, P), dtype=torch.int64, device=c.device)
    call("spr_direction_map", P, ptr(vis), ptr(c), D, ptr(od), ptr(w),
         ptr(p2o), stream_ptr())
    return p2o


def point_direction_index(points, patches_center, out_dirs, wall_ids):
    """``get_scattering_data_receiver_index(centers, point, ...)``
    (RadiosityFast.py:728-730) for each point -> int64 [NQ,P]."""
    q = _f64(points).reshape(-1, 3)
    c = _f64(patches_center)
    od = _f64(out_dirs)
    w = _i64(wall_ids)
    NQ, P, D = q.shape[0], c.shape[0], od.shape[1]
    idx = torch.empty((NQ, P), dtype=torch.int64, device=c.device)
    call("spr_point_direction_index", NQ, ptr(q), P, ptr(c), D, ptr(od),
         ptr(w), ptr(idx), stream_ptr())
    return idx


def form_factors_with_directivity_dim(visibility_matrix, form_factors,
                                      n_bins, patches_center, patches_area,
                                      air_attenuation, patch_to_wall_ids,
                                      scattering, scattering_index, sources,
                                      receivers):
    """RadiosityFast.py:1234-1272 -> float64 [P,P,D,B] (dense)."""
    vis = _u8(visibility_matrix)
    ff = _f64(form_factors)
    c = _f64(patches_center)
    area = _f64(patches_area)
    air = _f64(air_attenuation)
    w = _i64(patch_to_wall_ids)
    P = c.shape[0]
    if scattering is None:
        S, D = 1, 1
        brdf = bidx = src = None
    else:
        brdf = _f64(scattering)
        bidx = _i64(scattering_index)
        src = _f64(sources)
        S, D = src.shape[1], int(np.shape(receivers)[1])
    out = torch.empty((P, P, D, n_bins), dtype=torch.float64, device=c.device)
    call("spr_fft_dense", P, ptr(vis), ptr(ff), n_bins, ptr(c), ptr(area),
         ptr(air), ptr(w), S, D, ptr(brdf), ptr(bidx), ptr(src), ptr(out),
         stream_ptr())
    return out


def source_energy(source_position, patches_center, patches_points,
                  source_visibility, air_attenuation, n_bins,
                  patch_to_wall_ids, sources, scattering, scattering_index,
                  n_directions):
    """universal.py:98-147 + RadiosityFast.py:988-1034.

    Returns (energy_0 [P,B], energy_0_dir [P,D,B], distance_0 [P]), float64.
    """
    s = _f64(np.asarray(source_position, float).reshape(3)
             if not isinstance(source_position, torch.Tensor)
             else source_position.reshape(3))
    c = _f64(patches_center)
    pts = _f64(patches_points)
    vis = _u8(source_visibility)
    air = _f64(air_attenuation)
    w = _i64(patch_to_wall_ids)
    src = _f64(sources)
    brdf = _f64(scattering)
    bidx = _i64(scattering_index)
    P, S, D = c.shape[0], src.shape[1], int(n_directions)
    e0 = torch.empty((P, n_bins), dtype=torch.float64, device=c.device)
    e0d = torch.empty((P, D, n_bins), dtype=torch.float64, device=c.device)
    d0 = torch.empty(P, dtype=torch.float64, device=c.device)
    call("spr_source_energy", ptr(s), P, ptr(c), ptr(pts), ptr(vis),
         ptr(air), n_bins, ptr(w), S, D, ptr(brdf), ptr(bidx), ptr(src),
         ptr(e0), ptr(e0d), ptr(d0), stream_ptr())
    return e0, e0d, d0


def receiver_energy(points, patches_points, receiver_visibility):
    """universal.py:149-160, batched over points -> float64 [NQ,P]."""
    q = _f64(points).reshape(-1, 3)
    pts = _f64(patches_points)
    vis = _u8(receiver_visibility).reshape(q.shape[0], -1)
    NQ, P = q.shape[0], pts.shape[0]
    out = torch.empty((NQ, P), dtype=torch.float64, device=pts.device)
    call("spr_receiver_energy", NQ, ptr(q), P, ptr(pts), ptr(vis), ptr(out),
         stream_ptr())
    return out


def collect(etc, points, patches_center, idx, factor, air_attenuation,
            speed_of_sound, time_resolution, patchwise=False):
    """RadiosityFast.py:686-752, 1148-1185 (and the patch sum at :592).

    etc float32 [P,D,B,T] -> float32 [NR,B,T] or [NR,P,B,T] (patchwise).
    """
    etc = as_dev(etc, torch.float32)
    q = _f64(points).reshape(-1, 3)
    c = _f64(patches_center)
    idx = _i64(idx)
    g = _f64(factor)
    air = _f64(air_attenuation)
    P, D, B, T = etc.shape
    NR = q.shape[0]
    shape = (NR, P, B, T) if patchwise else (NR, B, T)
    out = torch.empty(shape, dtype=torch.float32, device=etc.device)
    call("spr_collect", NR, P, D, B, T, ptr(etc), ptr(q), ptr(c), ptr(idx),
         ptr(g), ptr(air), float(speed_of_sound), float(time_resolution),
         1 if patchwise else 0, ptr(out), stream_ptr())
    return out


def energy_exchange(n_samples, energy_0_directivity, distance_0,
                    distance_ij, form_factors_tilde, patch_2_out_directions,
                    speed_of_sound, histogram_time_resolution, max_order,
                    visible_patches):
    """RadiosityFast.py:1073-1145 ``_energy_exchange`` with the reference's
    arguments -> float32 [P,D,B,T]."""
    e0d = _f64(energy_0_directivity)
    d0 = _f64(distance_0)
    dist = _f64(distance_ij)
    fft = _f64(form_factors_tilde)
    p2o = _i64(patch_2_out_directions)
    pairs = as_dev(visible_patches, torch.int32).reshape(-1, 2)
    P, D, B = e0d.shape
    T = int(n_samples)
    V = pairs.shape[0]
    etc = torch.empty((P, D, B, T), dtype=torch.float32, device=e0d.device)
    nws = int(_hip.lib().spr_energy_exchange_workspace(P, D, B, T, V))
    ws = torch.empty(nws, dtype=torch.uint8, device=e0d.device)
    call("spr_energy_exchange", T, P, D, B, ptr(e0d), ptr(d0), ptr(dist),
         ptr(fft), ptr(p2o), float(speed_of_sound),
         float(histogram_time_resolution), int(max_order), V, ptr(pairs),
         ptr(etc), ptr(ws), nws, stream_ptr())
    return etc


def energy_exchange_groups():
    """Groups of the grouped plan the last energy_exchange call on this
    thread ran from the dense F~ (0: the per-pair plan ran)."""
    return int(_hip.lib().spr_energy_exchange_groups())


@dataclasses.dataclass
class ExchangePlan:
    """Destination-major pair table of one scene (see sparrow_hip.h)."""

    P: int
    D: int
    B: int
    row_ptr: torch.Tensor
    src: torch.Tensor
    lag: torch.Tensor
    dir: torch.Tensor
    coef: torch.Tensor

    @property
    def Vd(self):
        return int(self.src.numel())

    @classmethod
    def rows(cls, vis):
        vis = _u8(vis)
        P = vis.shape[0]
        row_ptr = torch.empty(P + 1, dtype=torch.int64, device=vis.device)
        call("spr_plan_rows", P, ptr(vis), ptr(row_ptr), None, stream_ptr())
        Vd = int(row_ptr[P].item())
        src = torch.empty(max(Vd, 1), dtype=torch.int32, device=vis.device)
        if Vd:
            call("spr_plan_rows", P, ptr(vis), ptr(row_ptr), ptr(src),
                 stream_ptr())
        return row_ptr, src[:Vd]

    @classmethod
    def _alloc(cls, P, D, B, row_ptr, src):
        DP = int(_hip.lib().spr_padded_directions(D))
        if DP == 0:
            raise ValueError(f"{D} directions unsupported (max 64)")
        Vd = src.numel()
        dev = src.device
        lag = torch.empty(max(Vd, 1), dtype=torch.int32, device=dev)[:Vd]
        dirs = torch.empty(max(Vd, 1), dtype=torch.int32, device=dev)[:Vd]
        coef = torch.empty((max(Vd, 1), B, DP), dtype=torch.float64,
                           device=dev)[:Vd]
        return cls(P, D, B, row_ptr, src, lag, dirs, coef)

    @classmethod
    def from_dense(cls, vis, distance_ij, patch_2_out_directions,
                   form_factors_tilde, speed_of_sound, time_resolution):
        fft = _f64(form_factors_tilde)
        P, _, D, B = fft.shape
        row_ptr, src = cls.rows(vis)
        plan = cls._alloc(P, D, B, row_ptr, src)
        dist = _f64(distance_ij)
        p2o = _i64(patch_2_out_directions)
        call("spr_plan_from_dense", P, D, B, float(speed_of_sound),
             float(time_resolution), plan.Vd, ptr(row_ptr), ptr(src),
             ptr(dist), ptr(p2o), ptr(fft), ptr(plan.lag), ptr(plan.dir),
             ptr(plan.coef), stream_ptr())
        return plan

    @classmethod
    def from_geometry(cls, vis, patches_center, patches_area, form_factors,
                      p2o, air_attenuation, patch_to_wall_ids, sources,
                      scattering, scattering_index, n_directions, n_bins,
                      speed_of_sound, time_resolution):
        c = _f64(patches_center)
        P = c.shape[0]
        D, B = int(n_directions), int(n_bins)
        row_ptr, src = cls.rows(vis)
        plan = cls._alloc(P, D, B, row_ptr, src)
        area = _f64(patches_area)
        ff = _f64(form_factors)
        p2o = _i64(p2o)
        air = _f64(air_attenuation)
        w = _i64(patch_to_wall_ids)
        if scattering is None:
            S, brdf, bidx, sd = 1, None, None, None
        else:
            sd = _f64(sources)
            S = sd.shape[1]
            brdf = _f64(scattering)
            bidx = _i64(scattering_index)
        call("spr_plan_from_geometry", P, D, B, float(speed_of_sound),
             float(time_resolution), plan.Vd, ptr(row_ptr), ptr(src), ptr(c),
             ptr(area), ptr(ff), ptr(p2o), ptr(air), ptr(w), S, ptr(brdf),
             ptr(bidx), ptr(sd), ptr(plan.lag), ptr(plan.dir),
             ptr(plan.coef), stream_ptr())
        return plan


@dataclasses.dataclass
class GroupedPlan:
    """Destination-major pair table with sources grouped by BRDF row.

    ``rtab[k, b, :]`` is the BRDF row k = bi*S + s; ``scal[b, e]`` the
    per-pair weight F~_ij * exp(-air_b |u_ij|); ``lag`` carries the
    last-in-group flag in bit 30; see sparrow_hip.h.
    """

    LAST_IN_GROUP = 0x40000000

    P: int
    D: int
    B: int
    row_ptr: torch.Tensor
    grp_ptr: torch.Tensor
    grp_start: torch.Tensor
    grp_key: torch.Tensor
    rowid: torch.Tensor
    lag: torch.Tensor
    scal: torch.Tensor
    rtab: torch.Tensor
    dscal: torch.Tensor = None  # diffuse weights (diffuse + specular plan)
    tiles: "TilePlan" = None    # merged destination tiles (tile_plan)

    @property
    def Vd(self):
        return int(self.rowid.numel())

    @property
    def diffuse_specular(self):
        return self.dscal is not None

    @property
    def n_groups(self):
        return int(self.grp_key.numel())

    MAX_ROW = 8192

    @classmethod
    def from_geometry(cls, vis, patches_center, patches_area, form_factors,
                      p2o, air_attenuation, patch_to_wall_ids, sources,
                      scattering, scattering_index, n_directions, n_bins,
                      speed_of_sound, time_resolution):
        c = _f64(patches_center)
        dev = c.device
        P = c.shape[0]
        D, B = int(n_directions), int(n_bins)
        DP = int(_hip.lib().spr_padded_directions(D))
        if DP == 0:
            raise ValueError(f"{D} directions unsupported (max 64)")
        row_ptr, src = ExchangePlan.rows(vis)
        Vd = src.numel()
        max_row = int((row_ptr[1:] - row_ptr[:-1]).max().item()) if P else 0
        if max_row > cls.MAX_ROW:
            raise ValueError("destination rows too long for the grouped plan")
        src = src.clone() if Vd else torch.empty(1, dtype=torch.int32,
                                                 device=dev)
        w = _i64(patch_to_wall_ids)
        if scattering is None:
            S, NK, brdf, bidx, sd = 1, 1, None, None, None
        else:
            sd = _f64(sources)
            S = sd.shape[1]
            brdf = _f64(scattering)
            bidx = _i64(scattering_index)
            NK = brdf.shape[0] * S
        key = torch.empty(max(Vd, 1), dtype=torch.int32, device=dev)
        grp_ptr = torch.empty(P + 1, dtype=torch.int64, device=dev)
        call("spr_plan_grouped_sort", P, max_row, ptr(row_ptr), ptr(src),
             ptr(c), ptr(w), S, ptr(bidx), ptr(sd), ptr(key), ptr(grp_ptr),
             stream_ptr())
        G = int(grp_ptr[P].item())
        grp_start = torch.empty(G + 1, dtype=torch.int64, device=dev)
        grp_key = torch.empty(max(G, 1), dtype=torch.int32, device=dev)[:G]
        rowid = torch.empty(max(Vd, 1), dtype=torch.int32, device=dev)
        lag = torch.empty(max(Vd, 1), dtype=torch.int32, device=dev)
        scal = torch.empty((B, max(Vd, 1)), dtype=torch.float64, device=dev)
        rtab = torch.empty((NK, B, DP), dtype=torch.float64, device=dev)
        call("spr_plan_grouped_fill", P, D, B, float(speed_of_sound),
             float(time_resolution), Vd, ptr(row_ptr), ptr(src), ptr(key),
             ptr(grp_ptr), ptr(c), ptr(_f64(patches_area)),
             ptr(_f64(form_factors)), ptr(_i64(p2o)),
             ptr(_f64(air_attenuation)), NK, ptr(brdf), ptr(grp_start),
             ptr(grp_key), ptr(rowid), ptr(lag), ptr(scal), ptr(rtab),
             stream_ptr())
        return cls(P, D, B, row_ptr, grp_ptr, grp_start, grp_key,
                   rowid[:Vd], lag[:Vd], scal[:, :Vd].contiguous(), rtab)


    # entry check of from_dense: |F~_e - s_e F~_key| <= ROW_TOL max|F~_e|
    # per band, element by element
    ROW_TOL = 2.0 ** -28

    @classmethod
    def from_dense(cls, vis, distance_ij, patch_2_out_directions,
                   form_factors_tilde, speed_of_sound, time_resolution,
                   max_group_fraction=0.25):
        """Grouped plan of a dense F~ [P,P,D,B] whose rows factorise, or
        None.

        F~[i, j, :, b] = f_ij exp(-air_b d_ij) brdf[wall_i, s_ij, :, b]
        (RadiosityFast.py:1234-1272), so the entries that share a BRDF row
        have rows that are band by band multiples of one vector.  Entries
        are keyed by the hash of their normalised row
        (spr_plan_dense_hash); keys are global, so destination tiles can
        merge entries of neighbouring destinations.  A key's table row is
        the F~ row of its first entry, and each entry's per-band scalar is
        s_eb = F~_e[piv] / F~_key[piv] at the key row's first element of
        largest magnitude.  Every entry is checked:
        |F~_e - s_e F~_key| <= 2^-28 max_d |F~_e| per band.  None when a
        check fails (a hash collision, or values that are not numbers) or
        when the groups exceed max_group_fraction of the entries, where the
        per-pair plan (ExchangePlan.from_dense) is the faster one.
        """
        base = ExchangePlan.from_dense(vis, distance_ij,
                                       patch_2_out_directions,
                                       form_factors_tilde, speed_of_sound,
                                       time_resolution)
        P, D, B, Vd = base.P, base.D, base.B, base.Vd
        if Vd == 0:
            return None
        fft = _f64(form_factors_tilde)
        dev = fft.device
        hkey = torch.empty(Vd, dtype=torch.int64, device=dev)
        call("spr_plan_dense_hash", P, D, B, ptr(base.row_ptr),
             ptr(base.src), ptr(fft), ptr(hkey), stream_ptr())
        _, key = torch.unique(hkey, return_inverse=True)
        NK = int(key.max().item()) + 1
        first = torch.full((NK,), Vd, dtype=torch.int64, device=dev)
        first.scatter_reduce_(0, key, torch.arange(Vd, device=dev), "amin")
        coef = base.coef                           # [Vd, B, DP]
        rtab = coef[first].contiguous()            # [NK, B, DP]
        piv = rtab[..., :D].abs().argmax(dim=-1)   # first maximum
        rpiv = rtab.gather(2, piv[..., None])[..., 0]
        scal = torch.empty((Vd, B), dtype=torch.float64, device=dev)
        step = max(1, (1 << 26) // (B * rtab.shape[2]))
        for e0 in range(0, Vd, step):
            k = key[e0:e0 + step]
            f = coef[e0:e0 + step, :, :D]
            rp = rpiv[k]
            fp = f.gather(2, piv[k][..., None])[..., 0]
            sc = torch.where(rp != 0, fp / torch.where(rp != 0, rp, 1.0),
                             torch.zeros_like(fp))
            err = (f - sc[..., None] * rtab[k, :, :D]).abs()
            bound = cls.ROW_TOL * f.abs().amax(dim=-1, keepdim=True)
            if not bool((err <= bound).all().item()):
                return None
            scal[e0:e0 + step] = sc
        # each destination's entries by key (stable: sources ascending)
        counts = base.row_ptr[1:] - base.row_ptr[:-1]
        j = torch.repeat_interleave(torch.arange(P, device=dev), counts)
        comp, perm = torch.sort(j * NK + key, stable=True)
        new = torch.ones(Vd, dtype=torch.bool, device=dev)
        new[1:] = comp[1:] != comp[:-1]
        starts = torch.nonzero(new)[:, 0]
        G = int(starts.numel())
        if G > max_group_fraction * Vd:
            return None
        grp_ptr = torch.zeros(P + 1, dtype=torch.int64, device=dev)
        grp_ptr[1:] = torch.cumsum(torch.bincount(comp[starts] // NK,
                                                  minlength=P), 0)
        grp_start = torch.empty(G + 1, dtype=torch.int64, device=dev)
        grp_start[:G] = starts
        grp_start[G] = Vd
        last = torch.ones(Vd, dtype=torch.bool, device=dev)
        last[:-1] = new[1:]
        src = base.src[perm].to(torch.int64)
        rowid = (src * D + base.dir[perm].to(torch.int64)).to(torch.int32)
        lag = base.lag[perm] | (last.to(torch.int32) * cls.LAST_IN_GROUP)
        grp_key = key[perm][starts].to(torch.int32)
        return cls(P, D, B, base.row_ptr, grp_ptr, grp_start, grp_key,
                   rowid.contiguous(), lag.contiguous(),
                   scal[perm].t().contiguous(), rtab)


@dataclasses.dataclass
class TilePlan:
    """Destination tiles of a GroupedPlan: JT rows merged into M distinct
    (group key, source row) entries with one delay word and one weight per
    slot; ``tptr`` is the tiles' CSR over the merged entries, ``mlag[M,
    JT]`` and ``mw[B, M, JT]`` (see sparrow_hip.h).  ``P`` is the number of
    slots (tiles x JT rows, empty rows included) and ``gstart[slot]`` the
    index of the slot's first group in the plan (where its sums go)."""

    JT: int
    M: int
    P: int
    gstart: torch.Tensor
    tptr: torch.Tensor
    mrow: torch.Tensor
    mlag: torch.Tensor
    mw: torch.Tensor
    # diffuse + specular plan: diffuse weights [B, M, JT] and each slot's
    # diffuse group (-1: none)
    mdw: torch.Tensor = None
    gdiff: torch.Tensor = None
    # 16-row tiles: bins per lane of the gather's workgroups (16 or 8) and
    # each merged entry's span [M, 4] (int32: source row, largest and
    # smallest delay of its slots, 0), from which the gather stages windows
    R: int = 16
    mspan: torch.Tensor = None


def _rows_in_order(plan, order):
    """The grouped plan's arrays with its destination rows in ``order``
    (destination index per new row, -1 for an empty row): a dict of
    row_ptr, grp_start, grp_key (grouped plan only), rowid, lag, scal,
    dscal and entry_key (diffuse + specular plan only), gstart and gend.
    Every row keeps its entries and groups in their order; gstart[r] and
    gend[r] are the plan's indices of the first and one past the last group
    of the row's destination (0 for an empty row)."""
    dev = plan.row_ptr.device
    order = torch.as_tensor(np.asarray(order, np.int64), device=dev)
    real = order >= 0
    zero = torch.zeros_like(order)
    src = torch.where(real, order, zero)
    rp, gp = plan.row_ptr, plan.grp_ptr
    cnt = torch.where(real, rp[src + 1] - rp[src], zero)
    row_ptr = torch.zeros(len(order) + 1, dtype=torch.int64, device=dev)
    row_ptr[1:] = torch.cumsum(cnt, 0)
    Vd = int(row_ptr[-1])
    # entry e of new row r comes from rp[src[r]] + (e - row_ptr[r])
    r_of_e = torch.repeat_interleave(torch.arange(len(order), device=dev),
                                     cnt)
    eidx = rp[src][r_of_e] + (torch.arange(Vd, device=dev) - row_ptr[r_of_e])
    out = dict(row_ptr=row_ptr, rowid=plan.rowid[eidx].contiguous(),
               lag=plan.lag[eidx].contiguous(),
               scal=plan.scal[:, eidx].contiguous(),
               gstart=torch.where(real, gp[src], zero),
               gend=torch.where(real, gp[src + 1], zero))
    if plan.diffuse_specular:
        out["dscal"] = plan.dscal[:, eidx].contiguous()
        out["entry_key"] = plan.entry_key[eidx].contiguous()
        return out
    gcnt = torch.where(real, gp[src + 1] - gp[src], zero)
    g_of_r = torch.repeat_interleave(torch.arange(len(order), device=dev),
                                     gcnt)
    gofs = torch.zeros(len(order) + 1, dtype=torch.int64, device=dev)
    gofs[1:] = torch.cumsum(gcnt, 0)
    G = int(gofs[-1])
    gidx = gp[src][g_of_r] + (torch.arange(G, device=dev) - gofs[g_of_r])
    # group g of the plan starts at entry grp_start[g] (every group of the
    # grouped plan has entries); in the new order at that entry's position
    new_pos = torch.empty(plan.Vd, dtype=torch.int64, device=dev)
    new_pos[eidx] = torch.arange(Vd, device=dev)
    grp_start = torch.empty(G + 1, dtype=torch.int64, device=dev)
    grp_start[:G] = new_pos[plan.grp_start[gidx]]
    grp_start[G] = Vd
    out["grp_start"] = grp_start
    out["grp_key"] = plan.grp_key[gidx].contiguous()
    return out


def tile_plan(plan, jt=4, min_merge=1.5, order=None):
    """TilePlan of a GroupedPlan, or None when tiles of ``jt`` rows merge
    fewer than ``min_merge`` entries into one.

    Grouped plans take ``jt`` of 4, 8 or 16 rows, diffuse + specular plans
    16 (merged by each entry's specular direction; the tiled gather also
    sums every row's diffuse group).  ``order`` (optional): the
    destination of each slot, tile after tile (-1 for an empty slot), e.g.
    from :func:`block_order`, so that a tile can be a compact block of
    patches rather than ``jt`` consecutive ids.  Every destination must
    appear once."""
    ds = plan.diffuse_specular
    jt = int(jt)
    if ds and jt != 16:
        raise ValueError("diffuse + specular tiles have 16 rows")
    B = plan.B
    dev = plan.row_ptr.device
    if order is None:
        order = np.arange(plan.P, dtype=np.int64)
    else:
        order = np.asarray(order, np.int64)
        real = np.sort(order[order >= 0])
        if len(real) != plan.P or not np.array_equal(real,
                                                    np.arange(plan.P)):
            raise ValueError("order must list every destination once")
    if len(order) % jt:
        order = np.concatenate([order, np.full(jt - len(order) % jt, -1)])
    if not ds and len(order) == plan.P + (-plan.P) % jt and \
            np.array_equal(order[:plan.P], np.arange(plan.P)):
        # consecutive rows: the plan's own arrays
        q = dict(row_ptr=torch.cat([plan.row_ptr, plan.row_ptr[-1:].expand(
                     len(order) - plan.P)]),
                 grp_start=plan.grp_start, grp_key=plan.grp_key,
                 rowid=plan.rowid, lag=plan.lag, scal=plan.scal)
        gstart = torch.cat([plan.grp_ptr[:-1], torch.zeros(
            len(order) - plan.P, dtype=torch.int64, device=dev)])
    else:
        q = _rows_in_order(plan, order)
        gstart = q["gstart"]
    P, Vd = len(order), int(q["rowid"].numel())
    n_tiles = P // jt
    tptr = torch.empty(n_tiles + 1, dtype=torch.int64, device=dev)
    first = torch.empty(Vd + 1, dtype=torch.int64, device=dev)
    work = [torch.empty(max(Vd, 1), dtype=torch.int64, device=dev)
            for _ in range(3)]
    skey, srt, payload = work
    if ds:
        # merge key: the entry's own specular direction (entries without
        # one have no group, so their group's key would be wrong)
        skey[:Vd] = (q["entry_key"].to(torch.int64) << 32) | \
            q["rowid"].to(torch.int64)
        call("spr_plan_tile_order", P, jt, Vd, 0, ptr(q["row_ptr"]), None,
             None, ptr(q["rowid"]), ptr(skey), ptr(srt), ptr(payload),
             ptr(first), ptr(tptr), stream_ptr())
    else:
        G = int(q["grp_key"].numel())
        call("spr_plan_tile_order", P, jt, Vd, G, ptr(q["row_ptr"]),
             ptr(q["grp_start"]), ptr(q["grp_key"]), ptr(q["rowid"]),
             ptr(skey), ptr(srt), ptr(payload), ptr(first), ptr(tptr),
             stream_ptr())
    M = int(first[Vd].item())
    if M == 0 or Vd < min_merge * M:
        return None
    del skey
    mrow = torch.empty(M, dtype=torch.int32, device=dev)
    mlag = torch.empty((M, jt), dtype=torch.int32, device=dev)
    mw = torch.empty((B, M, jt), dtype=torch.float64, device=dev)
    call("spr_plan_tile_fill", P, jt, B, Vd, M, ptr(q["row_ptr"]),
         ptr(q["rowid"]), ptr(q["lag"]), ptr(q["scal"]), ptr(srt),
         ptr(payload), ptr(first), ptr(mrow), ptr(mlag), ptr(mw),
         stream_ptr())
    tp = TilePlan(jt, M, P, gstart.contiguous(), tptr, mrow, mlag, mw)
    if jt == 16:
        tp.mspan = torch.empty((M, 4), dtype=torch.int32, device=dev)
        call("spr_plan_tile_span", M, jt, ptr(mrow), ptr(mlag),
             ptr(tp.mspan), stream_ptr())
    if ds:
        # the same fill with the diffuse weights (mrow and mlag are
        # written again with the same values)
        tp.mdw = torch.empty((B, M, jt), dtype=torch.float64, device=dev)
        call("spr_plan_tile_fill", P, jt, B, Vd, M, ptr(q["row_ptr"]),
             ptr(q["rowid"]), ptr(q["lag"]), ptr(q["dscal"]), ptr(srt),
             ptr(payload), ptr(first), ptr(mrow), ptr(mlag), ptr(tp.mdw),
             stream_ptr())
        has = q["gend"] > q["gstart"]
        tp.gdiff = torch.where(has, q["gend"] - 1,
                               torch.full_like(q["gend"], -1)).contiguous()
    return tp


def block_order(patch_points, patch_to_wall_ids, block=(4, 4), pad=True):
    """Destination order in which every run of block[0]*block[1] slots is
    one block of neighbouring patches of one wall (the patches of a wall
    form a regular grid; blocks at a wall's edge are padded with empty
    slots, -1).  With ``pad=False`` the blocks follow each other without
    padding (a tile may then hold the end of one block and the start of
    the next, but no empty slot).  For tile_plan(..., order=...)."""
    pts = np.asarray(patch_points, float)
    wall = np.asarray(patch_to_wall_ids)
    ctr = pts.mean(axis=1)
    a, b = int(block[0]), int(block[1])
    out = []
    for w in np.unique(wall):
        idx = np.nonzero(wall == w)[0]
        c = ctr[idx]
        ext = c.max(0) - c.min(0)
        ax = np.argsort(-ext, kind="stable")[:2]
        # grid coordinates along the wall's two in-plane axes
        size = np.abs(pts[idx, 2] - pts[idx, 0]).max(0)
        size = np.where(size > 0, size, 1.0)
        u = np.rint((c[:, ax[0]] - c[:, ax[0]].min()) / size[ax[0]])
        v = np.rint((c[:, ax[1]] - c[:, ax[1]].min()) / size[ax[1]])
        u, v = u.astype(np.int64), v.astype(np.int64)
        bu, bv = u // a, v // b
        for key in sorted(set(zip(bu.tolist(), bv.tolist()))):
            mask = (bu == key[0]) & (bv == key[1])
            sel = idx[mask][np.lexsort((v[mask], u[mask]))]
            out.extend(sel.tolist())
            if pad:
                out.extend([-1] * (a * b - len(sel)))
    return np.array(out, np.int64)


def diffuse_specular_tables(scattering, n_directions, n_bins, n_rows):
    """Split BRDF rows into a diffuse constant and one specular entry
    (``spr_brdf_diffuse_specular``).  Returns (a [NK,B], c [NK,B],
    spec_dir [NK]) device tensors, or None if a row has another form."""
    D, B = int(n_directions), int(n_bins)
    brdf = None if scattering is None else _f64(scattering)
    NK = int(n_rows)
    dev = device()
    a = torch.empty((NK, B), dtype=torch.float64, device=dev)
    c = torch.empty((NK, B), dtype=torch.float64, device=dev)
    m = torch.empty(NK, dtype=torch.int32, device=dev)
    ok = torch.empty(1, dtype=torch.int32, device=dev)
    call("spr_brdf_diffuse_specular", NK, D, B, ptr(brdf), ptr(a), ptr(c),
         ptr(m), ptr(ok), stream_ptr())
    return (a, c, m) if int(ok.item()) else None


def diffuse_specular_plan(vis, patches_center, patches_area, form_factors,
                          p2o, air_attenuation, patch_to_wall_ids, sources,
                          scattering, scattering_index, n_directions, n_bins,
                          speed_of_sound, time_resolution, keep_adjoint=False):
    """GroupedPlan of the diffuse + specular form (see sparrow_hip.h), or
    None if some BRDF row is not a constant plus one specular entry.
    ``keep_adjoint`` also keeps what the adjoint needs (``plan.adjoint``)."""
    c = _f64(patches_center)
    dev = c.device
    P = c.shape[0]
    D, B = int(n_directions), int(n_bins)
    DP = int(_hip.lib().spr_padded_directions(D))
    if DP == 0:
        raise ValueError(f"{D} directions unsupported (max 64)")
    if scattering is None:
        S, NK, bidx, sd = 1, 1, None, None
    else:
        sd = _f64(sources)
        S = sd.shape[1]
        bidx = _i64(scattering_index)
        NK = int(np.shape(scattering)[0]) * S
    tables = diffuse_specular_tables(scattering, D, B, NK)
    if tables is None:
        return None
    ta, tc, spec_dir = tables
    row_ptr, src = ExchangePlan.rows(vis)
    Vd = src.numel()
    max_row = int((row_ptr[1:] - row_ptr[:-1]).max().item()) if P else 0
    if max_row > GroupedPlan.MAX_ROW:
        raise ValueError("destination rows too long for the grouped plan")
    src = src.clone() if Vd else torch.empty(1, dtype=torch.int32,
                                             device=dev)
    w = _i64(patch_to_wall_ids)
    key = torch.empty(max(Vd, 1), dtype=torch.int32, device=dev)
    grp_ptr = torch.empty(P + 1, dtype=torch.int64, device=dev)
    call("spr_plan_ds_sort", P, D, max_row, ptr(row_ptr), ptr(src), ptr(c),
         ptr(w), S, ptr(bidx), ptr(sd), ptr(spec_dir), ptr(key),
         ptr(grp_ptr), stream_ptr())
    G = int(grp_ptr[P].item())
    grp_start = torch.empty(G + 1, dtype=torch.int64, device=dev)
    grp_key = torch.empty(max(G, 1), dtype=torch.int32, device=dev)[:G]
    rowid = torch.empty(max(Vd, 1), dtype=torch.int32, device=dev)
    lag = torch.empty(max(Vd, 1), dtype=torch.int32, device=dev)
    scal = torch.empty((B, max(Vd, 1)), dtype=torch.float64, device=dev)
    dscal = torch.empty((B, max(Vd, 1)), dtype=torch.float64, device=dev)
    rtab = torch.empty((D + 2, B, DP), dtype=torch.float64, device=dev)
    krow = wgt = None
    if keep_adjoint:
        krow = torch.empty(max(Vd, 1), dtype=torch.int32, device=dev)
        wgt = torch.empty((B, max(Vd, 1)), dtype=torch.float64, device=dev)
    call("spr_plan_ds_fill", P, D, B, float(speed_of_sound),
         float(time_resolution), Vd, ptr(row_ptr), ptr(src), ptr(key),
         ptr(grp_ptr), ptr(c), ptr(_f64(patches_area)),
         ptr(_f64(form_factors)), ptr(_i64(p2o)), ptr(_f64(air_attenuation)),
         ptr(w), S, ptr(bidx), ptr(sd), ptr(ta), ptr(tc), ptr(grp_start),
         ptr(grp_key), ptr(rowid), ptr(lag), ptr(scal), ptr(dscal),
         ptr(rtab), ptr(krow), ptr(wgt), stream_ptr())
    plan = GroupedPlan(P, D, B, row_ptr, grp_ptr, grp_start, grp_key,
                       rowid[:Vd], lag[:Vd], scal[:, :Vd].contiguous(), rtab,
                       dscal[:, :Vd].contiguous())
    plan.entry_key = key[:Vd]     # specular direction of each entry (D: none)
    if keep_adjoint:
        # what the adjoint needs besides the plan: entry keys (specular
        # direction, D for none), BRDF rows, pair factors, the tables
        plan.adjoint = dict(key=key[:Vd], krow=krow[:Vd],
                            wgt=wgt[:, :Vd].contiguous(), spec_dir=spec_dir,
                            max_row=max_row, S=S)
    return plan


class StatePool:
    """Device buffers that successive ExchangeStates (the chunks of a
    batched run) take their histograms from, instead of allocating their
    own.  ``begin(sizes)`` makes buffer ``k`` hold at least ``sizes[k]``
    float32 and rewinds; the ``k``-th ``take`` after it returns a view of
    buffer ``k``.  A later ``begin`` reuses the memory, so only one state
    built from the pool is live at a time.  Separate buffers (one per
    histogram) rather than one block: a device heap that earlier work left
    fragmented can still place pieces of this size."""

    def __init__(self, dev=None):
        self.dev = dev or device()
        self.bufs = []
        self.next = 0

    def begin(self, sizes):
        """Rewind; when a buffer must grow, every buffer is freed (and the
        allocator's cache emptied) before any is allocated again, at the
        larger of its old and new size, so old and new never coexist."""
        sizes = [max(int(n), 1) for n in sizes]
        old = [b.numel() for b in self.bufs]
        if len(old) < len(sizes) or any(n > m for n, m in zip(sizes, old)):
            keep = [max(n, old[k] if k < len(old) else 0)
                    for k, n in enumerate(sizes)] + old[len(sizes):]
            self.bufs = []
            torch.cuda.empty_cache()
            self.bufs = [torch.empty(n, dtype=torch.float32, device=self.dev)
                         for n in keep]
        self.next = 0

    def take(self, n):
        """A float32 view of ``n`` elements of the next buffer."""
        k = self.next
        if k >= len(self.bufs) or n > self.bufs[k].numel():
            raise RuntimeError(
                "StatePool: take %d of %d elements from buffer %d of %d"
                % (n, self.bufs[k].numel() if k < len(self.bufs) else 0, k,
                   len(self.bufs)))
        self.next += 1
        return self.bufs[k][:int(n)]


def gather_groups(plan, shape, prev, gsum):
    """Pass 1 of an order on a GroupedPlan: the per-group sums of ``prev``
    ([P, D, B, T] float32) into ``gsum`` ([G, B, T] float32), by the plan's
    destination tiles when it has them (the same sums, bit for bit)."""
    P, D, B, T = shape
    if plan.diffuse_specular and plan.tiles is not None:
        tp = plan.tiles
        call("spr_exchange_gather_tiled_ds", tp.P, tp.JT, B, T, tp.M,
             ptr(tp.tptr), ptr(tp.gstart), ptr(tp.gdiff), ptr(tp.mspan),
             ptr(tp.mlag), ptr(tp.mw), ptr(tp.mdw), ptr(prev), ptr(gsum),
             stream_ptr())
        return
    if plan.diffuse_specular:
        call("spr_exchange_gather_ds", P, B, T, plan.Vd,
             ptr(plan.row_ptr), ptr(plan.grp_ptr), ptr(plan.rowid),
             ptr(plan.lag), ptr(plan.scal), ptr(plan.dscal), ptr(prev),
             ptr(gsum), stream_ptr())
        return
    if plan.tiles is not None:
        tp = plan.tiles
        if tp.JT == 16:
            call("spr_exchange_gather_tiled_shape", tp.P, tp.JT, tp.R,
                 B, T, tp.M, ptr(tp.tptr), ptr(tp.gstart),
                 ptr(tp.mspan), ptr(tp.mlag), ptr(tp.mw), ptr(prev),
                 ptr(gsum), stream_ptr())
            return
        call("spr_exchange_gather_tiled", tp.P, tp.JT, B, T, tp.M,
             ptr(tp.tptr), ptr(tp.gstart), ptr(tp.mrow),
             ptr(tp.mlag), ptr(tp.mw), ptr(prev), ptr(gsum), stream_ptr())
        return
    call("spr_exchange_gather_grouped", P, B, T, plan.Vd,
         ptr(plan.row_ptr), ptr(plan.grp_ptr), ptr(plan.rowid),
         ptr(plan.lag), ptr(plan.scal), ptr(prev), ptr(gsum), stream_ptr())


class ExchangeState:
    """Order histograms and the compensated running total of one source.

    ``prev``/``cur`` are float32 [P,D,B,T]; the total is ``hi + lo``.
    With ``pool`` (a StatePool, rewound by its owner) the buffers, and the
    grouped plan's workspace, are slices of the pool's buffer.
    """

    def __init__(self, P, D, B, T, dev=None, pool=None):
        dev = dev or device()
        shape = (P, D, B, T)
        self.shape = shape
        self.pool = pool
        n = P * D * B * T

        def buf(name):
            if pool is not None:
                return pool.take(n).view(shape)
            return torch.empty(shape, dtype=torch.float32, device=dev)
        self.prev, self.cur = buf("prev"), buf("cur")
        self.hi, self.lo = buf("hi"), buf("lo")
        self.err = torch.zeros(1, dtype=torch.int32, device=dev)
        self.gsum = None  # per-group sums of the grouped plan (workspace)
        # grouped steps add orders to the running total in pairs; True while
        # the last order (in prev after the swap) is not yet in it
        self.pending = False

    def init(self, energy_0_directivity, distance_0, speed_of_sound,
             time_resolution):
        """Order 0 (RadiosityFast.py:1037-1070)."""
        e0d = _f64(energy_0_directivity)
        d0 = _f64(distance_0)
        P, D, B, T = self.shape
        self.err.zero_()
        self.pending = False
        call("spr_exchange_init", P, D, B, T, ptr(e0d), ptr(d0),
             float(speed_of_sound), float(time_resolution), ptr(self.prev),
             ptr(self.hi), ptr(self.lo), ptr(self.err), stream_ptr())

    def step(self, plan):
        """One reflection order on an ExchangePlan or GroupedPlan; swaps
        prev and cur."""
        P, D, B, T = self.shape
        if isinstance(plan, GroupedPlan):
            self.gather(plan)
            self.spread(plan)
        else:
            self.flush()
            call("spr_exchange_step", P, D, B, T, ptr(plan.row_ptr),
                 ptr(plan.src), ptr(plan.lag), ptr(plan.dir), ptr(plan.coef),
                 ptr(self.prev), ptr(self.cur), ptr(self.hi), ptr(self.lo),
                 stream_ptr())
        self.prev, self.cur = self.cur, self.prev

    def _gsum(self, plan):
        P, D, B, T = self.shape
        G = plan.n_groups
        if self.gsum is None or self.gsum.shape[0] < max(G, 1):
            if self.pool is not None:
                self.gsum = self.pool.take(max(G, 1) * B * T).view(
                    max(G, 1), B, T)
            else:
                self.gsum = torch.empty((max(G, 1), B, T),
                                        dtype=torch.float32,
                                        device=self.prev.device)
        return self.gsum

    def gather(self, plan):
        """Grouped plan, pass 1 of an order (per-group sums)."""
        gather_groups(plan, self.shape, self.prev, self._gsum(plan))

    def spread(self, plan):
        """Grouped plan, pass 2 of an order (directions; the running total
        every other order, for this order and the previous one); does not
        swap prev and cur."""
        P, D, B, T = self.shape
        if plan.diffuse_specular:
            pair = self.pending
            call("spr_exchange_spread_ds", P, D, B, T, ptr(plan.grp_ptr),
                 ptr(plan.grp_key), ptr(self._gsum(plan)),
                 ptr(self.prev) if pair else None, ptr(self.cur),
                 ptr(self.hi) if pair else None,
                 ptr(self.lo) if pair else None, stream_ptr())
        elif self.pending:
            call("spr_exchange_spread_grouped_pair", P, D, B, T,
                 ptr(plan.grp_ptr), ptr(plan.grp_key), ptr(plan.rtab),
                 ptr(self._gsum(plan)), ptr(self.prev), ptr(self.cur),
                 ptr(self.hi), ptr(self.lo), stream_ptr())
        else:
            call("spr_exchange_spread_grouped", P, D, B, T,
                 ptr(plan.grp_ptr), ptr(plan.grp_key), ptr(plan.rtab),
                 ptr(self._gsum(plan)), ptr(self.cur), None, None,
                 stream_ptr())
        self.pending = not self.pending

    def flush(self):
        """Add an order still left out of the running total (the last
        one, in prev after the swap)."""
        if self.pending:
            call("spr_exchange_total_add", self.hi.numel(), ptr(self.prev),
                 ptr(self.hi), ptr(self.lo), stream_ptr())
            self.pending = False

    def finalize(self, out=None):
        """float32 total (hi + lo); ``out`` may be ``self.hi``."""
        self.flush()
        out = torch.empty_like(self.hi) if out is None else out
        call("spr_exchange_finalize", self.hi.numel(), ptr(self.hi),
             ptr(self.lo), ptr(out), stream_ptr())
        return out

    def check(self):
        """Raise IndexError if a source delay fell beyond the histogram
        (the reference raises IndexError at RadiosityFast.py:1069)."""
        if int(self.err.item()):
            raise IndexError(
                "source-to-patch delay exceeds the histogram length")
