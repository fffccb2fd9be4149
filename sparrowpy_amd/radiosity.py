"""``DirectionalRadiosityFast`` on MI355X.

Same methods, arguments, call order and output shapes as the reference
class (``sparrowpy/classes/RadiosityFast.py:15-968``); every step of the
path runs in the HIP kernels of ``libsparrow_hip.so``:

=====================================  ======================================
reference (RadiosityFast.py)           here
=====================================  ======================================
``bake_geometry`` :369-433             spr_p2p_visibility, spr_visible_pairs,
                                       spr_form_factors, spr_direction_map
``init_source_energy`` :436-522        spr_point_visibility,
                                       spr_source_energy
``calculate_energy_exchange`` :524     spr_plan_from_geometry,
                                       spr_exchange_init/step/finalize
``collect_energy_receiver_*`` :570     spr_point_visibility,
                                       spr_receiver_energy,
                                       spr_point_direction_index, spr_collect
=====================================  ======================================

Large state stays on the GPU: the energy histograms (float32 [P,D,B,T]) and
the per-pair exchange table.  The directional form factor F~ is kept as the
exchange table (coefficients per visible pair) rather than the reference's
dense [P,P,D,B] array; ``_form_factors_tilde`` materialises the dense array
on request.  Attributes the reference exposes (``_visibility_matrix``,
``_form_factors``, ``_energy_exchange_etc`` ...) return numpy copies.
"""
from __future__ import annotations

import numpy as np
import torch

from . import geometry as geo
from . import ops
from .coordinates import Coordinates, FrequencyData, TimeData

__all__ = ["DirectionalRadiosityFast"]


def _cartesian(obj):
    if hasattr(obj, "cartesian"):
        return np.asarray(obj.cartesian, dtype=float).reshape(-1, 3)
    return np.asarray(obj, dtype=float).reshape(-1, 3)


def _to_numpy(t):
    if t is None:
        return None
    if isinstance(t, torch.Tensor):
        return t.detach().cpu().numpy()
    return t


class DirectionalRadiosityFast:
    """Directional radiosity (reference ``RadiosityFast.py:15``)."""

    def __init__(self, walls_points, walls_normal, walls_up_vector,
                 patches_points, n_patches, patch_to_wall_ids):
        self._walls_points = np.atleast_3d(np.asarray(walls_points, float))
        self._walls_normal = np.atleast_2d(np.asarray(walls_normal, float))
        self._walls_up_vector = np.atleast_2d(
            np.asarray(walls_up_vector, float))
        self._patches_points = np.atleast_3d(np.asarray(patches_points, float))
        self._n_patches = int(n_patches)
        self._patch_to_wall_ids = np.atleast_1d(
            np.array(patch_to_wall_ids, dtype=np.int64))
        self._frequencies = None
        self._brdf = None
        self._brdf_index = None
        self._brdf_incoming_directions = None
        self._brdf_outgoing_directions = None
        self._air_attenuation = None
        self._speed_of_sound = None
        self._etc_time_resolution = None
        self._etc_duration = None
        self._source = None
        self._dev = {}
        self._bake = None
        self._etc = None
        self.check()

    # -- construction ---------------------------------------------------
    @classmethod
    def from_polygon(cls, polygon_list, patch_size):
        """Subdivide the walls (reference ``RadiosityFast.py:340-367``)."""
        walls_points = np.array([p.pts for p in polygon_list])
        walls_normal = np.array([p.normal for p in polygon_list])
        walls_up_vector = np.array([p.up_vector for p in polygon_list])
        pts, _, n, ids = geo.process_patches(walls_points, walls_normal,
                                             patch_size, len(polygon_list))
        return cls(walls_points, walls_normal, walls_up_vector, pts, n, ids)

    def check(self):
        """Shape validation (reference ``RadiosityFast.py:211-337``)."""
        n_walls = self._walls_points.shape[0]
        if self._walls_points.shape[2] != 3:
            raise ValueError(
                "Walls need to be of shape (n_walls, n_points, 3)")
        if self._walls_up_vector.shape != (n_walls, 3):
            raise ValueError(
                "Up vector of walls need to be of shape (n_walls, 3)")
        if self._walls_normal.shape != (n_walls, 3):
            raise ValueError(
                "Normal of walls need to be of shape (n_walls, 3)")
        if (self._patches_points.shape[0] != self.n_patches
                or self._patches_points.shape[2] != 3):
            raise ValueError(
                "Patches need to be of shape (n_patches, n_points, 3)")
        if self._patch_to_wall_ids.shape != (self.n_patches,):
            raise ValueError(
                "patch_to_wall_ids need to be of shape (n_patches,)")
        ids = set(self._patch_to_wall_ids.tolist())
        if ids != set(range(n_walls)):
            raise ValueError(
                "patch_to_wall_ids does contain other ids than "
                "range(n_walls)")

    # -- materials --------------------------------------------------------
    def set_wall_brdf(self, wall_indexes, brdf, incoming_directions,
                      outgoing_directions):
        """Reference ``RadiosityFast.py:766-816``: rotate the direction sets
        into each wall's frame and store ``brdf.freq * pi``."""
        inc = _cartesian(incoming_directions)
        out = _cartesian(outgoing_directions)
        if not (inc[:, 2] >= 0).all():
            raise AssertionError("Sources must be in the positive half space")
        if not (out[:, 2] >= 0).all():
            raise AssertionError(
                "Receivers must be in the positive half space")
        self._check_set_frequency(brdf.frequencies)
        if self._brdf_incoming_directions is None:
            self._brdf_incoming_directions = np.empty(self.n_walls, object)
            self._brdf_outgoing_directions = np.empty(self.n_walls, object)
            self._brdf_index = np.full(self.n_walls, -1, dtype=np.int64)
            self._brdf = []
        for i in np.atleast_1d(wall_indexes):
            self._brdf_incoming_directions[i] = Coordinates.from_cartesian(
                geo.rotate_coords_to_normal(self._walls_normal[i],
                                            self._walls_up_vector[i], inc))
            self._brdf_outgoing_directions[i] = Coordinates.from_cartesian(
                geo.rotate_coords_to_normal(self._walls_normal[i],
                                            self._walls_up_vector[i], out))
        self._brdf.append(np.real(np.asarray(brdf.freq)) * np.pi)
        self._brdf_index[np.atleast_1d(wall_indexes)] = len(self._brdf) - 1

    def set_air_attenuation(self, air_attenuation):
        """Air attenuation in Np/m (reference ``RadiosityFast.py:754-764``)."""
        self._check_set_frequency(air_attenuation.frequencies)
        self._air_attenuation = np.atleast_1d(
            np.asarray(air_attenuation.freq, float).squeeze())

    def _check_set_frequency(self, frequencies):
        frequencies = np.atleast_1d(np.asarray(frequencies, float))
        if self._frequencies is None:
            self._frequencies = frequencies
        else:
            assert self._frequencies.size == frequencies.size, \
                "Number of frequency bins do not match"
            assert (self._frequencies == frequencies).all(), \
                "Frequencies do not match"

    def _brdf_tables(self):
        """Stacked BRDF tables and per-wall direction arrays (host)."""
        sources = np.array([s.cartesian for s in
                            self._brdf_incoming_directions])
        receivers = np.array([s.cartesian for s in
                              self._brdf_outgoing_directions])
        return (np.array(self._brdf), np.array(self._brdf_index), sources,
                receivers)

    # -- device geometry --------------------------------------------------
    def _geometry(self):
        if "centers" not in self._dev:
            self._dev["centers"] = ops.as_dev(self.patches_center,
                                              torch.float64)
            self._dev["points"] = ops.as_dev(self._patches_points,
                                             torch.float64)
            self._dev["normals"] = ops.as_dev(self.patches_normal,
                                              torch.float64)
            self._dev["areas"] = ops.as_dev(self.patches_area, torch.float64)
            self._dev["wall_ids"] = ops.as_dev(self._patch_to_wall_ids,
                                               torch.int64)
            self._dev["walls_points"] = ops.as_dev(self._walls_points,
                                                   torch.float64)
            self._dev["walls_normal"] = ops.as_dev(self._walls_normal,
                                                   torch.float64)
        return self._dev

    # -- bake -------------------------------------------------------------
    def bake_geometry(self):
        """Visibility, form factors and direction maps
        (reference ``RadiosityFast.py:369-433``)."""
        g = self._geometry()
        vis = ops.check_patch2patch_visibility(g["centers"], g["normals"],
                                               g["points"])
        pairs = ops.visible_pairs(vis)
        ff = ops.patch2patch_ff_universal(g["points"], g["normals"],
                                          g["areas"], pairs)
        if self._brdf_incoming_directions is not None:
            brdf, bidx, sources, receivers = self._brdf_tables()
            p2o = ops.direction_map(vis, g["centers"], receivers,
                                    g["wall_ids"])
            D = receivers.shape[1]
        else:
            brdf = bidx = sources = receivers = None
            P = self.n_patches
            p2o = torch.zeros((P, P), dtype=torch.int64,
                              device=g["centers"].device)
            D = 1
        n_bins = 1 if self._frequencies is None else self.n_bins
        self._bake = dict(
            vis=vis, pairs=pairs, ff=ff, p2o=p2o, D=D, n_bins=n_bins,
            air=None if self._air_attenuation is None else
            self._air_attenuation.copy(),
            brdf=brdf, brdf_index=bidx, sources=sources, receivers=receivers)
        self._plan = None

    # -- source -------------------------------------------------------------
    def init_source_energy(self, source):
        """Source-to-patch energy (reference ``RadiosityFast.py:436-522``).

        ``source``: a point (Coordinates of cshape (1,), or (3,) array).
        """
        pos = _cartesian(source)
        if pos.shape[0] != 1:
            raise ValueError('just one source position is allowed.')
        self._source = source
        self._source_position = pos[0]
        if self._brdf_incoming_directions is None:
            freqs = np.array([0.]) if self._frequencies is None else \
                self._frequencies
            self.set_wall_brdf(np.arange(self.n_walls),
                               FrequencyData(np.ones_like(freqs), freqs),
                               Coordinates(0, 0, 1, weights=1),
                               Coordinates(0, 0, 1, weights=1))
            self._frequencies = freqs
        if self._air_attenuation is None:
            freqs = np.array([0.]) if self._frequencies is None else \
                self._frequencies
            self.set_air_attenuation(FrequencyData(np.zeros_like(freqs),
                                                   freqs))
            self._frequencies = freqs
        g = self._geometry()
        brdf, bidx, sources, receivers = self._brdf_tables()
        vis = ops.check_point2patch_visibility(
            pos, g["centers"], g["walls_normal"], g["walls_points"])[0]
        e0, e0d, d0 = ops.source_energy(
            pos[0], g["centers"], g["points"], vis, self._air_attenuation,
            self.n_bins, g["wall_ids"], sources, brdf, bidx,
            receivers.shape[1])
        self._dev["source_visibility"] = vis
        self._dev["energy_0"] = e0
        self._dev["energy_init_source"] = e0d
        self._dev["distance_patches_to_source"] = d0

    # -- exchange -----------------------------------------------------------
    # Use the grouped plan when groups are at most this fraction of pairs
    # (each group costs D FMAs per bin, each pair one FMA).
    GROUPED_MAX_GROUP_FRACTION = 0.25

    def _exchange_plan(self, speed_of_sound, time_resolution, kind=None):
        """Plan of the exchange (cached per speed of sound and bin width).

        ``kind``: "grouped", "generic" or None (automatic: grouped when the
        sources of a destination share few BRDF rows).
        """
        key = (float(speed_of_sound), float(time_resolution), kind)
        if getattr(self, "_plan", None) is not None and self._plan[0] == key:
            return self._plan[1]
        b = self._bake
        g = self._geometry()
        args = (b["vis"], g["centers"], g["areas"], b["ff"], b["p2o"],
                b["air"], g["wall_ids"], b["sources"], b["brdf"],
                b["brdf_index"], b["D"], b["n_bins"], speed_of_sound,
                time_resolution)
        plan = None
        if kind in (None, "grouped"):
            try:
                plan = ops.GroupedPlan.from_geometry(*args)
            except ValueError:
                if kind == "grouped":
                    raise
            if (plan is not None and kind is None and plan.Vd and
                    plan.n_groups > self.GROUPED_MAX_GROUP_FRACTION * plan.Vd):
                plan = None
        if plan is None:
            plan = ops.ExchangePlan.from_geometry(*args)
        self._plan = (key, plan)
        return plan

    def calculate_energy_exchange(self, speed_of_sound, etc_time_resolution,
                                  etc_duration, max_reflection_order=-1,
                                  recalculate=False):
        """Reflection-order energy exchange
        (reference ``RadiosityFast.py:524-568``)."""
        n_samples = int(etc_duration / etc_time_resolution)
        if self._etc is None or recalculate:
            e0d = self._dev["energy_init_source"]
            d0 = self._dev["distance_patches_to_source"]
            P, D, B = e0d.shape
            if max_reflection_order >= 1 and self._bake is None:
                raise RuntimeError("bake_geometry() must run before an "
                                   "exchange with reflections")
            if max_reflection_order >= 1 and (self._bake["D"], self._bake[
                    "n_bins"]) != (D, B):
                raise ValueError(
                    "baked form factors and source energy disagree in "
                    "directions or bands")
            state = ops.ExchangeState(P, D, B, n_samples)
            state.init(e0d, d0, speed_of_sound, etc_time_resolution)
            if max_reflection_order >= 1:
                plan = self._exchange_plan(speed_of_sound,
                                           etc_time_resolution)
                for _ in range(int(max_reflection_order)):
                    state.step(plan)
            etc = state.finalize(out=state.hi)
            state.check()
            del state
            self._etc = etc
        self._etc_time_resolution = float(etc_time_resolution)
        self._speed_of_sound = float(speed_of_sound)
        self._etc_duration = float(etc_duration)

    # -- receivers ------------------------------------------------------------
    def _receiver_terms(self, receivers):
        g = self._geometry()
        pts = _cartesian(receivers)
        vis = ops.check_point2patch_visibility(
            pts, g["centers"], g["walls_normal"], g["walls_points"])
        factor = ops.receiver_energy(pts, g["points"], vis)
        _, _, _, out_dirs = self._brdf_tables()
        idx = ops.point_direction_index(pts, g["centers"], out_dirs,
                                        g["wall_ids"])
        return pts, vis, factor, idx

    def collect_energy_receiver_patchwise(self, receivers):
        """Per-patch histograms at the receivers, (R, P, B, T)
        (reference ``RadiosityFast.py:660-684``)."""
        if not hasattr(receivers, "cartesian"):
            raise ValueError(
                "Receiver positions must be of type pf.Coordinates")
        pts, _, factor, idx = self._receiver_terms(receivers)
        g = self._geometry()
        out = ops.collect(self._etc, pts, g["centers"], idx, factor,
                          self._air_attenuation, self._speed_of_sound,
                          self._etc_time_resolution, patchwise=True)
        data = out.double().cpu().numpy()
        times = np.arange(data.shape[-1]) * self._etc_time_resolution
        return TimeData(data, times)

    def collect_energy_receiver_mono(self, receivers, direct_sound=False):
        """Histograms at the receivers, (R, B, T)
        (reference ``RadiosityFast.py:570-602``)."""
        if not isinstance(direct_sound, bool):
            raise ValueError("direct_sound must be of type boolean")
        if not hasattr(receivers, "cartesian"):
            raise ValueError(
                "Receiver positions must be of type pf.Coordinates")
        pts, _, factor, idx = self._receiver_terms(receivers)
        g = self._geometry()
        out = ops.collect(self._etc, pts, g["centers"], idx, factor,
                          self._air_attenuation, self._speed_of_sound,
                          self._etc_time_resolution, patchwise=False)
        data = out.double().cpu().numpy()
        if direct_sound:
            ds, n = self.calculate_direct_sound(receivers)
            data[np.arange(len(n)), :, n] += ds
        times = np.arange(data.shape[-1]) * self._etc_time_resolution
        return TimeData(data, times)

    def calculate_direct_sound(self, receivers):
        """Direct sound 1/(4 pi r^2) exp(-m r) at bin int(r/c/dt)
        (reference ``RadiosityFast.py:605-657``)."""
        if not hasattr(receivers, "cartesian"):
            raise ValueError(
                "Receiver positions must be of type pf.Coordinates")
        pts = _cartesian(receivers)
        src = self._source_position
        d = pts - src
        r = np.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2 + d[:, 2] ** 2)
        direct = np.ones((pts.shape[0], self.n_bins), dtype=float)
        direct *= (1 / (4 * np.pi * r ** 2))[:, np.newaxis]
        if self._air_attenuation is not None:
            for i in range(self.n_bins):
                direct[:, i] *= np.exp(-self._air_attenuation[i] * r)
        n_delay = np.array(r / self.speed_of_sound /
                           self._etc_time_resolution, dtype=int)
        return direct, n_delay

    # -- reference attributes (host copies) -----------------------------------
    @property
    def _visibility_matrix(self):
        return None if self._bake is None else \
            _to_numpy(self._bake["vis"]).astype(bool)

    @property
    def _visible_patches(self):
        return None if self._bake is None else _to_numpy(self._bake["pairs"])

    @property
    def _form_factors(self):
        return None if self._bake is None else _to_numpy(self._bake["ff"])

    @property
    def _patch_2_brdf_outgoing_index(self):
        return None if self._bake is None else _to_numpy(self._bake["p2o"])

    @property
    def _form_factors_tilde(self):
        if self._bake is None:
            return None
        b = self._bake
        g = self._geometry()
        return _to_numpy(ops.form_factors_with_directivity_dim(
            b["vis"], b["ff"], b["n_bins"], g["centers"], g["areas"],
            b["air"], g["wall_ids"], b["brdf"], b["brdf_index"],
            b["sources"], b["receivers"]))

    @property
    def _source_visibility(self):
        return _to_numpy(self._dev.get("source_visibility")).astype(bool)

    @property
    def _energy_init_source(self):
        return _to_numpy(self._dev.get("energy_init_source"))

    @property
    def _distance_patches_to_source(self):
        return _to_numpy(self._dev.get("distance_patches_to_source"))

    @property
    def _energy_exchange_etc(self):
        return None if self._etc is None else _to_numpy(self._etc)

    @property
    def energy_exchange_etc_device(self):
        """The total histogram as a device tensor, float32 [P,D,B,T]."""
        return self._etc

    # -- reference properties -------------------------------------------------
    @property
    def n_bins(self):
        return None if self._frequencies is None else \
            self._frequencies.shape[0]

    @property
    def n_walls(self):
        return self._walls_points.shape[0]

    @property
    def n_patches(self):
        return self._n_patches

    @property
    def form_factors(self):
        return self._form_factors

    @property
    def visibility_matrix(self):
        return self._visibility_matrix

    @property
    def walls_area(self):
        return geo.calculate_area(self._walls_points)

    @property
    def walls_points(self):
        return self._walls_points

    @property
    def walls_normal(self):
        return self._walls_normal

    @property
    def walls_center(self):
        return geo.calculate_center(self._walls_points)

    @property
    def walls_up_vector(self):
        return self._walls_up_vector

    @property
    def patches_area(self):
        return geo.calculate_area(self._patches_points)

    @property
    def patches_center(self):
        return geo.calculate_center(self._patches_points)

    @property
    def patches_size(self):
        return geo.calculate_size(self._patches_points)

    @property
    def patches_points(self):
        return self._patches_points

    @property
    def patches_normal(self):
        return self._walls_normal[self._patch_to_wall_ids]

    @property
    def speed_of_sound(self):
        return self._speed_of_sound
