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


def _direction_sets(sets):
    """Per-wall direction sets (None, or Coordinates / [n,3] arrays / None
    per wall) -> object array of Coordinates (None where unset)."""
    if sets is None:
        return None
    out = np.empty(len(sets), object)
    for i, c in enumerate(sets):
        if c is None or (isinstance(c, str) and c == "None"):
            continue
        out[i] = Coordinates.from_cartesian(_cartesian(c))
    return out


def _as_array(v):
    return None if v is None else np.asarray(v)


def _as_state_value(v):
    return "None" if v is None else v


def _state_equal(a, b):
    """Deep equality of two state values (dicts, lists, arrays, scalars)."""
    if isinstance(a, dict) or isinstance(b, dict):
        return (isinstance(a, dict) and isinstance(b, dict)
                and a.keys() == b.keys()
                and all(_state_equal(a[k], b[k]) for k in a))
    if isinstance(a, (list, tuple)) or isinstance(b, (list, tuple)):
        return (isinstance(a, (list, tuple)) and isinstance(b, (list, tuple))
                and len(a) == len(b)
                and all(_state_equal(x, y) for x, y in zip(a, b)))
    if isinstance(a, str) or isinstance(b, str):
        return a == b
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype.kind == b.dtype.kind and \
        np.array_equal(a, b)


class DirectionalRadiosityFast:
    """Directional radiosity (reference ``RadiosityFast.py:15``)."""

    def __init__(self, walls_points, walls_normal, walls_up_vector,
                 patches_points, n_patches, patch_to_wall_ids,
                 visibility_matrix=None, visible_patches=None,
                 form_factors=None, form_factors_tilde=None,
                 frequencies=None, brdf=None, brdf_index=None,
                 brdf_incoming_directions=None,
                 brdf_outgoing_directions=None,
                 patch_2_brdf_outgoing_index=None, air_attenuation=None,
                 speed_of_sound=None, etc_time_resolution=None,
                 etc_duration=None, distance_patches_to_source=None,
                 energy_init_source=None, energy_exchange_etc=None):
        """Reference ``RadiosityFast.py:27-209``: geometry plus, optionally,
        a saved state (the keys of :meth:`to_dict`).  Saved bake, source
        and histogram arrays are uploaded to the GPU."""
        self._walls_points = np.atleast_3d(np.asarray(walls_points, float))
        self._walls_normal = np.atleast_2d(np.asarray(walls_normal, float))
        self._walls_up_vector = np.atleast_2d(
            np.asarray(walls_up_vector, float))
        self._patches_points = np.atleast_3d(np.asarray(patches_points, float))
        self._n_patches = int(n_patches)
        self._patch_to_wall_ids = np.atleast_1d(
            np.array(patch_to_wall_ids, dtype=np.int64))
        self._frequencies = None if frequencies is None else \
            np.atleast_1d(np.asarray(frequencies, float))
        self._brdf = None if brdf is None else \
            [np.asarray(b, float) for b in brdf]
        self._brdf_index = None if brdf_index is None else \
            np.asarray(brdf_index, dtype=np.int64)
        self._brdf_incoming_directions = _direction_sets(
            brdf_incoming_directions)
        self._brdf_outgoing_directions = _direction_sets(
            brdf_outgoing_directions)
        self._air_attenuation = None if air_attenuation is None else \
            np.atleast_1d(np.asarray(air_attenuation, float))
        self._speed_of_sound = None if speed_of_sound is None else \
            float(speed_of_sound)
        self._etc_time_resolution = None if etc_time_resolution is None \
            else float(etc_time_resolution)
        self._etc_duration = None if etc_duration is None else \
            float(etc_duration)
        self._source = None
        self._dev = {}
        self._bake = None
        self._plan = None
        self._etc = None
        # saved arrays, validated by check() before they are uploaded
        self._form_factors_in = _as_array(form_factors)
        self._form_factors_tilde_in = _as_array(form_factors_tilde)
        self._distance_patches_to_source_in = _as_array(
            distance_patches_to_source)
        self._energy_init_source_in = _as_array(energy_init_source)
        self._energy_exchange_etc_in = _as_array(energy_exchange_etc)
        if (energy_init_source is None) != (
                distance_patches_to_source is None):
            raise ValueError("energy_init_source and "
                             "distance_patches_to_source go together")
        self.check()
        self._form_factors_in = self._form_factors_tilde_in = None
        self._distance_patches_to_source_in = None
        self._energy_init_source_in = self._energy_exchange_etc_in = None
        if visibility_matrix is not None:
            self._restore_bake(visibility_matrix, visible_patches,
                               form_factors, form_factors_tilde,
                               patch_2_brdf_outgoing_index)
        if energy_init_source is not None:
            self._dev["energy_init_source"] = ops.as_dev(
                np.asarray(energy_init_source, float), torch.float64)
            self._dev["distance_patches_to_source"] = ops.as_dev(
                np.asarray(distance_patches_to_source, float),
                torch.float64)
        if energy_exchange_etc is not None:
            self._etc = ops.as_dev(np.asarray(energy_exchange_etc),
                                   torch.float32)

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

    def _n_outgoing_directions(self):
        """D of the set BRDF direction sets (1 without BRDFs)."""
        for c in (self._brdf_outgoing_directions
                  if self._brdf_outgoing_directions is not None else []):
            if c is not None:
                return int(c.cartesian.shape[0])
        return 1

    def check(self):
        """Input validation (reference ``RadiosityFast.py:211-337``).

        Raises the reference's ValueErrors for the geometry, frequencies,
        form factors, BRDF index and direction sets, F~, air attenuation,
        speed of sound, time resolution, duration and the saved source
        distances, source energies and histograms.  Walls without a BRDF
        (``set_wall_brdf`` on a subset of walls) keep ``None`` direction
        sets, which the reference's type check would refuse; every set
        wall must have the same number of directions.  The BRDF tables must
        match the direction sets and bands, and BRDF indices must name a
        table, since the GPU kernels index them without bounds checks.
        """
        n_walls = self._walls_points.shape[0]
        if self._walls_points.ndim != 3 or self._walls_points.shape[2] != 3:
            raise ValueError(
                "Walls need to be of shape (n_walls, n_points, 3)")
        if self._walls_up_vector.shape != (n_walls, 3):
            raise ValueError(
                "Up vector of walls need to be of shape (n_walls, 3)")
        if self._walls_normal.shape != (n_walls, 3):
            raise ValueError(
                "Normal of walls need to be of shape (n_walls, 3)")
        if (self._patches_points.ndim != 3
                or self._patches_points.shape[0] != self.n_patches
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
        n_bins = 1
        if self._frequencies is not None:
            if self._frequencies.ndim != 1:
                raise ValueError("Frequencies need to be of shape (n_bins,)")
            n_bins = self._frequencies.size
        n_dirs = self._n_outgoing_directions()
        if self._brdf_index is not None:
            if self._brdf_index.ndim != 1 or len(self._brdf_index) != n_walls:
                raise ValueError("brdf_index need to be of shape (n_walls,)")
        for name, sets in (
                ("brdf_incoming_directions", self._brdf_incoming_directions),
                ("brdf_outgoing_directions", self._brdf_outgoing_directions)):
            if sets is None:
                continue
            if len(sets) != n_walls:
                raise ValueError(f"{name} need to be a list of n_walls "
                                 "Coordinates")
            if name == "brdf_outgoing_directions" and any(
                    c is not None and c.cartesian.shape[0] != n_dirs
                    for c in sets):
                raise ValueError(f"{name} need the same number of "
                                 "directions on every wall")
        if self._brdf is not None:
            n_in = None
            if self._brdf_incoming_directions is not None:
                n_in = {c.cartesian.shape[0]
                        for c in self._brdf_incoming_directions
                        if c is not None}
            # the kernels read each table as flat [S, D, B] memory
            sizes = ({k * n_dirs * n_bins for k in n_in} if n_in else None)
            for b in self._brdf:
                if (b.size % (n_dirs * n_bins) != 0 or b.size == 0
                        or (sizes and b.size not in sizes)):
                    raise ValueError(
                        "brdf need to be a list of arrays of shape "
                        "(n_incoming_directions, n_outgoing_directions, "
                        "n_bins)")
            if self._brdf_index is not None and (
                    (self._brdf_index >= len(self._brdf)).any()
                    or (self._brdf_index < -1).any()):
                raise ValueError("brdf_index must name an entry of brdf "
                                 "(or -1 for a wall without a BRDF)")
        if self._form_factors_in is not None:
            if self._form_factors_in.shape != (self.n_patches,
                                               self.n_patches):
                raise ValueError(
                    "form_factors need to be of shape (n_patches, n_patches)")
        if self._form_factors_tilde_in is not None:
            if self._form_factors_tilde_in.shape != (
                    self.n_patches, self.n_patches, n_dirs, n_bins):
                raise ValueError(
                    "form_factors_tilde need to be of shape "
                    "(n_patches, n_patches, n_outgoing_directions, n_bins)")
        if self._air_attenuation is not None:
            if (self._air_attenuation.ndim != 1
                    or self._air_attenuation.shape[0] != n_bins):
                raise ValueError(
                    "Air attenuation need to be of shape (n_bins,)")
        if self._speed_of_sound is not None and self._speed_of_sound <= 0:
            raise ValueError("Speed of sound must be positive and non-zero")
        if (self._etc_time_resolution is not None
                and self._etc_time_resolution <= 0):
            raise ValueError("Time resolution must be positive and non-zero")
        if self._etc_duration is not None and self._etc_duration <= 0:
            raise ValueError("Duration must be positive and non-zero")
        if self._distance_patches_to_source_in is not None:
            if self._distance_patches_to_source_in.shape != (self.n_patches,):
                raise ValueError(
                    "distance_patches_to_source need to be of shape "
                    "(n_patches,)")
        if self._energy_init_source_in is not None:
            if self._energy_init_source_in.shape != (
                    self.n_patches, n_dirs, n_bins):
                raise ValueError(
                    "energy_init_source need to be of shape "
                    "(n_patches, n_outgoing_directions, n_bins)")
        if self._energy_exchange_etc_in is not None:
            if self._etc_duration is None or \
                    self._etc_time_resolution is None:
                raise ValueError(
                    "energy_exchange_etc needs etc_duration and "
                    "etc_time_resolution")
            n_samples = int(self._etc_duration / self._etc_time_resolution)
            if self._energy_exchange_etc_in.shape != (
                    self.n_patches, n_dirs, n_bins, n_samples):
                raise ValueError(
                    "energy_exchange_etc need to be of shape "
                    "(n_patches, n_outgoing_directions, n_bins, n_samples)")

    def _restore_bake(self, vis, pairs, ff, fft_dense, p2o):
        """Upload a saved bake (the reference's visibility matrix, visible
        pair list, form factors, direction map and, optionally, F~).

        Every array the kernels index is checked first: shapes against
        n_patches, the direction count and the bands, pair indices within
        the patches and above the diagonal, direction indices within
        [0, D] (D marks pairs that see each other in neither orientation).
        """
        P = self.n_patches
        vis = np.asarray(vis)
        if vis.shape != (P, P):
            raise ValueError("visibility_matrix must be (n_patches, "
                             "n_patches)")
        vis = vis.astype(bool)
        if ff is None or np.asarray(ff).shape != (P, P):
            raise ValueError(
                "form_factors need to be of shape (n_patches, n_patches)")
        if self._brdf_outgoing_directions is not None:
            brdf, bidx, sources, receivers = self._brdf_tables()
            D = receivers.shape[1]
        else:
            brdf = bidx = sources = receivers = None
            D = 1
        n_bins = 1 if self._frequencies is None else self.n_bins
        if pairs is not None:
            pairs = np.asarray(pairs).reshape(-1, 2) if np.size(pairs) else \
                np.zeros((0, 2), np.int64)
            if ((pairs < 0).any() or (pairs >= P).any()
                    or (pairs[:, 0] >= pairs[:, 1]).any()):
                raise ValueError("visible_patches need to be pairs (i, j) "
                                 "of patch indices with i < j")
            if not vis[pairs[:, 0], pairs[:, 1]].all() or \
                    len(pairs) != int(np.triu(vis, 1).sum()):
                raise ValueError("visible_patches must list the visible "
                                 "pairs of visibility_matrix")
        if p2o is not None:
            p2o = np.asarray(p2o, dtype=np.int64)
            if p2o.shape != (P, P):
                raise ValueError("patch_2_brdf_outgoing_index need to be of "
                                 "shape (n_patches, n_patches)")
            if (p2o < 0).any() or (p2o > D).any():
                raise ValueError("patch_2_brdf_outgoing_index values must be "
                                 "outgoing direction indices (0..D)")
        if fft_dense is not None and np.asarray(fft_dense).shape != (
                P, P, D, n_bins):
            raise ValueError(
                "form_factors_tilde need to be of shape "
                "(n_patches, n_patches, n_outgoing_directions, n_bins)")
        self._bake = dict(
            vis=ops.as_dev(vis, torch.uint8),
            pairs=ops.as_dev(np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
                             if pairs is not None else
                             np.argwhere(np.triu(vis, 1)).astype(np.int32),
                             torch.int32),
            ff=ops.as_dev(np.asarray(ff, float), torch.float64),
            p2o=ops.as_dev(p2o if p2o is not None
                           else np.zeros((P, P), np.int64), torch.int64),
            D=D, n_bins=n_bins,
            air=None if self._air_attenuation is None else
            self._air_attenuation.copy(),
            brdf=brdf, brdf_index=bidx, sources=sources, receivers=receivers,
            fft_dense=None if fft_dense is None else
            np.asarray(fft_dense, float))

    # -- state (reference RadiosityFast.py:828-886) ------------------------
    # Dense F~ [P,P,D,B] float64 is written only up to this size; larger
    # states store "None" for it (it is rebuilt exactly from the bake).
    MAX_DENSE_FFT_BYTES = 1 << 32

    def to_dict(self) -> dict:
        """The state under the reference's keys.

        Arrays stay numpy arrays (the reference converts them to lists for
        pyfar's writer); per-wall direction sets are lists of [n, 3] arrays
        (None for walls without a BRDF); ``energy_exchange_etc`` is the
        float32 total histogram; unset values are the string "None".
        """
        def dirs(sets):
            if sets is None:
                return None
            return [None if c is None else np.asarray(c.cartesian)
                    for c in sets]

        fft = None
        if self._bake is not None:
            b = self._bake
            P = self.n_patches
            if b.get("fft_dense") is not None:
                fft = b["fft_dense"]
            elif P * P * b["D"] * b["n_bins"] * 8 <= self.MAX_DENSE_FFT_BYTES:
                fft = self._form_factors_tilde
        out = {
            "walls_points": self._walls_points,
            "walls_normal": self._walls_normal,
            "walls_up_vector": self._walls_up_vector,
            "patches_points": self._patches_points,
            "n_patches": self._n_patches,
            "patch_to_wall_ids": self._patch_to_wall_ids,
            "visibility_matrix": self._visibility_matrix,
            "visible_patches": self._visible_patches,
            "form_factors": self._form_factors,
            "form_factors_tilde": fft,
            "frequencies": self._frequencies,
            "brdf": self._brdf,
            "brdf_index": self._brdf_index,
            "brdf_incoming_directions": dirs(self._brdf_incoming_directions),
            "brdf_outgoing_directions": dirs(self._brdf_outgoing_directions),
            "patch_2_brdf_outgoing_index": self._patch_2_brdf_outgoing_index,
            "air_attenuation": self._air_attenuation,
            "speed_of_sound": self._speed_of_sound,
            "etc_time_resolution": self._etc_time_resolution,
            "etc_duration": self._etc_duration,
            "distance_patches_to_source": self._distance_patches_to_source,
            "energy_init_source": self._energy_init_source,
            "energy_exchange_etc": self._energy_exchange_etc,
        }
        return {k: _as_state_value(v) for k, v in out.items()}

    @classmethod
    def from_dict(cls, input_dict: dict):
        """Inverse of :meth:`to_dict` (reference ``RadiosityFast.py:880``);
        the string "None" means unset."""
        kw = {k: (None if isinstance(v, str) and v == "None" else v)
              for k, v in input_dict.items()}
        return cls(**kw)

    def __eq__(self, other):
        """Equal states (reference ``RadiosityFast.py:871-876``)."""
        if not isinstance(other, DirectionalRadiosityFast):
            return False
        return _state_equal(self.to_dict(), other.to_dict())

    __hash__ = None

    def write(self, filename, compress=True):
        """Write the state to a numpy ``.npz`` archive (the reference writes
        a pyfar ``.far`` file with the same keys, ``RadiosityFast.py:828``).

        Lists (``brdf``, direction sets) are stored element-wise as
        ``key/<i>`` with ``key/len``; nothing is pickled.
        """
        flat = {}
        for key, value in self.to_dict().items():
            if isinstance(value, list):
                flat[f"{key}/len"] = np.array(len(value))
                for i, v in enumerate(value):
                    flat[f"{key}/{i}"] = np.array("None") if v is None \
                        else np.asarray(v)
            else:
                flat[key] = np.asarray(value)
        save = np.savez_compressed if compress else np.savez
        with open(filename, "wb") as f:
            save(f, **flat)

    @classmethod
    def from_read(cls, filename):
        """Read a state written by :meth:`write`."""
        with np.load(filename, allow_pickle=False) as z:
            data = {k: z[k] for k in z.files}
        out = {}
        for key in [k for k in data if "/" not in k]:
            v = data[key]
            out[key] = str(v) if v.dtype.kind == "U" else v
        lists = {k.split("/")[0] for k in data if k.endswith("/len")}
        for key in lists:
            n = int(data[f"{key}/len"])
            items = []
            for i in range(n):
                v = data[f"{key}/{i}"]
                items.append(None if v.dtype.kind == "U" else v)
            out[key] = items
        for key in ("n_patches",):
            out[key] = int(out[key])
        for key in ("speed_of_sound", "etc_time_resolution",
                    "etc_duration"):
            if not isinstance(out[key], str):
                out[key] = float(out[key])
        return cls.from_dict(out)

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
        """Stacked BRDF tables and per-wall direction arrays (host).

        Every wall needs a BRDF here (the kernels index the tables by
        wall); the reference fails there too, on its unset walls."""
        if (any(s is None for s in self._brdf_incoming_directions)
                or any(s is None for s in self._brdf_outgoing_directions)
                or (np.asarray(self._brdf_index) < 0).any()):
            raise ValueError("set_wall_brdf must cover every wall before "
                             "the bake, the source or the receivers")
        if len({s.cartesian.shape for s in
                self._brdf_incoming_directions}) != 1:
            raise ValueError("brdf_incoming_directions need the same number "
                             "of directions on every wall")
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
            brdf=brdf, brdf_index=bidx, sources=sources, receivers=receivers,
            fft_dense=None)
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
    # Rows per destination tile of the grouped plan's gather (4, 8 or 16; 0
    # gathers row by row); tiles are used when they merge at least
    # TILE_MIN_MERGE entries into one.  16-row tiles stage each merged
    # source row in LDS once for all 16 rows (one wave per row).
    TILE_ROWS = 16
    TILE_MIN_MERGE = 1.5
    # Tiles as (a, b) blocks of neighbouring patches of one wall instead of
    # TILE_ROWS consecutive patch ids (a * b must equal TILE_ROWS); None:
    # consecutive ids; "auto": 4 x 4 blocks for 16 rows, 2 x 4 for 8, 2 x 2
    # for 4.
    TILE_BLOCK = "auto"
    # 16-row tiles: bins per lane of the gather (8: two workgroups per CU;
    # or 16)
    TILE_R = 8
    # tile blocks padded to whole blocks at wall edges (empty slots), or
    # run together
    TILE_PAD = True

    _AUTO_BLOCKS = {16: (4, 4), 8: (2, 4), 4: (2, 2)}

    def _tile_order(self):
        """Destination order of the tiles (None: consecutive patch ids)."""
        blk = self.TILE_BLOCK
        if isinstance(blk, str):
            if blk != "auto":
                raise ValueError('TILE_BLOCK must be None, "auto" or (a, b)')
            blk = self._AUTO_BLOCKS.get(self.TILE_ROWS)
        if blk is None:
            return None
        a, bb = blk
        if a * bb != self.TILE_ROWS:
            raise ValueError("TILE_BLOCK must hold TILE_ROWS patches")
        return ops.block_order(self._patches_points,
                               self._patch_to_wall_ids, (a, bb),
                               pad=self.TILE_PAD)

    def _exchange_plan(self, speed_of_sound, time_resolution, kind=None):
        """Plan of the exchange (cached per speed of sound and bin width).

        ``kind``: "grouped", "ds" (diffuse + specular BRDF rows), "generic"
        or None (automatic: "grouped" when the sources of a destination
        share few BRDF rows, else "ds" when every BRDF row is a constant
        plus one specular entry, else "generic").  On a saved dense F~
        (restored state) the plan is "grouped" when its rows factorise
        into few groups and "generic" otherwise.
        """
        key = (float(speed_of_sound), float(time_resolution), kind,
               self.TILE_ROWS, self.TILE_MIN_MERGE, self.TILE_BLOCK,
               self.TILE_R, self.TILE_PAD)
        if getattr(self, "_plan", None) is not None and self._plan[0] == key:
            return self._plan[1]
        b = self._bake
        g = self._geometry()
        plan = None
        if b.get("fft_dense") is not None:
            # a saved F~ is used as given (reference: the attribute read
            # by _energy_exchange): grouped when its rows factorise into
            # few groups, else per pair
            ctr = self.patches_center
            d = ctr[:, None, :] - ctr[None, :, :]
            dist = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1])
                           + d[..., 2] * d[..., 2])
            if kind in (None, "grouped"):
                plan = ops.GroupedPlan.from_dense(
                    b["vis"], dist, b["p2o"], b["fft_dense"], speed_of_sound,
                    time_resolution, self.GROUPED_MAX_GROUP_FRACTION)
                if plan is None and kind == "grouped":
                    raise ValueError("the saved form_factors_tilde rows do "
                                     "not factorise into few groups")
                if plan is not None and self.TILE_ROWS:
                    self._add_tiles(plan)
            if plan is None:
                plan = ops.ExchangePlan.from_dense(
                    b["vis"], dist, b["p2o"], b["fft_dense"], speed_of_sound,
                    time_resolution)
            self._plan = (key, plan)
            return plan
        args = (b["vis"], g["centers"], g["areas"], b["ff"], b["p2o"],
                b["air"], g["wall_ids"], b["sources"], b["brdf"],
                b["brdf_index"], b["D"], b["n_bins"], speed_of_sound,
                time_resolution)
        if kind in (None, "grouped"):
            try:
                plan = ops.GroupedPlan.from_geometry(*args)
            except ValueError:
                if kind == "grouped":
                    raise
            if (plan is not None and kind is None and plan.Vd and
                    plan.n_groups > self.GROUPED_MAX_GROUP_FRACTION * plan.Vd):
                plan = None
            if plan is not None and self.TILE_ROWS:
                self._add_tiles(plan)
        if plan is None and kind in (None, "ds"):
            try:
                plan = ops.diffuse_specular_plan(*args)
            except ValueError:
                if kind == "ds":
                    raise
            if plan is None and kind == "ds":
                raise ValueError("BRDF rows are not diffuse + specular")
            if plan is not None and self.TILE_ROWS == 16:
                # 16-row tiles, the size the diffuse + specular tiled
                # gather supports
                plan.tiles = ops.tile_plan(plan, 16, self.TILE_MIN_MERGE,
                                           self._tile_order())
        if plan is None:
            plan = ops.ExchangePlan.from_geometry(*args)
        self._plan = (key, plan)
        return plan

    def _add_tiles(self, plan):
        """Destination tiles of TILE_ROWS rows for a grouped plan (kept only
        when they merge at least TILE_MIN_MERGE entries into one)."""
        plan.tiles = ops.tile_plan(plan, self.TILE_ROWS, self.TILE_MIN_MERGE,
                                   self._tile_order())
        if plan.tiles is not None:
            plan.tiles.R = self.TILE_R

    def calculate_energy_exchange(self, speed_of_sound, etc_time_resolution,
                                  etc_duration, max_reflection_order=-1,
                                  recalculate=False):
        """Reflection-order energy exchange
        (reference ``RadiosityFast.py:524-568``)."""
        if speed_of_sound <= 0:
            raise ValueError("Speed of sound must be positive and non-zero")
        if etc_time_resolution <= 0:
            raise ValueError("Time resolution must be positive and non-zero")
        if etc_duration <= 0:
            raise ValueError("Duration must be positive and non-zero")
        n_samples = int(etc_duration / etc_time_resolution)
        if self._etc is None or recalculate:
            if "energy_init_source" not in self._dev:
                raise RuntimeError("init_source_energy() must run before "
                                   "the exchange")
            e0d = self._dev["energy_init_source"]
            d0 = self._dev["distance_patches_to_source"]
            P, D, B = e0d.shape
            if P != self.n_patches or tuple(d0.shape) != (P,):
                raise ValueError(
                    "energy_init_source need to be of shape "
                    "(n_patches, n_outgoing_directions, n_bins)")
            if n_samples < 1:
                raise ValueError("etc_duration must hold at least one "
                                 "etc_time_resolution")
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
        if b.get("fft_dense") is not None:
            return b["fft_dense"]
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
