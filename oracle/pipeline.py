"""The reference class's method bodies, composed from the C oracle.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``OracleRadiosity`` mirrors ``DirectionalRadiosityFast`` method by method
(``sparrowpy/classes/RadiosityFast.py``): the numpy glue between the numba
kernels (visible-pair list, outgoing-direction map, pair distances,
receiver gather) is restated here in numpy, the kernels themselves are the
C functions of ``oracle``.  Everything is FP64.
"""
from __future__ import annotations

import numpy as np

import oracle as orc


def _norm_rows(v):
    v = np.asarray(v, dtype=float)
    return np.sqrt((v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1])
                   + v[..., 2] * v[..., 2])


class OracleRadiosity:
    """FP64 restatement of ``DirectionalRadiosityFast`` (methods at
    ``RadiosityFast.py:369-752``)."""

    def __init__(self, walls_points, walls_normal, walls_up_vector,
                 patches_points, patch_to_wall_ids, rotate_fn):
        self.walls_points = np.asarray(walls_points, float)
        self.walls_normal = np.asarray(walls_normal, float)
        self.walls_up_vector = np.asarray(walls_up_vector, float)
        self.patches_points = np.asarray(patches_points, float)
        self.patch_to_wall_ids = np.asarray(patch_to_wall_ids, np.int64)
        self.n_patches = self.patches_points.shape[0]
        self.n_walls = self.walls_points.shape[0]
        self._rotate = rotate_fn
        self.frequencies = None
        self.brdf = None
        self.brdf_index = None
        self.brdf_in = None
        self.brdf_out = None
        self.air = None
        self.source = None

    # geometry.py:412-448
    @property
    def patches_center(self):
        return np.sum(self.patches_points, axis=-2) / 4

    @property
    def patches_area(self):
        out = np.zeros(self.n_patches)
        for i, p in enumerate(self.patches_points):
            for t in range(2):
                a = p[t + 1] - p[0]
                b = p[t + 2] - p[0]
                c = np.array([a[1] * b[2] - a[2] * b[1],
                              a[2] * b[0] - a[0] * b[2],
                              a[0] * b[1] - a[1] * b[0]])
                out[i] += .5 * _norm_rows(c)
        return out

    @property
    def patches_normal(self):
        return self.walls_normal[self.patch_to_wall_ids]

    @property
    def n_bins(self):
        return None if self.frequencies is None else self.frequencies.size

    # RadiosityFast.py:766-816
    def set_wall_brdf(self, wall_indexes, brdf, incoming, outgoing,
                      frequencies):
        incoming = np.asarray(incoming, float).reshape(-1, 3)
        outgoing = np.asarray(outgoing, float).reshape(-1, 3)
        self._check_frequencies(frequencies)
        if self.brdf_in is None:
            self.brdf_in = np.zeros((self.n_walls, incoming.shape[0], 3))
            self.brdf_out = np.zeros((self.n_walls, outgoing.shape[0], 3))
            self.brdf_index = np.full(self.n_walls, -1, dtype=np.int64)
            self.brdf = []
        for w in np.atleast_1d(wall_indexes):
            self.brdf_in[w] = self._rotate(self.walls_normal[w],
                                           self.walls_up_vector[w], incoming)
            self.brdf_out[w] = self._rotate(self.walls_normal[w],
                                            self.walls_up_vector[w], outgoing)
        self.brdf.append(np.asarray(brdf, float) * np.pi)
        self.brdf_index[np.atleast_1d(wall_indexes)] = len(self.brdf) - 1

    def set_air_attenuation(self, air, frequencies):
        self._check_frequencies(frequencies)
        self.air = np.atleast_1d(np.asarray(air, float))

    def _check_frequencies(self, frequencies):
        frequencies = np.atleast_1d(np.asarray(frequencies, float))
        if self.frequencies is None:
            self.frequencies = frequencies
        else:
            assert np.array_equal(self.frequencies, frequencies)

    # RadiosityFast.py:369-433
    def bake_geometry(self, nthreads=None):
        P = self.n_patches
        centers = self.patches_center
        self.visibility_matrix = orc.p2p_visibility(
            centers, self.patches_normal, self.patches_points, nthreads)
        ii, jj = np.nonzero(self.visibility_matrix)
        self.visible_patches = np.stack([ii, jj], axis=-1).astype(np.int32)
        self.form_factors = orc.form_factors(
            self.patches_points, self.patches_normal, self.patches_area,
            self.visible_patches, nthreads)
        if self.brdf_in is not None:
            D = self.brdf_out.shape[1]
            self.p2o = np.full((P, P), D, dtype=np.int64)
            vv = self.visibility_matrix | self.visibility_matrix.T
            for j in range(P):
                vis = np.nonzero(vv[:, j])[0]
                if vis.size:
                    self.p2o[vis, j] = orc.receiver_index(
                        centers[vis], centers[j], self.brdf_out,
                        self.patch_to_wall_ids[vis])
            brdf = np.array(self.brdf)
            brdf_index, sources = self.brdf_index, self.brdf_in
            receivers = self.brdf_out
        else:
            self.p2o = np.zeros((P, P), dtype=np.int64)
            brdf = brdf_index = sources = receivers = None
        n_bins = 1 if self.frequencies is None else self.n_bins
        self.fft = orc.fft(self.visibility_matrix, self.form_factors, n_bins,
                           centers, self.patches_area, self.air,
                           self.patch_to_wall_ids, brdf, brdf_index, sources,
                           receivers, nthreads)

    # RadiosityFast.py:436-522 (omni sources)
    def init_source_energy(self, source):
        source = np.asarray(source, float).reshape(3)
        self.source = source
        if self.brdf_in is None:
            freqs = np.array([0.]) if self.frequencies is None else \
                self.frequencies
            self.set_wall_brdf(np.arange(self.n_walls),
                               np.ones((1, 1, freqs.size)), [[0, 0, 1]],
                               [[0, 0, 1]], freqs)
        if self.air is None:
            freqs = np.array([0.]) if self.frequencies is None else \
                self.frequencies
            self.set_air_attenuation(np.zeros(freqs.size), freqs)
        centers = self.patches_center
        self.source_visibility = orc.point_visibility(
            source, centers, self.walls_normal, self.walls_points)
        e0, d0 = orc.source_energy(source, centers, self.patches_points,
                                   self.source_visibility, self.air,
                                   self.n_bins)
        self.energy_0 = e0
        self.energy_init_source = orc.add_directional(
            e0, source, centers, self.patch_to_wall_ids, self.brdf_in,
            self.brdf_out, np.array(self.brdf), self.brdf_index)
        self.distance_patches_to_source = d0

    def pair_distances(self):
        c = self.patches_center
        return _norm_rows(c[:, None, :] - c[None, :, :])

    # RadiosityFast.py:524-568
    def calculate_energy_exchange(self, speed_of_sound, etc_time_resolution,
                                  etc_duration, max_reflection_order=-1,
                                  nthreads=1):
        T = int(etc_duration / etc_time_resolution)
        self.n_samples = T
        if max_reflection_order < 1:
            self.etc = orc.init_energy(
                T, self.energy_init_source, self.distance_patches_to_source,
                speed_of_sound, etc_time_resolution)
        else:
            self.etc = orc.exchange(
                T, self.energy_init_source, self.distance_patches_to_source,
                self.pair_distances(), self.fft, self.p2o, speed_of_sound,
                etc_time_resolution, max_reflection_order,
                self.visible_patches, nthreads)
        self.speed_of_sound = float(speed_of_sound)
        self.etc_time_resolution = float(etc_time_resolution)

    # RadiosityFast.py:686-752
    def collect_energy_receiver_patchwise(self, receivers):
        receivers = np.asarray(receivers, float).reshape(-1, 3)
        P = self.n_patches
        centers = self.patches_center
        T = self.etc.shape[-1]
        out = np.zeros((receivers.shape[0], P, self.n_bins, T))
        for r, rp in enumerate(receivers):
            vis = orc.point_visibility(rp, centers, self.walls_normal,
                                       self.walls_points)
            g = orc.receiver_energy(rp, self.patches_points, vis)
            idx = orc.receiver_index(centers, rp, self.brdf_out,
                                     self.patch_to_wall_ids)
            E = self.etc[np.arange(P), idx] * g[:, None, None]
            dist = _norm_rows(centers - rp)
            out[r] = orc.collect_receiver_energy(
                E, dist, self.speed_of_sound, self.etc_time_resolution,
                self.air)
        return out

    # RadiosityFast.py:570-657
    def collect_energy_receiver_mono(self, receivers, direct_sound=False):
        out = np.sum(self.collect_energy_receiver_patchwise(receivers),
                     axis=1)
        if direct_sound:
            ds, n = self.calculate_direct_sound(receivers)
            out[np.arange(len(n)), :, n] += ds
        return out

    def calculate_direct_sound(self, receivers):
        receivers = np.asarray(receivers, float).reshape(-1, 3)
        r = _norm_rows(receivers - self.source)
        ds = np.ones((receivers.shape[0], self.n_bins))
        ds *= (1 / (4 * np.pi * r ** 2))[:, None]
        for b in range(self.n_bins):
            ds[:, b] *= np.exp(-self.air[b] * r)
        n = np.array(r / self.speed_of_sound / self.etc_time_resolution,
                     dtype=int)
        return ds, n
