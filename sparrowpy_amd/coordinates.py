"""Direction, frequency and histogram containers for the radiosity path.

sparrowpy passes pyfar objects across its API: ``pyfar.Coordinates`` for
source/receiver positions and BRDF direction sets, ``pyfar.FrequencyData`` for
BRDFs and air attenuation, ``pyfar.TimeData`` for the returned energy
histograms (reference: ``sparrowpy/classes/RadiosityFast.py:449-452``,
``:475-478``, ``:683-684``, ``:754-816``).  pyfar is not a dependency of this
package, so these classes carry exactly the attributes that path reads, with
pyfar's names and shapes.  Any object exposing the same attributes
(``.cartesian``, ``.weights``, ``.freq``, ``.frequencies``) is accepted by
:class:`sparrowpy_amd.DirectionalRadiosityFast`, so real pyfar objects work too.
"""
from __future__ import annotations

import numpy as np

__all__ = ["Coordinates", "FrequencyData", "TimeData"]


def _as_cart(x, y, z):
    x = np.atleast_1d(np.asarray(x, dtype=float))
    y = np.atleast_1d(np.asarray(y, dtype=float))
    z = np.atleast_1d(np.asarray(z, dtype=float))
    x, y, z = np.broadcast_arrays(x, y, z)
    return np.stack([x, y, z], axis=-1).reshape(-1, 3).copy()


class Coordinates:
    """Cartesian point/direction set with a one-dimensional cshape.

    Mirrors the parts of ``pyfar.Coordinates`` used by sparrowpy: cartesian,
    spherical (colatitude/elevation) views, weights, subtraction, nearest
    neighbour search and rotation.
    """

    def __init__(self, x=None, y=None, z=None, weights=None):
        if x is None:
            self._xyz = np.zeros((0, 3))
        else:
            self._xyz = _as_cart(x, y, z)
        self.weights = weights

    # construction -----------------------------------------------------
    @classmethod
    def from_cartesian(cls, xyz, weights=None):
        xyz = np.atleast_2d(np.asarray(xyz, dtype=float))
        obj = cls()
        obj._xyz = xyz.reshape(-1, 3).copy()
        obj.weights = weights
        return obj

    @classmethod
    def from_spherical_colatitude(cls, azimuth, colatitude, radius,
                                  weights=None):
        az = np.atleast_1d(np.asarray(azimuth, dtype=float))
        col = np.atleast_1d(np.asarray(colatitude, dtype=float))
        r = np.atleast_1d(np.asarray(radius, dtype=float))
        az, col, r = np.broadcast_arrays(az, col, r)
        x = r * np.sin(col) * np.cos(az)
        y = r * np.sin(col) * np.sin(az)
        z = r * np.cos(col)
        return cls(x, y, z, weights=weights)

    @classmethod
    def from_spherical_elevation(cls, azimuth, elevation, radius,
                                 weights=None):
        return cls.from_spherical_colatitude(
            azimuth, np.pi / 2 - np.asarray(elevation, dtype=float), radius,
            weights)

    # shape ------------------------------------------------------------
    @property
    def cshape(self):
        return (self._xyz.shape[0],)

    @property
    def csize(self):
        return self._xyz.shape[0]

    @property
    def cdim(self):
        return 1

    # cartesian --------------------------------------------------------
    @property
    def cartesian(self):
        return self._xyz

    @cartesian.setter
    def cartesian(self, value):
        self._xyz = np.asarray(value, dtype=float).reshape(-1, 3).copy()

    @property
    def x(self):
        return self._xyz[:, 0]

    @property
    def y(self):
        return self._xyz[:, 1]

    @property
    def z(self):
        return self._xyz[:, 2]

    # spherical --------------------------------------------------------
    @property
    def radius(self):
        return np.sqrt(self.x ** 2 + self.y ** 2 + self.z ** 2)

    @radius.setter
    def radius(self, value):
        self._set_spherical(radius=value)

    @property
    def azimuth(self):
        return np.mod(np.arctan2(self.y, self.x), 2 * np.pi)

    @azimuth.setter
    def azimuth(self, value):
        self._set_spherical(azimuth=value)

    @property
    def colatitude(self):
        r = self.radius
        with np.errstate(invalid="ignore", divide="ignore"):
            c = np.where(r > 0, self.z / np.where(r > 0, r, 1.0), 1.0)
        return np.arccos(np.clip(c, -1.0, 1.0))

    @colatitude.setter
    def colatitude(self, value):
        self._set_spherical(colatitude=value)

    @property
    def elevation(self):
        return np.pi / 2 - self.colatitude

    def _set_spherical(self, radius=None, azimuth=None, colatitude=None):
        r = self.radius if radius is None else np.broadcast_to(
            np.asarray(radius, dtype=float), self.cshape)
        az = self.azimuth if azimuth is None else np.broadcast_to(
            np.asarray(azimuth, dtype=float), self.cshape)
        col = self.colatitude if colatitude is None else np.broadcast_to(
            np.asarray(colatitude, dtype=float), self.cshape)
        self._xyz = _as_cart(r * np.sin(col) * np.cos(az),
                             r * np.sin(col) * np.sin(az),
                             r * np.cos(col))

    # weights ----------------------------------------------------------
    @property
    def weights(self):
        return self._weights

    @weights.setter
    def weights(self, value):
        if value is None:
            self._weights = None
        else:
            w = np.atleast_1d(np.asarray(value, dtype=float)).copy()
            self._weights = np.broadcast_to(w, self.cshape).copy()

    # operations -------------------------------------------------------
    def copy(self):
        obj = Coordinates.from_cartesian(self._xyz.copy())
        obj._weights = None if self._weights is None else \
            self._weights.copy()
        return obj

    def __getitem__(self, index):
        xyz = self._xyz[index]
        obj = Coordinates.from_cartesian(np.atleast_2d(xyz))
        if self._weights is not None:
            obj._weights = np.atleast_1d(self._weights[index]).copy()
        return obj

    def __len__(self):
        return self.csize

    def __sub__(self, other):
        return Coordinates.from_cartesian(
            self._xyz - np.asarray(other.cartesian))

    def __add__(self, other):
        return Coordinates.from_cartesian(
            self._xyz + np.asarray(other.cartesian))

    def find_nearest(self, find, k=1):
        """Index of the nearest point (euclidean) for each point in ``find``.

        Returns ``((index,), distance)`` like pyfar; ties keep the first
        point.
        """
        pts = np.asarray(find.cartesian)
        d2 = np.sum((self._xyz[None, :, :] - pts[:, None, :]) ** 2, axis=-1)
        idx = np.argmin(d2, axis=-1)
        dist = np.sqrt(d2[np.arange(pts.shape[0]), idx])
        return (idx,), dist

    def rotate(self, rotation, value=None, degrees=True, inverse=False):
        """Rotate the points, ``rotation`` as in scipy's ``from_euler``."""
        from scipy.spatial.transform import Rotation
        rot = Rotation.from_euler(rotation, value, degrees=degrees)
        if inverse:
            rot = rot.inv()
        self._xyz = rot.apply(self._xyz).reshape(-1, 3)

    def __repr__(self):
        return f"Coordinates(cshape={self.cshape})"


class FrequencyData:
    """Frequency-domain data whose last axis is the frequency bin."""

    def __init__(self, data, frequencies):
        freqs = np.atleast_1d(np.asarray(frequencies, dtype=float))
        data = np.asarray(data)
        if data.ndim == 0:
            data = data.reshape(1, 1)
        elif data.ndim == 1:
            data = data.reshape(1, -1)
        if data.shape[-1] != freqs.size:
            raise ValueError(
                "the last dimension of data must match the number of "
                "frequencies")
        self._data = np.array(data, dtype=complex if np.iscomplexobj(data)
                              else float)
        self._frequencies = freqs

    @property
    def freq(self):
        return self._data

    @freq.setter
    def freq(self, value):
        self._data = np.asarray(value)

    @property
    def frequencies(self):
        return self._frequencies

    @property
    def n_bins(self):
        return self._frequencies.size

    @property
    def cshape(self):
        return self._data.shape[:-1]

    def __repr__(self):
        return f"FrequencyData(cshape={self.cshape}, n_bins={self.n_bins})"


class TimeData:
    """Energy histograms: the last axis is time, sampled at ``times``."""

    def __init__(self, data, times):
        self.time = np.asarray(data)
        self._times = np.asarray(times, dtype=float)

    @property
    def times(self):
        return self._times

    @property
    def n_samples(self):
        return self._times.size

    @property
    def cshape(self):
        return self.time.shape[:-1]

    def __getitem__(self, index):
        sub = self.time[index]
        if sub.ndim == 1:
            sub = sub[None, :]
        return TimeData(sub, self._times)

    def __repr__(self):
        return f"TimeData(cshape={self.cshape}, n_samples={self.n_samples})"
