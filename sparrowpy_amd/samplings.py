"""Spherical direction samplings used to discretise BRDFs.

sparrowpy's tests and examples build BRDF direction sets with
``pyfar.samplings.sph_gaussian`` (``tests/conftest.py:61-62``) and
``pyfar.samplings.sph_equal_angle`` (``tests/test_DRadiosityFast.py:64``),
keeping the upper hemisphere (``z > 0``).  These are restated here from the
underlying quadrature rules: a Gauss-Legendre rule in cos(colatitude) times
an equiangular azimuth rule, and an equal-angle grid.
"""
from __future__ import annotations

import numpy as np

from .coordinates import Coordinates

__all__ = ["sph_gaussian", "sph_equal_angle", "hemisphere"]


def sph_gaussian(n_points=None, sh_order=None, radius=1.0):
    """Gaussian sampling: Gauss-Legendre colatitudes x equiangular azimuths.

    ``sh_order=N`` gives ``N+1`` colatitudes and ``2(N+1)`` azimuths; points
    are ordered azimuth-major (all colatitudes of the first azimuth first).
    Weights are the product quadrature weights normalised to sum 1.
    """
    if n_points is None and sh_order is None:
        raise ValueError("either n_points or sh_order is required")
    if sh_order is not None:
        n_theta, n_phi = sh_order + 1, 2 * (sh_order + 1)
    else:
        n_points = np.atleast_1d(np.asarray(n_points, dtype=int))
        if n_points.size == 1:
            n_theta, n_phi = int(n_points[0]), 2 * int(n_points[0])
        else:
            n_theta, n_phi = int(n_points[0]), int(n_points[1])
    nodes, w = np.polynomial.legendre.leggauss(n_theta)
    theta = np.arccos(nodes)
    phi = np.arange(n_phi) * (2 * np.pi / n_phi)
    th, ph = np.meshgrid(theta, phi)
    weights = np.tile(w, n_phi)
    weights = weights / np.sum(weights)
    return Coordinates.from_spherical_colatitude(
        ph.reshape(-1), th.reshape(-1), radius, weights=weights)


def sph_equal_angle(delta_angles, radius=1.0):
    """Equal-angle grid with one point at each pole.

    ``delta_angles`` (degrees) may be a scalar or (delta_azimuth,
    delta_colatitude).
    """
    d = np.atleast_1d(np.asarray(delta_angles, dtype=float))
    d_az, d_col = (d[0], d[0]) if d.size == 1 else (d[0], d[1])
    az = np.arange(0.0, 360.0, d_az)
    col = np.arange(0.0, 180.0 + d_col / 2, d_col)
    az_all, col_all = [], []
    for c in col:
        if c == 0.0 or c == 180.0:
            az_all.append(0.0)
            col_all.append(c)
        else:
            az_all.extend(az)
            col_all.extend([c] * az.size)
    return Coordinates.from_spherical_colatitude(
        np.deg2rad(az_all), np.deg2rad(col_all), radius)


def hemisphere(coords: Coordinates) -> Coordinates:
    """Keep the directions with ``z > 0``."""
    return coords[coords.z > 0]
