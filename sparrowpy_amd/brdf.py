"""Discretised BRDFs built from scattering and absorption coefficients.

Host-side input producers for the radiosity path (their tables are uploaded
once and indexed by the GPU kernels).  Restated from
``sparrowpy/brdf.py:8-130`` (``create_from_scattering``) and ``:133-224``
(``create_from_directional_scattering``); SOFA file writing is not part of
this package.
"""
from __future__ import annotations

import numpy as np

from .coordinates import Coordinates, FrequencyData

__all__ = ["create_from_scattering", "create_from_directional_scattering"]


def _has_coords(obj):
    return hasattr(obj, "cartesian") and hasattr(obj, "csize")


def create_from_scattering(source_directions, receiver_directions,
                           scattering_coefficient,
                           absorption_coefficient=None):
    r"""BRDF of a surface with scattering coefficient ``s`` and absorption
    ``alpha`` (reference ``brdf.py:8-130``).

    .. math::
        \rho(\Omega_i,\Omega_o) = \frac{s}{\pi}
        + \frac{1-s}{\cos\theta\, w_o}\,\delta(\Omega_o - M(\Omega_i)),
        \quad \rho \mathrel{*}= (1-\alpha)

    ``w_o`` are the receiver weights normalised to sum ``2*pi``; the specular
    receiver of source ``k`` is the receiver nearest to ``k`` mirrored in
    azimuth.  The specular denominator is taken from the (receiver, source)
    outer product indexed ``[k, specular(k)]``, exactly as the reference
    does (``brdf.py:110-118``).
    """
    if not isinstance(scattering_coefficient, FrequencyData) and not (
            hasattr(scattering_coefficient, "freq")
            and hasattr(scattering_coefficient, "frequencies")):
        raise TypeError(
            "scattering_coefficient must be a FrequencyData object with "
            "shape (1,)")
    if tuple(scattering_coefficient.cshape) != (1,):
        raise TypeError(
            "scattering_coefficient must be a FrequencyData object with "
            "shape (1,)")
    if not _has_coords(source_directions):
        raise TypeError("source_directions must be a Coordinates object")
    if not _has_coords(receiver_directions):
        raise TypeError("receiver_directions must be a Coordinates object")
    if absorption_coefficient is None:
        absorption_coefficient = FrequencyData(
            np.zeros_like(scattering_coefficient.frequencies),
            scattering_coefficient.frequencies)

    n_bins = scattering_coefficient.frequencies.size
    brdf = np.zeros((source_directions.csize, receiver_directions.csize,
                     n_bins))
    receiver_weights = np.array(receiver_directions.weights, dtype=float)
    receiver_weights *= 2 * np.pi / np.sum(receiver_weights)
    s = np.real(np.asarray(scattering_coefficient.freq).flatten())
    image = Coordinates.from_cartesian(source_directions.cartesian)
    image.azimuth = image.azimuth + np.pi
    i_receiver = _nearest(receiver_directions.cartesian, image.cartesian)
    cos_factor = (np.cos(_colatitude(source_directions.cartesian)[
        np.newaxis, ...]) * receiver_weights[..., np.newaxis])
    scattering_factor = 1 - s[np.newaxis, ...]
    brdf[:, :, :] += s / np.pi
    i_sources = np.arange(source_directions.csize)
    brdf[i_sources, i_receiver, :] += scattering_factor / cos_factor[
        i_sources, i_receiver, np.newaxis]
    brdf *= (1 - np.asarray(absorption_coefficient.freq).flatten())
    return FrequencyData(brdf, scattering_coefficient.frequencies)


def create_from_directional_scattering(source_directions,
                                       receiver_directions,
                                       directional_scattering,
                                       absorption_coefficient=None):
    r"""BRDF from a directional scattering coefficient ``s_d`` (reference
    ``brdf.py:133-224``): ``rho = (1-alpha) s_d / (cos(theta_o) w_o)``."""
    if not _has_coords(source_directions):
        raise TypeError("source_directions must be a Coordinates object")
    if not _has_coords(receiver_directions):
        raise TypeError("receiver_directions must be a Coordinates object")
    if tuple(directional_scattering.cshape) != (
            source_directions.csize, receiver_directions.csize):
        raise TypeError(
            "directional_scattering must be a FrequencyData object with "
            f"cshape ({source_directions.csize}, "
            f"{receiver_directions.csize})")
    if absorption_coefficient is None:
        absorption_coefficient = FrequencyData(
            np.zeros_like(directional_scattering.frequencies),
            directional_scattering.frequencies)
    cos_receiver = np.cos(_colatitude(receiver_directions.cartesian))[
        np.newaxis, :, np.newaxis]
    receiver_weights = np.array(receiver_directions.weights, dtype=float)
    receiver_weights *= 2 * np.pi / np.sum(receiver_weights)
    receiver_factor = receiver_weights[..., np.newaxis]
    brdf = np.asarray(directional_scattering.freq) / receiver_factor \
        / cos_receiver
    brdf = brdf * (1 - np.asarray(absorption_coefficient.freq).flatten())
    return FrequencyData(brdf, absorption_coefficient.frequencies)


def _colatitude(xyz):
    r = np.sqrt(np.sum(xyz ** 2, axis=-1))
    return np.arccos(np.clip(xyz[..., 2] / r, -1.0, 1.0))


def _nearest(points, queries):
    d2 = np.sum((points[None, :, :] - queries[:, None, :]) ** 2, axis=-1)
    return np.argmin(d2, axis=-1)
