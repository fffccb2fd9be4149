"""Test scenes and closed-form form factors.

``shoebox_room_stub`` follows ``sparrowpy/testing/stub_utils.py:5-48`` (wall
order, corner order, up vectors and inward normals).  The closed-form form
factors are the textbook solutions for rectangles (Howell's catalogue), the
same ones the reference tests against (``sparrowpy/testing/
exact_ff_solutions.py``).
"""
from __future__ import annotations

import numpy as np

from .geometry import Polygon

__all__ = ["shoebox_room_stub", "ff_parallel_patches",
           "ff_perpendicular_shared_edge", "ff_perpendicular_shared_vertex",
           "contour_form_factor"]


def shoebox_room_stub(length_x, length_y, length_z):
    """Six inward-facing walls of an axis-aligned box at the origin."""
    X, Y, Z = length_x, length_y, length_z
    return [
        Polygon([[0, 0, 0], [X, 0, 0], [X, 0, Z], [0, 0, Z]],
                [1, 0, 0], [0, 1, 0]),
        Polygon([[0, Y, 0], [X, Y, 0], [X, Y, Z], [0, Y, Z]],
                [1, 0, 0], [0, -1, 0]),
        Polygon([[0, 0, 0], [X, 0, 0], [X, Y, 0], [0, Y, 0]],
                [1, 0, 0], [0, 0, 1]),
        Polygon([[0, 0, Z], [X, 0, Z], [X, Y, Z], [0, Y, Z]],
                [1, 0, 0], [0, 0, -1]),
        Polygon([[0, 0, 0], [0, 0, Z], [0, Y, Z], [0, Y, 0]],
                [0, 0, 1], [1, 0, 0]),
        Polygon([[X, 0, 0], [X, 0, Z], [X, Y, Z], [X, Y, 0]],
                [0, 0, 1], [-1, 0, 0]),
    ]


def ff_parallel_patches(a, b, c):
    """Two directly opposed a x b rectangles at distance c."""
    x, y = a / c, b / c
    g = 0.5 * np.log((1 + x * x) * (1 + y * y) / (1 + x * x + y * y))
    h = x * np.sqrt(1 + y * y) * np.arctan(x / np.sqrt(1 + y * y))
    i = y * np.sqrt(1 + x * x) * np.arctan(y / np.sqrt(1 + x * x))
    return 2 / (np.pi * x * y) * (g + h + i - x * np.arctan(x)
                                  - y * np.arctan(y))


def ff_perpendicular_shared_edge(w, h, length):
    """Perpendicular rectangles sharing an edge of length ``length``.

    Emitter has width ``h`` (height), receiver width ``w``.
    """
    W, H = w / length, h / length
    z2 = W * W + H * H
    z = np.sqrt(z2)
    t1 = W * np.arctan(1 / W) + H * np.arctan(1 / H) - z * np.arctan(1 / z)
    a = (1 + W * W) * (1 + H * H) / (1 + z2)
    b = (W * W * (1 + z2) / ((1 + W * W) * z2)) ** (W * W)
    c = (H * H * (1 + z2) / ((1 + H * H) * z2)) ** (H * H)
    return (t1 + 0.25 * np.log(a * b * c)) / (np.pi * H)


def ff_perpendicular_shared_vertex(aa, bb, cc, dd):
    """Perpendicular rectangles meeting at one corner.

    Computed by contour (Stokes) integration of ``ln r`` with high-order
    Gauss-Legendre quadrature on each edge pair, which converges to the
    exact value for rectangles that touch only at a vertex.  Emitter: the
    ``cc x bb`` rectangle in the z=0 plane; receiver: the ``aa x dd``
    rectangle in the x=0 plane on the far side of the shared vertex.
    """
    emitter = np.array([[0, 0, 0], [cc, 0, 0], [cc, bb, 0], [0, bb, 0.]])
    receiver = np.array([[0, 0, 0], [0, 0, aa], [0, -dd, aa],
                         [0, -dd, 0.]])
    return contour_form_factor(emitter, receiver, n=64)


def contour_form_factor(pi, pj, n=64):
    """Form factor F_{i->j} of two planar polygons by the double contour
    integral ``1/(2 pi A_i) sum_edges int int ln r dx_i . dx_j``.

    Gauss-Legendre with ``n`` nodes per edge; an independent reference for
    the product's Boole-rule contour quadrature.
    """
    pi = np.asarray(pi, float)
    pj = np.asarray(pj, float)
    x, w = np.polynomial.legendre.leggauss(n)
    s = 0.5 * (x + 1)
    w = 0.5 * w
    total = 0.0
    for a in range(len(pi)):
        p0, p1 = pi[a], pi[(a + 1) % len(pi)]
        da = p1 - p0
        qa = p0[None, :] + s[:, None] * da[None, :]
        for b in range(len(pj)):
            q0, q1 = pj[b], pj[(b + 1) % len(pj)]
            db = q1 - q0
            dot = float(np.dot(da, db))
            if dot == 0.0:
                continue
            qb = q0[None, :] + s[:, None] * db[None, :]
            r = np.linalg.norm(qa[:, None, :] - qb[None, :, :], axis=-1)
            with np.errstate(divide="ignore"):
                lr = np.log(r)
            lr[~np.isfinite(lr)] = 0.0
            total += dot * np.sum(w[:, None] * w[None, :] * lr)
    area = 0.0
    for t in range(len(pi) - 2):
        area += 0.5 * np.linalg.norm(np.cross(pi[t + 1] - pi[0],
                                              pi[t + 2] - pi[0]))
    return abs(total / (2 * np.pi * area))
