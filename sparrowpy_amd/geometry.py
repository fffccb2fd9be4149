"""Scene geometry on the host: walls, patch subdivision, BRDF frames.

These are the input producers of the radiosity path (wall polygons in,
patch corner arrays out); the per-pair geometry (visibility, form factors,
delays) runs in the HIP kernels.

Reference: ``sparrowpy/geometry.py`` -- ``Polygon`` (:15-230),
``_process_patches`` (:290-339), ``_total_number_of_patches`` (:341-371),
``_create_patches`` (:373-410), ``_calculate_center`` (:412-413),
``_calculate_size`` (:415-418), ``_polygon_area`` (:420-440),
``_calculate_area`` (:442-448); and
``sparrowpy/classes/RadiosityFast.py:971-986`` (``_rotate_coords_to_normal``).
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "Polygon", "process_patches", "total_number_of_patches",
    "create_patches", "calculate_center", "calculate_size",
    "polygon_area", "calculate_area", "rotate_coords_to_normal",
]


class Polygon:
    """Planar convex polygon (a wall) with an up vector and a normal.

    Same constructor contract as the reference (``geometry.py:25-62``):
    points must be coplanar and the normal perpendicular to them.
    """

    def __init__(self, points, up_vector, normal):
        self.pts = np.array(points, dtype=float)
        normal = np.array(normal, dtype=float)
        if self.pts.shape[1] < 3:
            raise ValueError("You need at least 3 points to build a Polygon")
        if self.n_points > 3:
            x_0 = self.pts[0]
            for i in range(1, self.n_points - 2):
                det = np.linalg.det(
                    [x_0 - self.pts[i], x_0 - self.pts[i + 1],
                     x_0 - self.pts[i + 2]])
                if not abs(det) < 1e-12:
                    raise ValueError(
                        "Points must be in a plane to create a Polygon")
        up = np.array(up_vector, dtype=float)
        self.up_vector = up / np.sqrt(np.dot(up, up))
        calc = np.cross(self.pts[0] - self.pts[1], self.pts[0] - self.pts[2])
        calc = calc / np.sqrt(np.dot(calc, calc))
        if not np.all(np.abs(np.cross(normal, calc)) < 1e-10):
            raise ValueError(
                "The normal vector is not perpendicular to the polygon")
        self._normal = normal

    @property
    def normal(self):
        return self._normal

    @property
    def n_points(self):
        return self.pts.shape[0]

    @property
    def center(self):
        return np.sum(self.pts, axis=0) / self.n_points

    @property
    def area(self):
        if self.n_points == 3:
            tris = [self.pts]
        elif self.n_points == 4:
            tris = [self.pts[0:3], [self.pts[2], self.pts[3], self.pts[0]]]
        else:
            c = self.center
            tris = [[self.pts[i], self.pts[(i + 1) % self.n_points], c]
                    for i in range(self.n_points)]
        area = 0.0
        for t in tris:
            t = np.asarray(t)
            area += .5 * np.linalg.norm(np.cross(t[1] - t[0], t[2] - t[0]))
        return area

    @property
    def size(self):
        vec1 = self.pts[0] - self.pts[1]
        vec2 = self.pts[1] - self.pts[2]
        return np.abs(vec1 - vec2)

    def to_dict(self):
        return {"up_vector": self.up_vector.tolist(),
                "pts": self.pts.tolist(), "normal": self._normal.tolist()}

    @classmethod
    def from_dict(cls, d):
        return cls(d["pts"], d["up_vector"], d["normal"])


def _patch_nums(polygon_points, max_size):
    size = polygon_points.max(axis=0) - polygon_points.min(axis=0)
    nums = np.array([int(n) for n in size / max_size])
    x_idx = y_idx = None
    # later tests override earlier ones, as in the reference
    if nums[2] == 0:
        x_idx, y_idx = 0, 1
    if nums[1] == 0:
        x_idx, y_idx = 0, 2
    if nums[0] == 0:
        x_idx, y_idx = 1, 2
    if x_idx is None:
        raise ValueError(
            "walls must be axis-aligned rectangles (one extent must be "
            "smaller than the patch size)")
    return size, nums, x_idx, y_idx


def total_number_of_patches(polygon_points, max_size):
    """Patches of one wall: ``int(extent/max_size)`` per in-plane axis."""
    _, nums, x_idx, y_idx = _patch_nums(np.asarray(polygon_points, float),
                                        max_size)
    return int(nums[x_idx] * nums[y_idx])


def create_patches(polygon_points, max_size):
    """Split an axis-aligned rectangle into a regular grid of patches.

    Patch corners are ordered (x0,y0), (x1,y0), (x1,y1), (x0,y1) in the wall's
    two in-plane axes; the out-of-plane coordinate is copied from the wall's
    corners (reference ``geometry.py:373-410``).
    """
    polygon_points = np.asarray(polygon_points, dtype=float)
    size, nums, x_idx, y_idx = _patch_nums(polygon_points, max_size)
    with np.errstate(divide="ignore", invalid="ignore"):
        real_size = size / nums
    x_min = np.min(polygon_points.T[x_idx])
    y_min = np.min(polygon_points.T[y_idx])
    n = int(nums[x_idx] * nums[y_idx])
    out = np.empty((n, 4, 3))
    k = 0
    for i_x in range(nums[x_idx]):
        for i_y in range(nums[y_idx]):
            p = polygon_points.copy()
            p[0, x_idx] = x_min + i_x * real_size[x_idx]
            p[0, y_idx] = y_min + i_y * real_size[y_idx]
            p[1, x_idx] = x_min + (i_x + 1) * real_size[x_idx]
            p[1, y_idx] = y_min + i_y * real_size[y_idx]
            p[3, x_idx] = x_min + i_x * real_size[x_idx]
            p[3, y_idx] = y_min + (i_y + 1) * real_size[y_idx]
            p[2, x_idx] = x_min + (i_x + 1) * real_size[x_idx]
            p[2, y_idx] = y_min + (i_y + 1) * real_size[y_idx]
            out[k] = p
            k += 1
    return out


def process_patches(walls_points, walls_normal, patch_size, n_walls=None):
    """Subdivide every wall; returns (points, normals, n_patches, wall_ids).

    Reference ``geometry.py:290-339``.
    """
    walls_points = np.asarray(walls_points, dtype=float)
    walls_normal = np.asarray(walls_normal, dtype=float)
    blocks = [create_patches(walls_points[i], patch_size)
              for i in range(walls_points.shape[0])]
    pts = np.concatenate(blocks, axis=0) if blocks else np.zeros((0, 4, 3))
    ids = np.concatenate([np.full(b.shape[0], i, dtype=np.int64)
                          for i, b in enumerate(blocks)])
    return pts, walls_normal[ids, :], pts.shape[0], ids


def calculate_center(points):
    points = np.asarray(points, dtype=float)
    return np.sum(points, axis=-2) / points.shape[-2]


def calculate_size(points):
    points = np.asarray(points, dtype=float)
    vec1 = points[..., 0, :] - points[..., 1, :]
    vec2 = points[..., 1, :] - points[..., 2, :]
    return np.abs(vec1 - vec2)


def polygon_area(pts):
    """Fan-triangulated area of a convex polygon (``geometry.py:420-440``)."""
    pts = np.asarray(pts, dtype=float)
    area = 0.0
    for tri in range(pts.shape[0] - 2):
        c = np.cross(pts[tri + 1] - pts[0], pts[tri + 2] - pts[0])
        area += .5 * np.sqrt((c[0] * c[0] + c[1] * c[1]) + c[2] * c[2])
    return area


def calculate_area(points):
    points = np.asarray(points, dtype=float)
    return np.array([polygon_area(p) for p in points])


def rotate_coords_to_normal(wall_normal, wall_up_vector, directions):
    """Rotate BRDF directions from the BRDF frame into a wall's frame.

    BRDFs are tabulated with normal ``(0, 0, 1)`` and up vector
    ``(1, 0, 0)``.  The reference composes two view/up orientations and
    applies their difference as Euler angles
    (``RadiosityFast.py:971-986``); the resulting rotation maps the BRDF
    normal to the wall normal and the BRDF up vector to the wall up vector,
    which is the rotation built here directly: columns ``(up, normal x up,
    normal)``.  Returned directions have unit length.
    """
    n = np.asarray(wall_normal, dtype=float)
    n = n / np.linalg.norm(n)
    u = np.asarray(wall_up_vector, dtype=float)
    u = u - np.dot(u, n) * n
    u = u / np.linalg.norm(u)
    rot = np.stack([u, np.cross(n, u), n], axis=1)
    d = np.asarray(directions, dtype=float).reshape(-1, 3) @ rot.T
    return d / np.linalg.norm(d, axis=-1, keepdims=True)
