"""Convex-hull geometry from simulation points (the GPU form of the
reference's SimulationGeometry, src/Geometry.cpp:92-132).

The reference builds the geometry of a simulated jet from the (z, r) points
of a field file: 36 points on the circle of radius r at height z for every
row, their 3-D convex hull (CGAL::convex_hull_3), an AABB tree over its
facets (src/main.cpp:873-897, coordinates in parsec).  Here the hull is
computed once on the host (scipy's Qhull) and handed to the kernels as face
planes: a ray's segment is the line's entry and exit through the planes
(jsx_hull_upload, JSX_GEOM_CONVEX_HULL).
"""
import math

import numpy as np

PC = 3.0856775814671913e+18


def ring_points(z_cm, r_cm, n_circle=36):
    """Points of src/main.cpp:883-892 in parsec: for every (z, r) row,
    (r sin(2 pi j / n), r cos(2 pi j / n), z) for j < n_circle."""
    z = np.asarray(z_cm, dtype=np.float64).reshape(-1) / PC
    r = np.asarray(r_cm, dtype=np.float64).reshape(-1) / PC
    j = np.arange(n_circle)
    ang = j * 2 * math.pi / n_circle
    x = r[:, None] * np.sin(ang)[None, :]
    y = r[:, None] * np.cos(ang)[None, :]
    zz = np.broadcast_to(z[:, None], x.shape)
    return np.stack([x.ravel(), y.ravel(), zz.ravel()], axis=1)


def planes_from_points(points_pc, decimals=12):
    """Face planes [n, 4] (unit outward normal, h in cm) of the convex hull of
    points given in parsec; coplanar facets (Qhull triangulates faces) appear
    once."""
    from scipy.spatial import ConvexHull
    hull = ConvexHull(np.asarray(points_pc, dtype=np.float64))
    eq = hull.equations                      # normal . x + offset <= 0 inside
    scale = np.max(np.abs(points_pc)) or 1.0
    key = np.round(np.concatenate([eq[:, :3], eq[:, 3:] / scale], axis=1), decimals)
    _, first = np.unique(key, axis=0, return_index=True)
    eq = eq[np.sort(first)]
    return np.concatenate([eq[:, :3], -eq[:, 3:] * PC], axis=1)


def simulation_hull(z_cm, r_cm, n_circle=36):
    """Planes of the reference's simulation geometry for field-file rows
    (z [cm], r [cm]) -- e.g. textio.read_field_text(path)[:, :2].T."""
    return planes_from_points(ring_points(z_cm, r_cm, n_circle))
