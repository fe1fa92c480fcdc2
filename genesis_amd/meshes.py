"""Collision meshes: convex hulls, their mass properties, and the support-field direction grid of the narrowphase.

* `convex_hull_vertices` keeps the vertices of the convex hull of a mesh (the reference convexifies collision meshes
  with trimesh's hull, genesis/utils/mesh.py:299-475; both are qhull). Vertices stay in file order.
* `hull_mass_properties`: mass, centre of mass and inertia about it of the solid hull at a density (signed
  tetrahedra from the hull's triangles against an interior point), as rigid_link.py:125-188 does with trimesh's
  mass properties of the collision geometry.
* `support_field`: for every direction of the 180 x 180 (theta, phi) grid of support_field_decomp.py:24-94, the
  vertex with the largest projection (first on ties), as float32 positions and vertex indices; the device looks up
  four neighbouring cells per query (support_field_decomp.py:145-186).
"""
from __future__ import annotations

import math

import numpy as np

SUPPORT_RES = 180  # support_field_decomp.py:20


def convex_hull_vertices(verts: np.ndarray) -> np.ndarray:
    from scipy.spatial import ConvexHull
    v = np.asarray(verts, np.float64)
    hull = ConvexHull(v)
    return v[np.sort(hull.vertices)]


def hull_triangles(verts: np.ndarray):
    """(n, 3) triangle vertex indices of the convex hull of `verts` (all of them hull vertices), outward-facing."""
    from scipy.spatial import ConvexHull
    v = np.asarray(verts, np.float64)
    hull = ConvexHull(v)
    tri = hull.simplices.copy()
    c = v.mean(0)
    for k, t in enumerate(tri):
        n = np.cross(v[t[1]] - v[t[0]], v[t[2]] - v[t[0]])
        if n @ (v[t[0]] - c) < 0:
            tri[k] = t[[0, 2, 1]]
    return tri


def hull_mass_properties(verts: np.ndarray, rho: float):
    """(mass, centre of mass, inertia about the centre of mass) of the solid convex hull of `verts` at density
    rho: the sum over hull triangles of the tetrahedra (triangle, interior point), with the closed-form second
    moments of a tetrahedron."""
    v = np.asarray(verts, np.float64)
    tri = hull_triangles(v)
    o = v.mean(0)
    vol, first, second = 0.0, np.zeros(3), np.zeros((3, 3))
    for t in tri:
        a, b, c = v[t[0]] - o, v[t[1]] - o, v[t[2]] - o
        dv = np.dot(a, np.cross(b, c)) / 6.0
        vol += dv
        first += dv * (a + b + c) / 4.0
        # integral of x x^T over the tetrahedron (0, a, b, c): dv / 20 (sum_i p_i p_i^T + s s^T), s = a + b + c
        s = a + b + c
        second += dv / 20.0 * (np.outer(a, a) + np.outer(b, b) + np.outer(c, c) + np.outer(s, s))
    mass = rho * vol
    com_o = first / vol
    cov = rho * second - mass * np.outer(com_o, com_o)  # second moment about the centre of mass
    inertia = np.trace(cov) * np.eye(3) - cov
    return mass, com_o + o, inertia


def direction_grid(res: int = SUPPORT_RES) -> np.ndarray:
    """(res * res, 3) unit directions, theta-major: cell i * res + j has theta = 2 pi i / res - pi and
    phi = pi j / res (support_field_decomp.py:24-37)."""
    theta = np.arange(res) / res * 2 * math.pi - math.pi
    phi = np.arange(res) / res * math.pi
    x = np.sin(phi)[None, :] * np.cos(theta)[:, None]
    y = np.sin(phi)[None, :] * np.sin(theta)[:, None]
    z = np.broadcast_to(np.cos(phi)[None, :], (res, res))
    return np.stack([x, y, z], -1).reshape(-1, 3)


def support_field(verts32: np.ndarray, res: int = SUPPORT_RES):
    """Support vertex of every grid direction for one geom's float32 vertices: (positions (res*res, 3) float32,
    vertex indices (res*res,) int32); np.argmax keeps the first maximal vertex."""
    d = direction_grid(res)
    p = np.asarray(verts32, np.float32).astype(np.float64)
    idx = np.empty(len(d), np.int64)
    step = max(1, int(5e8 // max(len(p), 1)))
    for i in range(0, len(d), step):
        idx[i:i + step] = np.argmax(d[i:i + step] @ p.T, axis=1)
    return np.asarray(verts32, np.float32)[idx], idx.astype(np.int32)
