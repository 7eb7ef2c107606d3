"""Structured Tet4 / Hex8 meshes and the benchmark scenes of this solver path.

The reference obtains Tet4 grids from Newton's ``ModelBuilder.add_soft_grid``
(third party, not vendored) and Hex8 grids from ``tools/bake_cantilever.py``.
Here the cube split is defined explicitly: the Kuhn (Freudenthal) split of
every cube into 6 positively oriented tetrahedra sharing the main diagonal,
which is conforming across cubes. Vertex ids are x-major,
``i * (ny + 1) * (nz + 1) + j * (nz + 1) + k``, as in
``tools/bake_cantilever.py:26-29``.
"""

from __future__ import annotations

from itertools import permutations

import numpy as np

# Kuhn simplices: a path from corner 000 to corner 111 through one axis at a time.
_AXIS = np.eye(3, dtype=np.int64)


def _kuhn_local():
    tets = []
    for order in permutations(range(3)):
        path = [np.zeros(3, dtype=np.int64)]
        for axis in order:
            path.append(path[-1] + _AXIS[axis])
        corners = [int(p[0] * 4 + p[1] * 2 + p[2]) for p in path]
        edges = np.array([path[1] - path[0], path[2] - path[0], path[3] - path[0]], dtype=np.float64)
        if np.linalg.det(edges) < 0:
            corners[1], corners[2] = corners[2], corners[1]
        tets.append(corners)
    return np.asarray(tets, dtype=np.int64)


KUHN_TETS = _kuhn_local()


def grid_vertices(nx, ny, nz, cell, origin=(0.0, 0.0, 0.0)):
    cell = np.broadcast_to(np.asarray(cell, dtype=np.float64), (3,))
    idx = np.array([(i, j, k) for i in range(nx + 1) for j in range(ny + 1) for k in range(nz + 1)], dtype=np.float64)
    return np.asarray(origin, dtype=np.float64) + idx * cell


def _node(i, j, k, ny, nz):
    return i * (ny + 1) * (nz + 1) + j * (nz + 1) + k


def _cube_corners(nx, ny, nz):
    """[cells, 8] vertex ids in binary (x, y, z) corner order, cells x-major."""
    cells = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                cells.append([_node(i + a, j + b, k + c, ny, nz) for a in (0, 1) for b in (0, 1) for c in (0, 1)])
    return np.asarray(cells, dtype=np.int32)


def hex_grid(nx, ny, nz, cell, origin=(0.0, 0.0, 0.0)):
    """Vertices [N,3] float64 and Hex8 cells [E,8] in binary corner order."""
    return grid_vertices(nx, ny, nz, cell, origin), _cube_corners(nx, ny, nz)


def tet_poses(vertices, tets):
    """Inverse rest edge matrices Dm^{-1} (Dm columns x1-x0, x2-x0, x3-x0), float32 like Newton's tet_poses."""
    x = np.asarray(vertices, dtype=np.float64)[tets]
    dm = np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]], axis=2)
    return np.linalg.inv(dm).astype(np.float32)


def kuhn_tet_grid(nx, ny, nz, cell, origin=(0.0, 0.0, 0.0)):
    """Vertices [N,3], tets [6*cells, 4] (positive orientation) and float32 poses."""
    vertices = grid_vertices(nx, ny, nz, cell, origin)
    cubes = _cube_corners(nx, ny, nz)
    tets = cubes[:, KUHN_TETS].reshape(-1, 4).astype(np.int32)
    return vertices, tets, tet_poses(vertices, tets)


def surface_vertices(tets):
    """Vertices of faces used by exactly one tetrahedron, ascending (DP/contact.py:676-681)."""
    tets = np.asarray(tets)
    faces = tets[:, [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]].reshape(-1, 3)
    keys, counts = np.unique(np.sort(faces, axis=1), axis=0, return_counts=True)
    return np.unique(keys[counts == 1])


def twisted_positions(rest, angle):
    """Cross-sections rotated about their centre linearly from 0 at x_min to ``angle`` at x_max.

    Follows ``twisted_positions`` of ``examples/_diff_phd_cantilever.py:64-77``;
    the root section keeps its rest position exactly.
    """
    rest = np.asarray(rest, dtype=np.float64)
    low, high = rest.min(axis=0), rest.max(axis=0)
    theta = float(angle) * (rest[:, 0] - low[0]) / (high[0] - low[0])
    center = 0.5 * (low[1:] + high[1:])
    offset = rest[:, 1:] - center
    q = rest.copy()
    q[:, 1] = center[0] + np.cos(theta) * offset[:, 0] - np.sin(theta) * offset[:, 1]
    q[:, 2] = center[1] + np.sin(theta) * offset[:, 0] + np.cos(theta) * offset[:, 1]
    root = rest[:, 0] == low[0]
    q[root] = rest[root]
    return q


def lame(young, poisson):
    young = np.asarray(young, dtype=np.float64)
    poisson = np.asarray(poisson, dtype=np.float64)
    mu = young / (2.0 * (1.0 + poisson))
    lam = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson))
    return mu, lam


def icosphere(radius, subdivisions, center=(0.0, 0.0, 0.0)):
    """Triangulated sphere: an icosahedron with every face split in four ``subdivisions`` times.

    Returns vertices [V,3] float64 on the sphere and outward faces [F,3] int32.
    """
    t = (1.0 + 5.0**0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    points = [np.asarray(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    for _ in range(subdivisions):
        midpoint = {}

        def middle(a, b):
            key = (min(a, b), max(a, b))
            if key not in midpoint:
                m = points[a] + points[b]
                points.append(m / np.linalg.norm(m))
                midpoint[key] = len(points) - 1
            return midpoint[key]

        split = []
        for a, b, c in faces:
            ab, bc, ca = middle(a, b), middle(b, c), middle(c, a)
            split += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = split
    vertices = np.asarray(center, dtype=np.float64) + radius * np.asarray(points)
    return vertices, np.asarray(faces, dtype=np.int32)
