"""NumPy restatement of the reference's per-solver setup (TEST ORACLE ONLY).

Test infrastructure: only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg use it, as the checker.

Restates (paths under ``src/newton_solver_diff_phd/_src/solvers/diff_phd/``):

* ``tet_gradients``      solver.py:214-223 (gradients, volumes, mean volume)
* ``hex_setup``          hex.py:30-92 (validation, Gauss gradients, incidence)
* ``tet_incidence``      kernels.py:34-72 (count / fill / insertion sort)
* ``element_weights``    kernels.py:112-133 and hex.py:95-112
* ``assemble_rows``      solver.py:521-545 (row dictionaries of A)

Pinned by: the closed forms of the reference tests (hex mass 1.0 for 0.1 m
cubes at density 1000, ``tests/test_diff_phd_hex.py:117-118``; cantilever
invariants ``tests/test_cantilever.py:19-48``), the isoparametric Gauss
gradients of ``HexOracle`` (``tests/test_diff_phd_hex.py:52-64``), and brute
force on tiny meshes (incidence lists every (element, corner) once, sorted).
Dot products are written out as ``(a0*b0 + a1*b1) + a2*b2``.
"""

import numpy as np


def tet_gradients(poses):
    p = np.asarray(poses, dtype=np.float32).astype(np.float64)
    g0 = -((p[:, 0, :] + p[:, 1, :]) + p[:, 2, :])
    grads = np.concatenate([g0[:, None, :], p], axis=1)
    det = (
        p[:, 0, 0] * (p[:, 1, 1] * p[:, 2, 2] - p[:, 1, 2] * p[:, 2, 1])
        - p[:, 0, 1] * (p[:, 1, 0] * p[:, 2, 2] - p[:, 1, 2] * p[:, 2, 0])
        + p[:, 0, 2] * (p[:, 1, 0] * p[:, 2, 1] - p[:, 1, 1] * p[:, 2, 0])
    )
    volumes = np.abs(1.0 / det) / 6.0
    return grads, volumes


def hex_setup(vertices, cells):
    vertices = np.asarray(vertices, dtype=np.float64)
    cells = np.asarray(cells, dtype=np.int64)
    rest = vertices[cells]
    e = rest[:, [1, 2, 4]] - rest[:, :1]
    cross = np.stack(
        [
            e[:, 1, 1] * e[:, 2, 2] - e[:, 1, 2] * e[:, 2, 1],
            e[:, 1, 2] * e[:, 2, 0] - e[:, 1, 0] * e[:, 2, 2],
            e[:, 1, 0] * e[:, 2, 1] - e[:, 1, 1] * e[:, 2, 0],
        ],
        axis=1,
    )
    volumes = np.abs((e[:, 0, 0] * cross[:, 0] + e[:, 0, 1] * cross[:, 1]) + e[:, 0, 2] * cross[:, 2])
    total = 0.0
    for v in volumes:
        total += float(v)
    mean = total / len(cells)
    dx = mean ** (1.0 / 3.0)
    corners = np.array([[a, b, c] for a in range(2) for b in range(2) for c in range(2)], dtype=np.float64)
    if not np.allclose(rest, rest[:, :1] + dx * corners[None], rtol=0.0, atol=1.0e-9 * dx):
        raise ValueError("Reference Hex8 requires equally sized axis-aligned cubes in binary corner order")
    inv = 1.0 / dx
    r = 1.0 / np.sqrt(3.0)
    grads = np.empty((8, 8, 3))
    for s, c in enumerate(corners):
        nx, ny, nz = (((c - 0.5) * r + 0.5) * dx) * inv
        cx, cy, cz = 1.0 - nx, 1.0 - ny, 1.0 - nz
        grads[s] = [
            [-inv * cy * cz, cx * -inv * cz, cx * cy * -inv],
            [-inv * cy * nz, cx * -inv * nz, cx * cy * inv],
            [-inv * ny * cz, cx * inv * cz, cx * ny * -inv],
            [-inv * ny * nz, cx * inv * nz, cx * ny * inv],
            [inv * cy * cz, nx * -inv * cz, nx * cy * -inv],
            [inv * cy * nz, nx * -inv * nz, nx * cy * inv],
            [inv * ny * cz, nx * inv * cz, nx * ny * -inv],
            [inv * ny * nz, nx * inv * nz, nx * ny * inv],
        ]
    # Stable grouping of (element, sample, corner) entries by vertex.
    by_entry = np.repeat(cells, 8, axis=0).ravel()
    counts = np.bincount(by_entry, minlength=len(vertices))
    offsets = np.concatenate(([0], np.cumsum(counts))).astype(np.int32)
    entries = np.argsort(by_entry, kind="stable").astype(np.int32)
    return {"grads": grads, "volumes": volumes, "mean_volume": mean, "dx": dx, "inc_off": offsets, "inc_ent": entries}


def tet_incidence(tets, n):
    tets = np.asarray(tets)
    lists = [[] for _ in range(n)]
    for e, tet in enumerate(tets):
        for k, v in enumerate(tet):
            lists[int(v)].append(4 * e + k)
    offsets = np.zeros(n + 1, dtype=np.int32)
    entries = []
    for v in range(n):
        lst = lists[v]
        # insertion sort: the reference's fixed summation order
        for a in range(1, len(lst)):
            value, b = lst[a], a - 1
            while b >= 0 and lst[b] > value:
                lst[b + 1] = lst[b]
                b -= 1
            lst[b + 1] = value
        entries.extend(lst)
        offsets[v + 1] = len(entries)
    return offsets, np.asarray(entries, dtype=np.int32)


def element_weights(materials, material, sample_volume):
    m = np.asarray(materials, dtype=np.float64)
    mu, lam = m[:, 0], m[:, 1]
    zero = np.zeros(len(m))
    if material == "pd":
        return 2.0 * mu * sample_volume, lam * sample_volume, zero
    return zero, zero, 2.0 * mu * sample_volume


def assemble_rows(n, cells, grads, shared, weights, free, inertia, fibre_dir=None, fibre_w=None, planes=None):
    """Row dictionaries of A in the reference's accumulation order."""
    rows = [{i: inertia if free[i] else 1.0} for i in range(n)]
    for e, nodes in enumerate(cells):
        samples = grads if shared else grads[e][None]
        for g in samples:
            k = len(nodes)
            block = np.empty((k, k))
            for a in range(k):
                for b in range(k):
                    block[a, b] = weights[e] * ((g[a, 0] * g[b, 0] + g[a, 1] * g[b, 1]) + g[a, 2] * g[b, 2])
            if fibre_w is not None and fibre_w[e] > 0:
                m = fibre_dir[e]
                fib = (g[:, 0] * m[0] + g[:, 1] * m[1]) + g[:, 2] * m[2]
                block = block + fibre_w[e] * np.outer(fib, fib)
            for a in range(k):
                i = int(nodes[a])
                if not free[i]:
                    continue
                for b in range(a, k):
                    j = int(nodes[b])
                    if not free[j]:
                        continue
                    value = float(block[a, b])
                    rows[i][j] = rows[i].get(j, 0.0) + value
                    if i != j:
                        rows[j][i] = rows[j].get(i, 0.0) + value
    if planes is not None:
        for i, stiffness in zip(*planes):
            if free[i]:
                rows[i][i] += float(stiffness)
    return rows


def rows_to_full_csr(rows):
    off = [0]
    col, val = [], []
    for r in rows:
        for j in sorted(r):
            col.append(j)
            val.append(r[j])
        off.append(len(col))
    return np.asarray(off, np.int32), np.asarray(col, np.int32), np.asarray(val, np.float64)


def surface_vertices(tets):
    """Vertices of faces that belong to exactly one tetrahedron, ascending (contact.py:679-684)."""
    tets = np.asarray(tets)
    faces = np.concatenate([tets[:, [0, 1, 2]], tets[:, [0, 1, 3]], tets[:, [0, 2, 3]], tets[:, [1, 2, 3]]])
    unique, counts = np.unique(np.sort(faces, axis=1), axis=0, return_counts=True)
    return np.unique(unique[counts == 1]).astype(np.int32)
