"""CPU restatement of the reference's exact sparse inverse factor (TEST ORACLE ONLY).

This module is test infrastructure. Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg may use it, and only as a checker: the
product never imports it.

It restates, step by step, the ordering and factorization of the reference's
``SparseInverse`` (``src/newton_solver_diff_phd/_src/solvers/diff_phd/sparse.py``):

* ``order_min_degree``      follows ``_minimum_degree``     (sparse.py:18-28)
* ``bfs_levels``            follows ``_distances``          (sparse.py:31-40)
* ``order_dissection``      follows ``_nested_dissection``  (sparse.py:43-70)
* ``check_rows``            follows ``_validate_rows``      (sparse.py:73-95)
* ``ldlt_right_looking``    follows ``_factor_ldlt``        (sparse.py:98-156)
* ``inverse_factor_rows``   follows ``_inverse_rows``       (sparse.py:159-177)
* ``rows_to_csr``           follows ``_csr``                (sparse.py:180-191)

The numeric updates use NumPy element-wise operations in the reference's
order (product, then subtraction), so results are bitwise those of the
reference for the same input rows. Pure-Python loops: small matrices only.

Pinning: the exact 3x3 known answer of ``tests/test_diff_phd_sparse.py:21-35``
(permutation, offsets, columns and 1/sqrt values), dense-solve agreement on a
weighted grid with disconnected rows (``:37-67``), and brute-force checks of
the ordering's defining properties (every vertex once, separators split).
"""

from collections import deque

import numpy as np


def order_min_degree(subset, neighbours):
    """Greedy minimum degree on the induced subgraph, ties broken by index."""
    live = {v: set(neighbours[v]) & subset for v in subset}
    out = []
    while live:
        best = None
        for v in live:
            key = (len(live[v]), v)
            if best is None or key < best:
                best = key
        v = best[1]
        around = live.pop(v)
        out.append(v)
        for w in around:
            live[w].discard(v)
            live[w] |= around - {w}
    return out


def bfs_levels(root, subset, neighbours):
    level = {root: 0}
    todo = deque([root])
    while todo:
        v = todo.popleft()
        for w in neighbours[v]:
            if w in subset and w not in level:
                level[w] = level[v] + 1
                todo.append(w)
    return level


def order_dissection(neighbours):
    """Nested dissection on BFS level sets; returns permutation[position] = vertex."""

    def split(subset):
        if len(subset) <= 24:
            return order_min_degree(subset, neighbours)
        degree = {v: len(neighbours[v] & subset) for v in subset}
        root = min(subset, key=lambda v: (degree[v], v))
        level = bfs_levels(root, subset, neighbours)
        if len(level) < len(subset):
            parts = [set(level)]
            left = subset - parts[0]
            while left:
                piece = set(bfs_levels(min(left), left, neighbours))
                parts.append(piece)
                left = left - piece
            out = []
            for piece in parts:
                out.extend(split(piece))
            return out
        far = max(level, key=lambda v: (level[v], -v))
        level = bfs_levels(far, subset, neighbours)
        counts = np.bincount(np.fromiter(level.values(), dtype=np.int64))
        cut = int(np.searchsorted(np.cumsum(counts), len(subset) / 2))
        low = {v for v, d in level.items() if d < cut}
        high = {v for v, d in level.items() if d > cut}
        middle = subset - low - high
        if not low or not high:
            return order_min_degree(subset, neighbours)
        return split(low) + split(high) + order_min_degree(middle, neighbours)

    return np.asarray(split(set(range(len(neighbours)))), dtype=np.int32)


def check_rows(rows):
    """Drop exact zeros; require finite, positive-diagonal, exactly symmetric rows."""
    size = len(rows)
    kept = []
    for i, row in enumerate(rows):
        clean = {}
        for j, value in row.items():
            if not 0 <= int(j) < size:
                raise ValueError("matrix column indices must be integers in [0, size)")
            value = float(value)
            if not np.isfinite(value):
                raise ValueError("matrix coefficients must be finite")
            if value != 0.0:
                clean[int(j)] = value
        if clean.get(i, 0.0) <= 0.0:
            raise ValueError("matrix must be positive definite with positive diagonal entries")
        kept.append(clean)
    for i, row in enumerate(kept):
        for j, value in row.items():
            if kept[j].get(i, 0.0) != value:
                raise ValueError("matrix_rows must contain both halves of an exactly symmetric matrix")
    return kept


def ldlt_right_looking(rows, perm):
    """Symbolic fill, then right-looking LDL^T; returns (lower dict rows, diagonal)."""
    n = len(rows)
    position = np.empty(n, dtype=np.int64)
    position[perm] = np.arange(n)
    lower_rows = []
    for i in range(n):
        src = rows[int(perm[i])]
        lower_rows.append({int(position[j]): v for j, v in src.items() if position[j] <= i})
    # Symbolic elimination: eliminating k connects every pair of its later rows.
    pattern = [set(r) for r in lower_rows]
    below = [set() for _ in range(n)]
    for i in range(n):
        for j in pattern[i]:
            if j < i:
                below[j].add(i)
    for k in range(n):
        seen = set()
        for i in sorted(below[k]):
            extra = seen - pattern[i]
            pattern[i] |= extra
            for j in extra:
                below[j].add(i)
            seen.add(i)
    cols = [np.array(sorted(p), dtype=np.int64) for p in pattern]
    vals = [np.zeros(len(c)) for c in cols]
    for i in range(n):
        for j, v in lower_rows[i].items():
            vals[i][int(np.searchsorted(cols[i], j))] = v
    col_rows = [[] for _ in range(n)]
    col_pos = [[] for _ in range(n)]
    for i in range(n):
        for p, k in enumerate(cols[i][:-1]):
            col_rows[int(k)].append(i)
            col_pos[int(k)].append(p)
    diag = np.empty(n)
    for k in range(n):
        pivot = vals[k][-1]
        if not np.isfinite(pivot) or pivot <= 0.0:
            raise ValueError(f"matrix is not positive definite: LDLT pivot {k} is {pivot}")
        diag[k] = pivot
        rws = np.array(col_rows[k], dtype=np.int64)
        snapshot = np.array([vals[i][p] for i, p in zip(col_rows[k], col_pos[k])])
        for t, (i, p) in enumerate(zip(col_rows[k], col_pos[k])):
            m = snapshot[t] / pivot
            vals[i][p] = m
            where = np.searchsorted(cols[i], rws[: t + 1])
            vals[i][where] -= m * snapshot[: t + 1]
    lower = [dict(zip(map(int, c[:-1]), v[:-1])) for c, v in zip(cols, vals)]
    return lower, diag


def inverse_factor_rows(lower, diag, perm):
    """Rows of D^{-1/2} L^{-1} P with only exact zeros dropped."""
    n = len(lower)
    row_cols, row_vals, out = [], [], []
    acc = np.zeros(n)
    for i in range(n):
        acc[: i + 1] = 0.0
        acc[i] = 1.0
        for k, coef in sorted(lower[i].items()):
            acc[row_cols[k]] -= coef * row_vals[k]
        nz = np.flatnonzero(acc[: i + 1])
        vals = acc[nz].copy()
        row_cols.append(nz)
        row_vals.append(vals)
        s = 1.0 / np.sqrt(diag[i])
        out.append(dict(zip(map(int, perm[nz]), vals * s)))
    return out


def rows_to_csr(rows):
    off = [0]
    cols, vals = [], []
    for row in rows:
        for j in sorted(row):
            cols.append(j)
            vals.append(row[j])
        off.append(len(cols))
    return np.asarray(off, dtype=np.int32), np.asarray(cols, dtype=np.int32), np.asarray(vals, dtype=np.float64)


def sparse_inverse(rows):
    """Return dict(permutation, s=(off, col, val), st=(off, col, val))."""
    rows = check_rows(rows)
    neighbours = [set(r) - {i} for i, r in enumerate(rows)]
    perm = order_dissection(neighbours)
    lower, diag = ldlt_right_looking(rows, perm)
    s_rows = inverse_factor_rows(lower, diag, perm)
    t_rows = [{} for _ in rows]
    for i, r in enumerate(s_rows):
        for j, v in r.items():
            t_rows[j][i] = v
    return {
        "permutation": perm,
        "s": rows_to_csr(s_rows),
        "st": rows_to_csr(t_rows),
        "factorization_nnz": len(rows) + sum(len(r) for r in lower),
        "matrix_nnz": sum(len(r) for r in rows),
    }


def apply_csr(csr, rhs):
    """Sequential CSR product in stored order (the reference's CPU SpMV)."""
    off, col, val = csr
    out = np.zeros((len(off) - 1,) + rhs.shape[1:])
    for r in range(len(off) - 1):
        acc = np.zeros(rhs.shape[1:])
        for p in range(off[r], off[r + 1]):
            acc = acc + val[p] * rhs[col[p]]
        out[r] = acc
    return out
