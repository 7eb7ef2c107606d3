"""Host setup of the native library against the oracle restatements (CPU).

Integer outputs (permutation, CSR structure, incidence) must be bit-exact;
the factor values follow the reference's operation order and are compared
bitwise too. Known answers come from tests/test_diff_phd_sparse.py:21-67 and
tests/test_diff_phd_hex.py of the reference.
"""

import numpy as np
import pytest

from diffphd_amd import _lib, meshes
from oracle import setup_ref, sparse_ref


def _csr_rows(matrix):
    return [{int(j): float(r[j]) for j in np.flatnonzero(r)} for r in matrix]


def _factor_pair(rows):
    off, col, val = setup_ref.rows_to_full_csr(rows)
    native = _lib.Factor(off, col, val)
    ref = sparse_ref.sparse_inverse(rows)
    return native, ref


def _assert_factor_equal(native, ref):
    np.testing.assert_array_equal(native.permutation, ref["permutation"])
    for a, b in zip(native.s, ref["s"]):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(native.st, ref["st"]):
        np.testing.assert_array_equal(a, b)
    assert native.factorization_nnz == ref["factorization_nnz"]
    assert native.matrix_nnz == ref["matrix_nnz"]


def test_exact_cancellation_known_answer():
    # A = L diag(2, 3, 5) L^T with unit lower L: entry (2, 0) of L^{-1} cancels exactly.
    matrix = np.asarray(((2.0, 2.0, 2.0), (2.0, 5.0, 5.0), (2.0, 5.0, 10.0)))
    native, ref = _factor_pair(_csr_rows(matrix))
    _assert_factor_equal(native, ref)
    np.testing.assert_array_equal(native.permutation, (0, 1, 2))
    np.testing.assert_array_equal(native.s[0], (0, 1, 3, 5))
    np.testing.assert_array_equal(native.s[1], (0, 0, 1, 1, 2))
    np.testing.assert_array_equal(
        native.s[2], (1 / np.sqrt(2), -1 / np.sqrt(3), 1 / np.sqrt(3), -1 / np.sqrt(5), 1 / np.sqrt(5))
    )


def test_grid_fill_and_disconnected_pins_match_dense_solve():
    nx, ny = 8, 7
    size = nx * ny + 3
    rng = np.random.default_rng(519)
    matrix = np.eye(size)
    for y in range(ny):
        for x in range(nx):
            i = y * nx + x
            for j in ([i + 1] if x + 1 < nx else []) + ([i + nx] if y + 1 < ny else []):
                w = rng.uniform(0.1, 5.0)
                matrix[i, i] += w
                matrix[j, j] += w
                matrix[i, j] = matrix[j, i] = -w
    native, ref = _factor_pair(_csr_rows(matrix))
    _assert_factor_equal(native, ref)
    assert native.factorization_nnz > (native.matrix_nnz + size) // 2
    rhs = rng.normal(size=(size, 3))
    x = sparse_ref.apply_csr(native.st, sparse_ref.apply_csr(native.s, rhs))
    np.testing.assert_allclose(x, np.linalg.solve(matrix, rhs), rtol=2e-12, atol=2e-13)
    np.testing.assert_array_equal(x[-3:], rhs[-3:])
    for off, col, val in (native.s, native.st):
        assert off[-1] == len(col) == len(val)
        assert np.all(val != 0.0)
        for i in range(size):
            assert np.all(np.diff(col[off[i] : off[i + 1]]) > 0)


def _tet_rows(k, seed, dt=1e-3):
    vertices, tets, poses = meshes.kuhn_tet_grid(k, k, k, 0.005)
    grads, volumes = setup_ref.tet_gradients(poses)
    rng = np.random.default_rng(seed)
    young = np.exp(rng.uniform(np.log(1e5), np.log(1e7), len(tets)))
    mu, lam = meshes.lame(young, 0.4)
    vbar = volumes.mean()
    wc, wv, wn = setup_ref.element_weights(np.stack([mu, lam], 1).astype(np.float32), "pd", vbar)
    free = vertices[:, 2] > 0
    inertia = 1000.0 * vbar / dt**2
    return vertices, tets, poses, grads, (wc + wv) + wn, free, inertia


def test_tet_geometry_and_assembly_bitwise():
    vertices, tets, poses, grads, weights, free, inertia = _tet_rows(4, 1)
    g_native, v_native = _lib.tet_geometry(poses)
    g_ref, v_ref = setup_ref.tet_gradients(poses)
    np.testing.assert_array_equal(g_native, g_ref)
    np.testing.assert_array_equal(v_native, v_ref)
    rows = setup_ref.assemble_rows(len(vertices), tets, grads, False, weights, free, inertia)
    ref = setup_ref.rows_to_full_csr(rows)
    got = _lib.assemble(len(vertices), tets, grads, False, weights, free, inertia)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("k", [3, 5])
def test_tet_grid_factor_bitwise(k):
    vertices, tets, poses, grads, weights, free, inertia = _tet_rows(k, k)
    rows = setup_ref.assemble_rows(len(vertices), tets, grads, False, weights, free, inertia)
    native, ref = _factor_pair(rows)
    _assert_factor_equal(native, ref)
    dense = np.zeros((len(vertices),) * 2)
    for i, r in enumerate(rows):
        for j, v in r.items():
            dense[i, j] = v
    rhs = np.random.default_rng(0).normal(size=(len(vertices), 3))
    x = sparse_ref.apply_csr(native.st, sparse_ref.apply_csr(native.s, rhs))
    np.testing.assert_allclose(x, np.linalg.solve(dense, rhs), rtol=1e-10, atol=1e-14)


def test_hex_gradients_equal_isoparametric_and_factor_bitwise():
    vertices, cells = meshes.hex_grid(3, 2, 2, 0.016)
    h = setup_ref.hex_setup(vertices, cells)
    native = _lib.hex_gradients(h["dx"])
    np.testing.assert_array_equal(native, h["grads"])
    # Independent isoparametric derivation (HexOracle, tests/test_diff_phd_hex.py:52-64).
    gauss, _ = np.polynomial.legendre.leggauss(2)
    signs = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], dtype=float)
    nodes = cells[0]
    for s, p in enumerate(np.array([[a, b, c] for a in gauss for b in gauss for c in gauss])):
        dn = np.empty((8, 3))
        for axis in range(3):
            rest = [b for b in range(3) if b != axis]
            dn[:, axis] = signs[:, axis] * np.prod(1 + signs[:, rest] * p[rest], axis=1) / 8
        jac = vertices[nodes].T @ dn
        np.testing.assert_allclose(native[s], dn @ np.linalg.inv(jac), rtol=0, atol=1e-12 / h["dx"])
    free = vertices[:, 0] > 0
    weights = np.full(len(cells), 3.0e4 * h["mean_volume"] / 8)
    rows = setup_ref.assemble_rows(len(vertices), cells, h["grads"], True, weights, free, 40.0)
    ref = setup_ref.rows_to_full_csr(rows)
    got = _lib.assemble(len(vertices), cells, h["grads"], True, weights, free, 40.0)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)
    native_f, ref_f = _factor_pair(rows)
    _assert_factor_equal(native_f, ref_f)


def test_incidence_is_the_reference_order():
    vertices, tets, _ = meshes.kuhn_tet_grid(3, 2, 2, 0.1)
    off, ent = _lib.incidence(len(vertices), tets)
    r_off, r_ent = setup_ref.tet_incidence(tets, len(vertices))
    np.testing.assert_array_equal(off, r_off)
    np.testing.assert_array_equal(ent, r_ent)
    # every (element, corner) exactly once, ascending per vertex
    assert sorted(ent.tolist()) == list(range(4 * len(tets)))
    for v in range(len(vertices)):
        seg = ent[off[v] : off[v + 1]]
        assert np.all(np.diff(seg) > 0)
        assert np.all(tets.reshape(-1)[seg] == v)
    hv, cells = meshes.hex_grid(2, 2, 1, 0.1)
    h = setup_ref.hex_setup(hv, cells)
    off8, ent8 = _lib.incidence(len(hv), cells)
    # The reference groups (cell, sample, corner); per cell the corner order is the same.
    for v in range(len(hv)):
        ref_cells = h["inc_ent"][h["inc_off"][v] : h["inc_off"][v + 1]] // 64
        ref_corner = h["inc_ent"][h["inc_off"][v] : h["inc_off"][v + 1]] % 8
        got = ent8[off8[v] : off8[v + 1]]
        np.testing.assert_array_equal(np.repeat(got // 8, 8), ref_cells)
        np.testing.assert_array_equal(np.repeat(got % 8, 8), ref_corner)


def test_kuhn_split_invariants():
    vertices, tets, poses = meshes.kuhn_tet_grid(2, 3, 2, 0.125)
    x = vertices[tets]
    vol = np.einsum("ij,ij->i", x[:, 1] - x[:, 0], np.cross(x[:, 2] - x[:, 0], x[:, 3] - x[:, 0])) / 6.0
    assert np.all(vol > 0)
    np.testing.assert_array_equal(vol.sum(), 2 * 3 * 2 * 0.125**3)  # dyadic cells: exact
    faces = np.sort(tets[:, [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]].reshape(-1, 3), axis=1)
    _, counts = np.unique(faces, axis=0, return_counts=True)
    assert counts.max() == 2  # conforming: interior faces shared by exactly two tets
    boundary = counts[counts == 1].size
    assert boundary == 2 * 2 * (2 * 3 + 2 * 2 + 3 * 2)  # two triangles per boundary square
    np.testing.assert_array_equal(poses.astype(np.float64), np.linalg.inv(np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]], 2)))
