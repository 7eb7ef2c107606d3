"""Mesh contact state force on the GPU (MeshContactStateForce, DP/mesh_contact.py).

Against the golden vectors of the reference's NumPy transcription and the
oracle, on the two-body scene of tests/test_diff_phd_mesh_contact.py.
"""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from scenes import golden, tet_model, two_bodies

pytestmark = pytest.mark.gpu

TIGHT = {"max_iterations": 3000, "abs_tol": 1.0e-12, "rel_tol": 1.0e-11}


def _solver(model, **kw):
    from diffphd_amd import SolverDiffPhD

    return SolverDiffPhD(model, density=1000.0, **kw)


def _scene(cuda, gap, requires_grad=False):
    vertices, tets, poses, materials, free, faces_a, faces_b = two_bodies(gap)
    model = tet_model(cuda, tets, poses, materials, free, vertices, (0.0, 0.0, 0.0), requires_grad=requires_grad)
    return model, vertices, faces_a, faces_b


def test_force_matches_reference_transcription(cuda):
    print("[case] mesh contact force", flush=True)
    g = golden("mesh_contact")
    model, vertices, faces_a, faces_b = _scene(cuda, -0.004)
    solver = _solver(model, mesh_contact=(*g["options"], faces_a, faces_b))
    solver.prepare(0.004)
    for trial, (pos, vel, expected) in enumerate(zip(g["positions"], g["velocities"], g["forces"])):
        got = solver.mesh_contact_force(torch.as_tensor(pos, device=cuda), torch.as_tensor(vel, device=cuda))
        got = got.cpu().numpy()
        scale = max(np.abs(expected).max(), 1e-300)
        np.testing.assert_allclose(got, expected, rtol=0, atol=1e-12 * scale, err_msg=f"trial {trial}")
        ref, pairs = O.mesh_contact_force(pos, vel, faces_a, faces_b, *g["options"])
        np.testing.assert_array_equal(got, ref, err_msg=f"trial {trial}")
        assert int(solver.mesh_contact_pairs[0]) == pairs


def test_step_matches_reference_with_state_force_in_the_target(cuda):
    print("[case] mesh contact step", flush=True)
    g = golden("mesh_contact")
    model, vertices, faces_a, faces_b = _scene(cuda, -0.003)
    solver = _solver(model, mesh_contact=(*g["step_options"], faces_a, faces_b), **TIGHT)
    s_in, s_out = model.state(), model.state()
    s_in.particle_qd = torch.as_tensor(g["step_qd"], device=cuda)
    solver.step(s_in, s_out, None, None, 0.004)
    assert solver.read_status() == 0
    x = solver.solution().cpu().numpy()
    np.testing.assert_allclose(x, g["step_x"], rtol=0, atol=1e-10 * np.abs(g["step_x"]).max())
    assert int(solver.mesh_contact_pairs[0]) > 0


def test_gradients_hold_the_state_force_fixed(cuda):
    print("[case] mesh contact gradients", flush=True)
    g = golden("mesh_contact")
    weights = torch.as_tensor(np.random.default_rng(5).standard_normal((40, 3)), device=cuda)
    results = []
    for frozen in (False, True):
        model, vertices, faces_a, faces_b = _scene(cuda, -0.003, requires_grad=True)
        options = (*g["step_options"], faces_a, faces_b)
        solver = _solver(model, mesh_contact=None if frozen else options, **TIGHT)
        s_in, s_out = model.state(), model.state()
        if frozen:
            q = vertices.astype(np.float64)
            force, _ = O.mesh_contact_force(q, np.zeros_like(q), faces_a, faces_b, *g["step_options"])
            s_in.particle_f = torch.as_tensor(force.astype(np.float32), device=cuda)
        for name in ("particle_q", "particle_qd", "particle_f"):
            getattr(s_in, name).requires_grad_(True)
        solver.step(s_in, s_out, None, None, 0.004)
        loss = (s_out.particle_q.to(torch.float64) * weights).sum()
        loss.backward()
        assert solver.read_status() == 0
        results.append([s_in.particle_q.grad.cpu().numpy(), s_in.particle_qd.grad.cpu().numpy()])
    for actual, expected in zip(results[0], results[1]):
        np.testing.assert_allclose(actual, expected, rtol=0, atol=1e-6 * max(np.abs(expected).max(), 1.0))


def test_invalid_mesh_contact_is_rejected(cuda):
    model, vertices, faces_a, faces_b = _scene(cuda, 0.0)
    n = len(vertices)
    for options in (
        (0.01, 5.0e3, 5.0e4, 50.0, 0.5, faces_a),
        (-0.01, 5.0e3, 5.0e4, 50.0, 0.5, faces_a, faces_b),
        (0.01, 5.0e3, 5.0e4, 50.0, 0.5, faces_a, np.zeros((0, 3), dtype=np.int32)),
        (0.01, 5.0e3, 5.0e4, 50.0, 0.5, faces_a, faces_b + n),
    ):
        with pytest.raises(ValueError):
            _solver(model, mesh_contact=options)


def test_broad_phase_keeps_every_pair_of_a_wide_scene(cuda):
    """A 12 x 12 slab under a 4 x 4 plate: most of the slab's vertices lie outside the plate's
    widened face boxes and most grid cells hold no pair, so the broad phase skips them; forces and
    pair counts must equal the oracle's brute-force evaluation on perturbed states."""
    from diffphd_amd import meshes, parity

    print("[case] mesh contact broad phase", flush=True)
    ra, ta, _, _, da = parity.grid(height=0.0, dims=(12, 12, 1), cell=0.01, fix_bottom=True)
    rb, tb, _, _, db = parity.grid(height=0.0, dims=(4, 4, 1), cell=0.01, origin=(0.035, 0.045, 0.0105))
    vertices = np.concatenate([ra, rb]).astype(np.float32)
    tets = np.concatenate([ta, tb + len(ra)])
    free = ~np.concatenate([da, db])
    faces = parity.boundary_faces(tets)
    in_b = faces >= len(ra)
    faces_a, faces_b = faces[~in_b.any(axis=1)], faces[in_b.all(axis=1)]
    assert len(faces_a) > 4 * 64 and len(faces_b) > 64
    materials = np.tile(np.array([[2.0e4, 5.0e4, 0.0]], np.float32), (len(tets), 1))
    model = tet_model(cuda, tets, meshes.tet_poses(vertices, tets), materials, free, vertices, (0.0, 0.0, 0.0))
    options = (0.004, 60.0, 150.0, 3.0, 0.5)
    solver = _solver(model, mesh_contact=(*options, faces_a, faces_b))
    solver.prepare(0.004)
    rng = np.random.default_rng(21)
    for trial in range(4):
        pos = vertices.astype(np.float64) + 0.001 * rng.standard_normal(vertices.shape)
        vel = 0.1 * rng.standard_normal(vertices.shape)
        got = solver.mesh_contact_force(torch.as_tensor(pos, device=cuda), torch.as_tensor(vel, device=cuda))
        ref, pairs = O.mesh_contact_force(pos, vel, faces_a, faces_b, *options)
        assert pairs > 0
        np.testing.assert_array_equal(got.cpu().numpy(), ref, err_msg=f"trial {trial}")
        ab, ba = solver.mesh_contact_pair_counts
        assert ab + ba == pairs == int(solver.mesh_contact_pairs[0])


def _slab_under(cuda, rb, tb, db):
    """The 12 x 12 slab of 1 cm cells (bottom pinned) with body B above it; both sides' boundary faces."""
    from diffphd_amd import meshes, parity

    ra, ta, _, _, da = parity.grid(height=0.0, dims=(12, 12, 1), cell=0.01, fix_bottom=True)
    vertices = np.concatenate([ra, rb]).astype(np.float32)
    tets = np.concatenate([ta, tb + len(ra)])
    free = ~np.concatenate([da, db])
    faces = parity.boundary_faces(tets)
    in_b = faces >= len(ra)
    faces_a, faces_b = faces[~in_b.any(axis=1)], faces[in_b.all(axis=1)]
    materials = np.tile(np.array([[2.0e4, 5.0e4, 0.0]], np.float32), (len(tets), 1))
    model = tet_model(cuda, tets, meshes.tet_poses(vertices, tets), materials, free, vertices, (0.0, 0.0, 0.0))
    return model, vertices, faces_a, faces_b


def _compare_forces(cuda, solver, pos, vel, faces_a, faces_b, options, label):
    got = solver.mesh_contact_force(torch.as_tensor(pos, device=cuda), torch.as_tensor(vel, device=cuda))
    got = got.cpu().numpy()
    ref, pairs = O.mesh_contact_force(pos, vel, faces_a, faces_b, *options)
    # The contact kernels round as the oracle does (no fused multiply-adds, the same order of every sum):
    # equal bit for bit, non-finite entries included.
    np.testing.assert_array_equal(got, ref, err_msg=label)
    ab, ba = solver.mesh_contact_pair_counts
    assert ab + ba == pairs == int(solver.mesh_contact_pairs[0]), label
    return pairs


def test_broad_phase_with_faces_far_larger_than_the_other_side(cuda):
    """One 12 cm Kuhn cube (six tets, faces 12 cm across) resting 0.5 mm above the slab of 1 cm cells: the
    grid over the cube's faces has cells as wide as the cube, so the slab's 169 top vertices share a few
    cells (long per-cell lists); forces and pair counts must still equal the oracle's brute force. (The
    large faces also amplify rounding in the closest-point arithmetic: a fused multiply-add anywhere in it
    moves forces by 1e-11 of their size.)"""
    from diffphd_amd import parity

    print("[case] mesh contact broad phase, large faces over a fine slab", flush=True)
    rb, tb, _, _, db = parity.grid(height=0.0, dims=(1, 1, 1), cell=0.12, origin=(0.0, 0.0, 0.0105))
    model, vertices, faces_a, faces_b = _slab_under(cuda, rb, tb, db)
    options = (0.004, 60.0, 150.0, 3.0, 0.5)
    solver = _solver(model, mesh_contact=(*options, faces_a, faces_b))
    solver.prepare(0.004)
    rng = np.random.default_rng(33)
    for trial in range(3):
        pos = vertices.astype(np.float64) + 0.0005 * rng.standard_normal(vertices.shape)
        vel = 0.1 * rng.standard_normal(vertices.shape)
        assert _compare_forces(cuda, solver, pos, vel, faces_a, faces_b, options, f"trial {trial}") > 100


def test_broad_phase_with_a_non_finite_position_tests_every_pair(cuda):
    """A NaN coordinate turns the grid off (every pair is tested, as the reference does): forces, their
    non-finite entries and the pair counts equal the oracle's on the wide scene."""
    from diffphd_amd import parity

    print("[case] mesh contact broad phase, non-finite position", flush=True)
    rb, tb, _, _, db = parity.grid(height=0.0, dims=(4, 4, 1), cell=0.01, origin=(0.035, 0.045, 0.0105))
    model, vertices, faces_a, faces_b = _slab_under(cuda, rb, tb, db)
    options = (0.004, 60.0, 150.0, 3.0, 0.5)
    solver = _solver(model, mesh_contact=(*options, faces_a, faces_b))
    solver.prepare(0.004)
    rng = np.random.default_rng(8)
    pos = vertices.astype(np.float64) + 0.001 * rng.standard_normal(vertices.shape)
    vel = 0.1 * rng.standard_normal(vertices.shape)
    _compare_forces(cuda, solver, pos, vel, faces_a, faces_b, options, "finite")
    for which, side in ((int(faces_a[7, 1]), "a slab vertex"), (int(faces_b[3, 2]), "a plate vertex")):
        bad = pos.copy()
        bad[which, 2] = np.nan
        _compare_forces(cuda, solver, bad, vel, faces_a, faces_b, options, side)
