"""Frictional boundary contact on the oracle (CPU) against the reference's dense transcription.

Golden rollouts of ``ReferenceStep`` with ``ReferenceContact`` /
``ReferenceMeshContact`` (the reference's tests/reference_oracle.py) on the
scenes of tests/test_diff_phd_contact.py and tests/test_diff_phd_mesh_boundary.py;
each step is fed the golden's own inputs, the multipliers carry over.
"""

import numpy as np
import pytest

from oracle import oracle as O
from scenes import contact_grid, golden, two_blocks

GRID_CASES = ["plane", "stick", "slide", "sphere", "chatter", "restitution"]
MESH_CASES = ["mesh", "mesh_stick", "mesh_restitution"]


def contact_body(name):
    g = golden(f"contact_{name}")
    options = dict(max_iterations=int(g["max_iterations"]), abs_tol=float(g["abs_tol"]), rel_tol=float(g["rel_tol"]))
    if name.startswith("mesh"):
        vertices, tets, poses, materials, free, split, left, right = two_blocks(0.012)
        boundary = ("mesh", (float(g["radius"]), int(g["split"]), g["faces_left"], g["faces_right"]))
    else:
        vertices, tets, poses, materials, free = contact_grid()
        boundary = (str(g["kind"]), tuple(g["params"]))
    body = O.OracleBody(
        "tet4", tets, poses=poses, materials=materials, density=1000.0, inverse_mass=free.astype(np.float32),
        gravity=(0.0, 0.0, -9.81), dt=0.004, frictional_boundary=boundary, frictional_vertices=g["nodes"],
        friction=float(g["friction"]), restitution=float(g["restitution"]), **options,
    )
    return g, body


@pytest.mark.parametrize("name", GRID_CASES + MESH_CASES)
def test_oracle_contact_rollout_matches_reference(name):
    g, body = contact_body(name)
    for t in range(len(g["x"])):
        out = body.forward(0.004, g["q"][t], g["qd"][t])
        assert out["iterations"] == int(g["iterations"][t]), (t, out["iterations"], int(g["iterations"][t]))
        assert (out["status"] == 0) == bool(g["converged"][t])
        if bool(g["converged"][t]):
            np.testing.assert_allclose(out["x"], g["x"][t], rtol=0, atol=1e-9, err_msg=f"step {t}")
            tol = dict(rtol=1e-5, atol=1e-6) if name.startswith("mesh") else dict(rtol=1e-6, atol=1e-9)
            np.testing.assert_allclose(body.contact["multiplier"], g["multiplier"][t], **tol, err_msg=f"step {t}")
    if name != "chatter":
        assert np.count_nonzero(g["multiplier"]) > 0


def woodbury_body(name):
    g = golden(f"contact_{name}")
    vertices, tets, poses, materials, free = contact_grid()
    body = O.OracleBody(
        "tet4", tets, poses=poses, materials=materials, density=1000.0, inverse_mass=free.astype(np.float32),
        gravity=(0.0, 0.0, -9.81), dt=0.004, muscle_directions=g["directions"], muscle_stiffness=4.0e4,
        frictional_boundary=("planar", (0.0, 0.0, 1.0, 0.0)), frictional_vertices=g["nodes"],
        friction=float(g["friction"]), max_iterations=int(g["max_iterations"]), abs_tol=float(g["abs_tol"]),
        rel_tol=float(g["rel_tol"]),
    )
    return g, body


@pytest.mark.parametrize("name", ["woodbury", "woodbury_friction"])
def test_oracle_woodbury_projection_matches_dense_reference(name):
    g, body = woodbury_body(name)
    for t in range(3):
        out = body.forward(0.004, g["q"][t], g["qd"][t], targets=g["targets"])
        assert out["status"] == 0
    np.testing.assert_allclose(out["x"], g["x"], rtol=0, atol=1e-9)
    bw = body.backward(0.004, out["x"], g["gq"], g["gqd"], targets=g["targets"], activations=True)
    assert bw["status"] & 8 == 0
    for key in ("q", "qd", "f", "materials", "activations"):
        ref = g[f"out_{key}"]
        scale = max(np.abs(ref).max(), 1e-30)
        np.testing.assert_allclose(bw[key], ref, rtol=0, atol=1e-5 * scale, err_msg=key)
