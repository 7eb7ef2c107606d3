"""Mesh contact state force on the oracle (CPU) against the reference's NumPy transcription.

Golden vectors: ``mesh_contact_force`` and ``ReferenceStep.step`` of the
reference's tests/reference_oracle.py on the two-body scene of
tests/test_diff_phd_mesh_contact.py (tests/golden/make_golden.py).
"""

import numpy as np

from oracle import oracle as O
from scenes import golden, two_bodies


def test_oracle_force_matches_reference_transcription():
    g = golden("mesh_contact")
    _, _, _, _, _, faces_a, faces_b = two_bodies(-0.004)
    for pos, vel, expected in zip(g["positions"], g["velocities"], g["forces"]):
        got, pairs = O.mesh_contact_force(pos, vel, faces_a, faces_b, *g["options"])
        scale = max(np.abs(expected).max(), 1e-300)
        np.testing.assert_allclose(got, expected, rtol=0, atol=1e-12 * scale)
        assert (pairs > 0) == (np.abs(expected).max() > 0)


def test_oracle_step_matches_reference_with_the_force_in_the_target():
    g = golden("mesh_contact")
    vertices, tets, poses, materials, free, faces_a, faces_b = two_bodies(-0.003)
    body = O.OracleBody(
        "tet4", tets, poses=poses, materials=materials, density=1000.0, inverse_mass=free.astype(np.float32),
        dt=0.004, max_iterations=3000, abs_tol=1e-12, rel_tol=1e-11,
        mesh_contact=(*g["step_options"], faces_a, faces_b),
    )
    out = body.forward(0.004, vertices, g["step_qd"])
    assert out["status"] == 0 and bool(g["step_converged"])
    np.testing.assert_allclose(out["x"], g["step_x"], rtol=0, atol=1e-10 * np.abs(g["step_x"]).max())
    assert abs(out["iterations"] - int(g["step_iterations"])) <= 1
