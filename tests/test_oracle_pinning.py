"""Pinning of the C oracle (CPU): golden vectors of the reference's NumPy oracle and closed forms."""

import re
from pathlib import Path

import numpy as np
import pytest

from diffphd_amd import meshes
from oracle import oracle as O
from scenes import cube_body, golden

ROOT = Path(__file__).resolve().parents[1]


def _body(g, material, **kw):
    extra = {}
    if "directions" in g:
        extra = dict(muscle_directions=g["directions"], muscle_stiffness=g["stiffness"])
    return O.OracleBody(
        "tet4", g["tets"], poses=g["poses"], materials=g["materials"], density=float(g["density"]),
        inverse_mass=g["free"].astype(np.float32), gravity=g["gravity"], dt=float(g["dt"]), material=material,
        max_iterations=int(g["max_iterations"]), abs_tol=float(g["abs_tol"]), rel_tol=float(g["rel_tol"]),
        **extra, **kw,
    )


@pytest.mark.parametrize(
    "name",
    [
        "tet_fwd_fixed_pd",
        "tet_fwd_fixed_nh",
        "tet_fwd_conv_c10_pd",
        "tet_fwd_conv_c10_nh",
        "tet_fwd_conv_c100_pd",
        "tet_fwd_conv_c100_nh",
        "tet_fwd_muscle_pd",
    ],
)
def test_forward_matches_golden(name):
    g = golden(name)
    body = _body(g, "pd" if name.endswith("_pd") else "neo_hookean")
    out = body.forward(float(g["dt"]), g["q"], g["qd"], g["f"], targets=g.get("targets"))
    assert out["iterations"] == int(g["iterations"])
    assert out["status"] == (0 if bool(g["converged"]) else 1)
    np.testing.assert_allclose(out["x"], g["x"], rtol=0, atol=1e-10 * np.abs(g["x"]).max())


@pytest.mark.parametrize("name", ["tet_bwd_pd_c2", "tet_bwd_pd_qd", "tet_bwd_nh", "tet_bwd_muscle_pd"])
def test_backward_matches_golden(name):
    g = golden(name)
    material = str(g["material"])
    body = _body(g, material)
    targets = g.get("targets")
    fw = body.forward(float(g["dt"]), g["q"], g["qd"], g["f"], targets=targets)
    assert fw["iterations"] == int(g["iterations"])
    bw = body.backward(float(g["dt"]), g["x"], g["gq"], g["gqd"], targets=targets, activations=targets is not None)
    if targets is not None:
        # d/du of the contraction u = 1 - target; zero where the fibre is disabled.
        ref = g["out_activations"]
        assert np.abs(ref).max() > 0 and np.all(bw["activations"][g["stiffness"] == 0] == 0)
        np.testing.assert_allclose(bw["activations"], ref, rtol=0, atol=2e-5 * np.abs(ref).max())
    assert bw["status"] & 8 == 0
    free = g["free"]
    # The golden mu comes from finite-difference projection Jacobians (eps 1e-6).
    np.testing.assert_allclose(bw["mu"], g["out_mu"], rtol=0, atol=2e-5 * np.abs(g["out_mu"]).max())
    for key in ("qd", "f"):
        ref = g[f"out_{key}"]
        np.testing.assert_allclose(bw[key], ref, rtol=0, atol=2e-5 * np.abs(ref).max())
    # The dense oracle has no Dirichlet gradient; the solver adds one on pinned rows.
    np.testing.assert_allclose(bw["q"][free], g["out_q"][free], rtol=0, atol=2e-5 * np.abs(g["out_q"]).max())
    if material == "pd":
        np.testing.assert_allclose(bw["materials"], g["out_materials"], rtol=0, atol=2e-5 * np.abs(g["out_materials"]).max())


def test_projections_match_golden():
    g = golden("projections")
    mu, lame = float(g["mu"]), float(g["lame"])
    out = O.projections(g["f"], g["directions"], mu, lame, 2 * mu, 0.8)
    np.testing.assert_allclose(out["sigma"], g["sigma"], rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(out["reconstruction"], g["f"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(out["rotation"], g["rotation"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(out["volume"], g["volume"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(out["neo_hookean"], g["neo_hookean"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(out["fibre"], g["fibre"], rtol=0, atol=1e-13)
    for r, m in zip(out["rotation"], g["f"]):
        assert abs(np.linalg.det(r) - np.sign(np.linalg.det(m))) < 1e-12


def test_jvps_match_finite_differences():
    rng = np.random.default_rng(13)
    f = np.array([np.diag([1.3, 0.9, 0.7]) + 0.2 * rng.standard_normal((3, 3)) for _ in range(10)])
    df = rng.standard_normal((10, 3, 3))
    d = rng.standard_normal((10, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    mu, lame, eps = 3.0e4, 8.0e4, 1e-6
    got = O.jvps(f, df, d, mu, lame, 0.8)
    plus = O.projections(f + eps * df, d, mu, lame, 2 * mu, 0.8)
    minus = O.projections(f - eps * df, d, mu, lame, 2 * mu, 0.8)
    for name in ("rotation", "volume", "neo_hookean", "fibre"):
        numeric = (plus[name] - minus[name]) / (2 * eps)
        np.testing.assert_allclose(got[name], numeric, rtol=0, atol=2e-6 * max(1.0, np.abs(numeric).max()))


def test_rigid_free_fall_is_the_inertial_target():
    # Dyadic cells: float32 rest shapes are exact, F = I, every projection returns I.
    vertices, tets, poses = meshes.kuhn_tet_grid(3, 2, 2, 0.125, origin=(0, 0, 1))
    free = np.ones(len(vertices), bool)
    mats = np.tile([2e4, 5e4, 0.0], (len(tets), 1))
    body = O.OracleBody("tet4", tets, poses=poses, materials=mats, density=1000.0,
                        inverse_mass=free.astype(np.float32), gravity=(0, 0, -9.81), dt=2.0**-7)
    v = np.tile([0.25, -0.125, 0.5], (len(vertices), 1)).astype(np.float32)
    out = body.forward(2.0**-7, vertices.astype(np.float32), v)
    expected = vertices + 2.0**-7 * v + 2.0**-14 * np.array([0, 0, -9.81])
    np.testing.assert_allclose(out["x"], expected, rtol=0, atol=1e-12)


def test_mirrored_rest_shape_is_stationary():
    # Unsigned SVD: a reflection is another rest shape of every projection.
    vertices, tets, poses = meshes.kuhn_tet_grid(3, 2, 2, 0.125, origin=(0, 0, 1))
    mats = np.tile([2e4, 5e4, 0.0], (len(tets), 1))
    free = np.ones(len(vertices), bool)
    reflected = vertices.copy()
    reflected[:, 0] *= -1
    for material in ("pd", "neo_hookean"):
        body = O.OracleBody("tet4", tets, poses=poses, materials=mats, density=1000.0, material=material,
                            inverse_mass=free.astype(np.float32), gravity=(0, 0, 0), dt=1e-3, max_iterations=20)
        out = body.forward(1e-3, reflected.astype(np.float32), np.zeros_like(reflected, dtype=np.float32))
        np.testing.assert_allclose(out["x"], reflected, rtol=0, atol=1e-12)


def test_element_mode_adjoint_matches_central_differences():
    """FP64 FD on the oracle itself: the per-element Jacobian is the exact derivative."""
    vertices, tets, poses, materials, free = cube_body(2, seed=7)
    dt = 1e-3
    rng = np.random.default_rng(1)
    q0 = (vertices + 2e-4 * rng.standard_normal(vertices.shape) * free[:, None]).astype(np.float32)
    w = rng.standard_normal(vertices.shape)

    def make(mats):
        return O.OracleBody("tet4", tets, poses=poses, materials=mats, density=1000.0,
                            inverse_mass=free.astype(np.float32), gravity=(0, 0, -9.81), dt=dt, rel_tol=1e-14,
                            abs_tol=1e-18, max_iterations=3000, jacobian_mode="element")

    body = make(materials)
    fw = body.forward(dt, q0, np.zeros_like(q0))
    bw = body.backward(dt, fw["x"], w.astype(np.float32), None)
    base = materials.astype(np.float64)
    for e, col, h in ((0, 0, 50.0), (5, 1, 100.0)):
        vals = []
        for sign in (1, -1):
            m = base.copy()
            m[e, col] += sign * h
            vals.append((make(m).forward(dt, q0, np.zeros_like(q0))["x"] * w).sum())
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(bw["materials"][e, col] - fd) <= 1e-4 * abs(fd), (e, col, bw["materials"][e, col], fd)


def test_activation_gradients_match_central_differences():
    """FP64 FD of the oracle's converged step in the fibre targets (u = 1 - target, so dL/du = -dL/dtarget)."""
    vertices, tets, poses, materials, free = cube_body(2, seed=3, log_e=(np.log(2e5), np.log(2e5)))
    dt = 1e-3
    rng = np.random.default_rng(4)
    directions = rng.standard_normal((len(tets), 3))
    stiffness = np.where(np.arange(len(tets)) % 3 == 0, 0.0, 3.0e5)
    q0 = (vertices + 3e-4 * rng.standard_normal(vertices.shape) * free[:, None]).astype(np.float32)
    w = rng.standard_normal(vertices.shape)
    body = O.OracleBody("tet4", tets, poses=poses, materials=materials, density=1000.0,
                        inverse_mass=free.astype(np.float32), gravity=(0, 0, -9.81), dt=dt, rel_tol=1e-14,
                        abs_tol=1e-18, max_iterations=3000, muscle_directions=directions, muscle_stiffness=stiffness)
    targets = np.where(stiffness > 0, 0.8, 1.0)
    fw = body.forward(dt, q0, np.zeros_like(q0), targets=targets)
    bw = body.backward(dt, fw["x"], w.astype(np.float32), None, materials=False, targets=targets, activations=True)
    assert bw["status"] == 0 and np.all(bw["activations"][stiffness == 0] == 0)
    h = 1e-4
    for e in (1, 5, 10):
        vals = []
        for sign in (1, -1):
            t = targets.copy()
            t[e] += sign * h
            vals.append((body.forward(dt, q0, np.zeros_like(q0), targets=t)["x"] * w).sum())
        fd = -(vals[0] - vals[1]) / (2 * h)
        assert abs(bw["activations"][e] - fd) <= 1e-4 * abs(fd), (e, bw["activations"][e], fd)


def test_library_exports_every_declared_entry_point():
    from diffphd_amd import _lib

    header = (ROOT / "include" / "diffphd.h").read_text()
    declared = set(re.findall(r"\b(dphd_[a-z0-9_]+)\s*\(", header))
    assert len(declared) >= 25
    lib = _lib.lib()
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert lib.dphd_version() == 6
    assert set(_lib.SIGNATURES) == declared


def test_solver_requires_a_hip_device():
    import torch

    from diffphd_amd import SolverDiffPhD
    from scenes import tet_model

    vertices, tets, poses, materials, free = cube_body(2)
    model = tet_model(torch.device("cpu"), tets, poses, materials, free, vertices, (0, 0, -9.81))
    with pytest.raises(RuntimeError, match="HIP device"):
        SolverDiffPhD(model, density=1000.0)
    with pytest.raises(ValueError):
        SolverDiffPhD(model, density=0.0)
    with pytest.raises(ValueError):
        SolverDiffPhD(model, density=1000.0, material="linear")
    with pytest.raises(ValueError):
        SolverDiffPhD(model, density=1000.0, anderson_window=9)
    with pytest.raises(ValueError):
        SolverDiffPhD(model, density=1000.0, max_iterations=0)
