"""Parity fixtures in the reference's schema newton.diff_phd_parity.v1 (docs/validation.md:451-472):
the named scenes pass the acceptance rules of tools/parity/compare.py:63-101 and survive an NPZ round
trip, the rules reject what compare.py rejects, and the oracle's fixture outputs have the meaning
the schema gives them (gradients of the linear objective, checked by central differences)."""

import numpy as np
import pytest

from diffphd_amd import parity
import parity_oracle


def _small(steps=2, seed=3):
    rest, tets, mu, lame, dirichlet = parity.grid(fix_left=True, dims=(2, 1, 1))
    rng = np.random.default_rng(seed)
    free = ~dirichlet
    q = rest + np.where(free[:, None], 0.01 * rng.standard_normal(rest.shape), 0.0)
    v = np.where(free[:, None], 0.2 * rng.standard_normal(rest.shape), 0.0)
    f = rng.uniform(-1.0, 1.0, (steps, len(rest), 3))
    return parity.make_fixture("small", rest, tets, mu, lame, dirichlet, steps=steps, q=q, v=v, f_ext=f, seed=seed)


def test_named_fixtures_pass_the_schema_checks_and_round_trip(tmp_path):
    names = set()
    for fx in parity.cases():
        assert parity.check(fx) == [], (fx["case"], parity.check(fx))
        names.add(fx["case"])
        path = tmp_path / f"{fx['case']}.fixture.npz"
        parity.save(path, fx)
        back = parity.load(path)
        assert set(back) == set(fx)
        for key, value in fx.items():
            if isinstance(value, np.ndarray):
                np.testing.assert_array_equal(back[key], value, err_msg=key)
            else:
                assert back[key] == value, key
    # The harness's case list (tools/parity/port_cases.py:240-282).
    assert {"homogeneous_free", "zero_volume_weight", "neo_hookean_large_deformation", "plane_impact",
            "sticking_friction", "sliding_friction", "multistep_pinned", "muscle_activation", "mesh_contact_poke",
            "mesh_boundary_poke", "mesh_contact_prescribed", "pinned_contrast_10x_per_element_small_dt"} <= names


def test_schema_checks_reject_what_the_harness_rejects():
    fx = _small()
    assert parity.check(fx) == []
    cases = {
        "schema": {"schema": "other"},
        "material": {"material": "stvk"},
        "contact kind": {"contact_kind": "capsule"},
        "q is not float32-representable": {"q": fx["q"] + 1e-12},
        "activation convention": {"activations": np.zeros((2, len(fx["tets"]))), "activation_convention": "a = u"},
        "velocity_damping": {"velocity_damping": 1.5},
        "f_ext must be (steps, N, 3)": {"f_ext": fx["f_ext"][:1]},
        "dirichlet_positions shape": {"dirichlet_positions": np.zeros((1, 2, 3))},
        "dirichlet_positions is not float32-representable": {
            "dirichlet_positions": np.broadcast_to(fx["q"], (int(fx["steps"]),) + fx["q"].shape) + 1e-12},
        "mesh_contact parameters": {"mesh_contact_parameters": np.array([1e-3, -1.0, 0.0, 0.0, 0.0]),
                                    "mesh_contact_faces_a": fx["tets"][:, :3], "mesh_contact_faces_b": fx["tets"][:, 1:]},
        "mesh_contact faces": {"mesh_contact_parameters": np.zeros(5), "mesh_contact_faces_a": fx["tets"][:, :3] + 10**6,
                               "mesh_contact_faces_b": fx["tets"][:, 1:]},
        "tet indices out of range": {"tets": fx["tets"] + len(fx["rest_vertices"])},
        "only Tet4 topology": {"tets": fx["tets"][:, :3]},
    }
    for reason, change in cases.items():
        reasons = parity.check({**fx, **change})
        assert any(reason in r for r in reasons), (reason, reasons)
    flipped = fx["tets"].copy()
    flipped[0, [1, 2]] = flipped[0, [2, 1]]
    assert "tetrahedra must be positively oriented" in parity.check({**fx, "tets": flipped})
    with pytest.raises(ValueError, match="fixture cannot be run"):
        parity.run({**fx, "material": "stvk"})


def test_oracle_fixture_gradients_are_the_objective_gradients():
    """Central differences of the oracle rollout's loss against its chained adjoint, for the initial
    velocity and each step's external force (the loss is on the float32 states, so the steps are
    large enough to dwarf their rounding)."""
    fx = _small()
    out = parity_oracle.run(fx)
    assert np.all(out["converged"]) and out["status_bits"][-1] == 0
    free = out["free"]
    rng = np.random.default_rng(0)
    checks = [("v", None, 0.05), ("f_ext", 0, 20.0), ("f_ext", 1, 20.0)]
    for key, step, h in checks:
        grad = out["dL_dv"] if key == "v" else out["dL_df_ext"][step]
        for _ in range(3):
            i = int(rng.choice(np.flatnonzero(free)))
            c = int(rng.integers(3))
            plus, minus = {k: np.copy(v) for k, v in fx.items() if isinstance(v, np.ndarray)}, None
            minus = {k: np.copy(v) for k, v in plus.items()}
            if key == "v":
                plus["v"][i, c] += h
                minus["v"][i, c] -= h
            else:
                plus["f_ext"][step, i, c] += h
                minus["f_ext"][step, i, c] -= h
            lp = parity_oracle.run({**fx, **plus})["loss"]
            lm = parity_oracle.run({**fx, **minus})["loss"]
            fd = (lp - lm) / (2 * h)
            assert abs(fd - grad[i, c]) <= 2e-3 * np.abs(grad[free]).max(), (key, step, i, c, fd, grad[i, c])
