"""The benchmark configurations on the GPU against the oracle, and their FP32-factor error.

* parity at reduced size (or full size where the oracle is fast): forward rollouts with equal
  iteration counts and positions, and rollout gradients (chained per-step adjoints on the oracle,
  autograd through the solver's steps on the GPU);
* at full size, the FP32-factor run against the FP64 run of the same GPU path; the error is
  printed as ``[result]``.
"""

import numpy as np
import pytest
import torch

from rollouts import gpu_rollout, hex_oracle, log_young_gradient, oracle_rollout, oracle_rollout_grad, tet_oracle
from scenes import config1, config1_force, config2, config3, config5, golden, tet_model, tip_rotation

pytestmark = pytest.mark.gpu


def _hex_solver(cuda, scene, precision=64, **kw):
    from diffphd_amd import Mesh, SolverDiffPhD
    from diffphd_amd.model import Model

    n = len(scene["vertices"])
    model = Model(
        particle_q=scene["q0"], particle_qd=np.zeros_like(scene["q0"]),
        particle_inv_mass=scene["free"].astype(np.float32), particle_world=np.zeros(n, np.int32),
        gravity=np.zeros((1, 3)), tet_indices=None, tet_poses=None, tet_materials=None, device=cuda,
    )
    solver = SolverDiffPhD(
        model, density=1000.0, hex_mesh=Mesh("beam", scene["vertices"], scene["hexes"]),
        hex_materials=scene["materials"], max_iterations=kw.pop("max_iterations", 10000),
        abs_tol=scene["abs_tol"], rel_tol=scene["rel_tol"], factor_precision=precision, **kw,
    )
    return model, solver


def _tet_solver(cuda, scene, precision=64, requires_grad=False, **kw):
    from diffphd_amd import SolverDiffPhD

    model = tet_model(cuda, scene["tets"], scene["poses"], scene["materials"], scene["free"], scene["vertices"],
                      (0.0, 0.0, -9.81), requires_grad=requires_grad)
    solver = SolverDiffPhD(
        model, density=scene.get("density", 1000.0), max_iterations=kw.pop("max_iterations", 10000),
        abs_tol=scene.get("abs_tol", 1e-9), rel_tol=scene["rel_tol"], factor_precision=precision, **kw,
    )
    return model, solver


def test_config1_full_size_matches_oracle(cuda):
    print("[case] config 1 Hex8 cantilever 20x5x5, 25 steps", flush=True)
    scene = config1()
    body = hex_oracle(scene, max_iterations=10000)
    refs = oracle_rollout(body, scene, scene["steps"], scene["q0"], np.zeros_like(scene["q0"]),
                          force=lambda t: config1_force(scene, t))
    scale = np.abs(scene["q0"]).max()
    finals = {}
    for precision in (64, 32):
        model, solver = _hex_solver(cuda, scene, precision)
        dev = cuda
        state = model.state()
        for t, ref in enumerate(refs):
            state.particle_f = torch.as_tensor(config1_force(scene, t), device=dev)
            out = model.state()
            solver.step(state, out, None, None, scene["dt"])
            assert solver.read_status() == 0
            it = int(solver.iterations[0])
            q = out.particle_q.cpu().numpy()
            if precision == 64:
                assert it == ref["iterations"], (t, it, ref["iterations"])
                np.testing.assert_allclose(q, ref["q"], rtol=0, atol=1e-7 * scale, err_msg=f"step {t}")
            else:
                assert abs(it - ref["iterations"]) <= 1, (t, it, ref["iterations"])
            state = out
        finals[precision] = state.particle_q.cpu().numpy().astype(np.float64)
    err = np.abs(finals[32] - finals[64]).max()
    print(f"[result] config 1: fp32-factor vs fp64 final position error {err:.3e} m", flush=True)
    assert err <= 1e-5 * scale


def test_config2_reduced_rollout_gradients_match_oracle(cuda):
    print("[case] config 2 Tet4 cube k=4, 5 steps, gradients", flush=True)
    scene = {**config2(k=4), "rel_tol": 1e-11, "abs_tol": 1e-14}
    steps = 5
    v0 = (np.random.default_rng(3).normal(0.0, 0.05, scene["vertices"].shape) * scene["free"][:, None]).astype(
        np.float32)
    body = tet_oracle(scene, max_iterations=5000, jacobian_mode="element")
    refs = oracle_rollout(body, scene, steps, scene["vertices"], v0)
    q_last = refs[-1]["q"].astype(np.float64)
    gq = np.zeros_like(q_last)
    gq[scene["top"]] = 2.0 * (q_last[scene["top"]] - scene["targets"])
    grad = oracle_rollout_grad(body, scene, refs, gq)
    model, solver = _tet_solver(cuda, scene, requires_grad=True, max_iterations=5000, jacobian_mode="element")
    first, last, iterations = gpu_rollout(solver, model, scene["dt"], steps, scene["vertices"], v0, track=True)
    assert [abs(a - r["iterations"]) <= 1 for a, r in zip(iterations, refs)] == [True] * steps
    np.testing.assert_allclose(last.particle_q.detach().cpu().numpy(), refs[-1]["q"], rtol=0, atol=1e-9)
    q = last.particle_q.to(torch.float64)
    top = torch.as_tensor(scene["top"], device=cuda)
    loss = ((q[top] - torch.as_tensor(scene["targets"], device=cuda)) ** 2).sum()
    loss.backward()
    assert solver.read_status() == 0
    got_m = model.tet_materials.grad[:, :2].cpu().numpy().astype(np.float64)
    for name, got, ref in (
        ("log E", log_young_gradient(got_m, scene["materials"]), log_young_gradient(grad["materials"], scene["materials"])),
        ("v0", first.particle_qd.grad.cpu().numpy(), grad["qd"]),
        ("q0", first.particle_q.grad.cpu().numpy(), grad["q"]),
    ):
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * np.abs(ref).max(), err_msg=name)


def test_config2_full_size_fp32_factor_error(cuda):
    print("[case] config 2 Tet4 cube 20^3 x 6, 50 steps fwd+adjoint, fp64 and fp32 factor", flush=True)
    scene = config2()
    results = {}
    for precision, rel in ((64, None), (32, None), ("tight", 1e-10)):
        s = scene if rel is None else {**scene, "rel_tol": rel, "abs_tol": 1e-13}
        model, solver = _tet_solver(cuda, s, 64 if precision == "tight" else precision, requires_grad=True,
                                    jacobian_mode="element", max_iterations=20000)
        first, last, iterations = gpu_rollout(solver, model, scene["dt"], scene["steps"], scene["vertices"],
                                              np.zeros_like(scene["vertices"]), track=True)
        q = last.particle_q.to(torch.float64)
        top = torch.as_tensor(scene["top"], device=cuda)
        loss = ((q[top] - torch.as_tensor(scene["targets"], device=cuda)) ** 2).sum()
        loss.backward()
        assert solver.read_status() == 0, precision
        g = log_young_gradient(model.tet_materials.grad[:, :2].cpu().numpy().astype(np.float64), scene["materials"])
        results[precision] = (last.particle_q.detach().cpu().numpy().astype(np.float64), g, iterations)
    err = np.abs(results[32][0] - results[64][0]).max()
    scale = np.abs(results["tight"][1]).max()
    gerr = np.abs(results[32][1] - results[64][1]).max() / scale
    gtol = np.abs(results[64][1] - results["tight"][1]).max() / scale
    dit = max(abs(a - b) for a, b in zip(results[32][2], results[64][2]))
    print(f"[result] config 2: fp32-factor vs fp64 final position error {err:.3e} m, iteration difference {dit}; "
          f"log E gradient: fp32 vs fp64 {gerr:.3e}, fp64 at rel 1e-6 vs rel 1e-10 {gtol:.3e} (relative to "
          f"the largest entry)", flush=True)
    assert err <= 1e-5 * np.abs(scene["vertices"]).max() and dit <= 1


def test_config3_reduced_rollout_gradient_matches_oracle(cuda):
    print("[case] config 3 Hex8 twist 16x4x4 (stiffness x 1e-2), 10 steps, tip rotation gradient", flush=True)
    # Stiffness scaled by 1e-2 so the 1e-10 gates are reached in tens of iterations.
    scene = {**config3(16, 4, 4), "rel_tol": 1e-10, "abs_tol": 1e-13}
    scene["materials"] = scene["materials"] * 1e-2
    steps = 10
    body = hex_oracle(scene, max_iterations=20000, jacobian_mode="element")
    refs = oracle_rollout(body, scene, steps, scene["q0"], np.zeros_like(scene["q0"]))
    rest = scene["vertices"]
    x = refs[-1]["q"].astype(np.float64)
    gq = np.zeros_like(x)
    for i in scene["tip"]:
        for d in (1, 2):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += 1e-7
            xm[i, d] -= 1e-7
            gq[i, d] = (tip_rotation(xp, rest, scene["tip"]) - tip_rotation(xm, rest, scene["tip"])) / 2e-7
    grad = oracle_rollout_grad(body, scene, refs, gq)
    model, solver = _hex_solver(cuda, scene, max_iterations=20000, jacobian_mode="element")
    solver.hex_materials.requires_grad_(True)
    first, last, iterations = gpu_rollout(solver, model, scene["dt"], steps, scene["q0"],
                                          np.zeros_like(scene["q0"]), track=True)
    assert [abs(a - r["iterations"]) <= 1 for a, r in zip(iterations, refs)] == [True] * steps
    np.testing.assert_allclose(last.particle_q.detach().cpu().numpy(), refs[-1]["q"], rtol=0, atol=1e-8)
    (last.particle_q.to(torch.float64) * torch.as_tensor(gq, device=cuda)).sum().backward()
    assert solver.read_status() & 8 == 0
    got = log_young_gradient(solver.hex_materials.grad.cpu().numpy(), scene["materials"])
    ref = log_young_gradient(grad["materials"], scene["materials"])
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * np.abs(ref).max())


def _tip_rotation_seed(scene, q):
    """dL/dq of the tip rotation at positions q (central differences per tip coordinate)."""
    rest = scene["vertices"]
    x = np.asarray(q, np.float64)
    gq = np.zeros_like(x)
    for i in scene["tip"]:
        for d in (1, 2):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += 1e-7
            xm[i, d] -= 1e-7
            gq[i, d] = (tip_rotation(xp, rest, scene["tip"]) - tip_rotation(xm, rest, scene["tip"])) / 2e-7
    return gq


def test_config3_reduced_at_config_stiffness_matches_oracle(cuda):
    """The twist release at the config's own stiffness (E 1e6, middle third 1e7) on 16x4x4 cells, 10 steps at
    rel 1e-10: ~3000 iterations per forward step and up to ~3200 per adjoint. Forward and adjoint iteration
    counts against the oracle's, positions, and the tip-rotation gradient w.r.t. log E (the reference records
    no Hex8 adjoint; the oracle's Hex8 adjoint is pinned by central differences)."""
    print("[case] config 3 Hex8 twist 16x4x4 at the config stiffness, 10 steps, rel 1e-10", flush=True)
    scene = {**config3(16, 4, 4), "rel_tol": 1e-10, "abs_tol": 0.0}
    steps = 10
    body = hex_oracle(scene, max_iterations=20000, jacobian_mode="element")
    refs = oracle_rollout(body, scene, steps, scene["q0"], np.zeros_like(scene["q0"]))
    gq = _tip_rotation_seed(scene, refs[-1]["q"])
    grad = oracle_rollout_grad(body, scene, refs, gq)
    assert grad["status"] == 0 and all(r["status"] == 0 for r in refs)
    model, solver = _hex_solver(cuda, scene, max_iterations=20000, jacobian_mode="element")
    solver.hex_materials.requires_grad_(True)
    solver.iteration_log = []
    first, last, iterations = gpu_rollout(solver, model, scene["dt"], steps, scene["q0"],
                                          np.zeros_like(scene["q0"]), track=True)
    (last.particle_q.to(torch.float64) * torch.as_tensor(gq, device=cuda)).sum().backward()
    adjoint = [int(it[0]) for kind, it in solver.iteration_log if kind == "adjoint"]
    assert solver.read_status() == 0
    fwd_ref = [r["iterations"] for r in refs]
    dfwd = max(abs(a - b) for a, b in zip(iterations, fwd_ref))
    dadj = max(abs(a - b) for a, b in zip(adjoint, grad["iterations"]))
    got = log_young_gradient(solver.hex_materials.grad.cpu().numpy(), scene["materials"])
    ref = log_young_gradient(grad["materials"], scene["materials"])
    gerr = np.abs(got - ref).max() / np.abs(ref).max()
    perr = np.abs(last.particle_q.detach().cpu().numpy() - refs[-1]["q"]).max()
    print(f"[result] config 3 reduced at config stiffness: forward iterations {iterations} (oracle {fwd_ref}), "
          f"adjoint {adjoint} (oracle {grad['iterations']}); largest differences {dfwd} / {dadj}; positions "
          f"{perr:.2e} m; log E gradient {gerr:.2e} of its largest entry", flush=True)
    assert dfwd <= 1 and dadj <= 1
    assert perr <= 1e-8
    assert gerr <= 1e-5


def test_config3_full_size_fwd_adjoint_fp32_factor_error(cuda):
    """64x16x16 cells, 100 steps forward and the tip-rotation adjoint w.r.t. per-cell log E, with the FP64
    and the FP32 factor at the config's gate (rel 1e-6) and the FP64 factor at rel 1e-10 (abs 0) as the
    yardstick of what the gate itself leaves in the gradient. Per-element Jacobian weights: with the
    reference's per-family mean weights on this two-stiffness beam the adjoint operator is not the
    derivative and its cascade guard fires (status 16) on the twist loss."""
    print("[case] config 3 Hex8 twist 64x16x16, 100 steps fwd+adjoint, fp64 / fp32 / tight fp64", flush=True)
    scene = config3()
    results = {}
    for precision, rel in ((64, None), (32, None), (64, 1e-10)):
        s = scene if rel is None else {**scene, "rel_tol": rel, "abs_tol": 0.0}
        model, solver = _hex_solver(cuda, s, precision, max_iterations=20000, jacobian_mode="element")
        solver.hex_materials.requires_grad_(True)
        solver.iteration_log = []
        _, last, iterations = gpu_rollout(solver, model, s["dt"], s["steps"], s["q0"], np.zeros_like(s["q0"]),
                                          track=True)
        q = last.particle_q.detach().cpu().numpy().astype(np.float64)
        gq = _tip_rotation_seed(s, q)
        (last.particle_q.to(torch.float64) * torch.as_tensor(gq, device=cuda)).sum().backward()
        bits = solver.read_status()
        adjoint = [int(it[0]) for kind, it in solver.iteration_log if kind == "adjoint"]
        g = log_young_gradient(solver.hex_materials.grad.cpu().numpy(), s["materials"])
        results[(precision, rel)] = (q, g, iterations, adjoint, bits)
        print(f"[result] config 3 full size fp{precision} rel {s['rel_tol']:g}: status {bits}, forward iterations "
              f"max {max(iterations)} mean {np.mean(iterations):.1f}, adjoint max {max(adjoint)} mean "
              f"{np.mean(adjoint):.1f}", flush=True)
        assert bits == 0, (precision, rel, bits)
    q64, g64, it64, _, _ = results[(64, None)]
    q32, g32, it32, _, _ = results[(32, None)]
    gtight = results[(64, 1e-10)][1]
    scale = np.abs(gtight).max()
    err = np.abs(q32 - q64).max()
    dit = max(abs(a - b) for a, b in zip(it32, it64))
    gerr = np.abs(g32 - g64).max() / scale
    gate = np.abs(g64 - gtight).max() / scale
    rot = np.degrees(tip_rotation(q64, scene["vertices"], scene["tip"]))
    print(f"[result] config 3: fp32-factor vs fp64 final positions {err:.3e} m, forward iteration difference "
          f"{dit}, final tip rotation {rot:.3f} deg; log E gradient fp32 vs fp64 {gerr:.3e}, fp64 at rel 1e-6 "
          f"vs rel 1e-10 {gate:.3e} (relative to the largest entry)", flush=True)
    assert err <= 1e-5 * np.abs(scene["vertices"]).max() and dit <= 1


def _config5_rollout(cuda, scene, steps, precision=64, track=False):
    model, solver = _tet_solver(cuda, scene, precision, requires_grad=track, jacobian_mode="element",
                                frictional_boundary=("spherical", scene["sphere"](0)),
                                frictional_vertices=scene["surface"], friction=scene["friction"])
    solver.prepare(scene["dt"])

    def move(t):
        solver.boundary_parameters = scene["sphere"](t)

    first, last, iterations = gpu_rollout(solver, model, scene["dt"], steps, scene["vertices"],
                                          np.zeros_like(scene["vertices"]), track=track, before=move)
    return model, solver, first, last, iterations


def test_config5_reduced_contact_rollout_matches_oracle(cuda):
    # The config's tolerance and friction, the first 8 steps of a 20-step press: later the contact set
    # changes and the single Newton step per iteration chatters (as in the reference) at any tolerance.
    print("[case] config 5 block 8x8x4 + moving sphere, 8 steps, gradients", flush=True)
    scene = {**config5(cells=(8, 8, 4), steps=20), "rel_tol": 1e-6, "abs_tol": 1e-13}
    steps = 8
    body = tet_oracle(scene, max_iterations=5000, jacobian_mode="element",
                      frictional_boundary=("spherical", scene["sphere"](0)), frictional_vertices=scene["surface"],
                      friction=scene["friction"])

    def move(t):
        body.contact["params"] = np.asarray(scene["sphere"](t), dtype=np.float64)

    refs = oracle_rollout(body, scene, steps, scene["vertices"], np.zeros_like(scene["vertices"]), before=move)
    assert all(r["status"] == 0 for r in refs)
    assert np.count_nonzero(body.contact["multiplier"]) > 0
    x = refs[-1]["q"].astype(np.float64)
    surface, targets = scene["surface"], scene["targets"]
    gq = np.zeros_like(x)
    gq[surface] = 2.0 * (x[surface] - targets)
    grad = oracle_rollout_grad(body, scene, refs, gq)
    model, solver, first, last, iterations = _config5_rollout(cuda, scene, steps, track=True)
    assert np.count_nonzero(solver.contact_multipliers.cpu().numpy()) > 0
    assert [abs(a - r["iterations"]) <= 1 for a, r in zip(iterations, refs)] == [True] * steps
    np.testing.assert_allclose(last.particle_q.detach().cpu().numpy(), refs[-1]["q"], rtol=0, atol=1e-8)
    q = last.particle_q.to(torch.float64)
    idx = torch.as_tensor(surface, device=cuda)
    ((q[idx] - torch.as_tensor(targets, device=cuda)) ** 2).sum().backward()
    assert solver.read_status() & 8 == 0
    got = model.tet_materials.grad[:, :2].cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(got, grad["materials"], rtol=0, atol=1e-4 * np.abs(grad["materials"]).max())


def test_config5_full_size_rollout(cuda):
    # The first 40 of the 200 steps (the whole press takes minutes: from step ~60 on the stick/slip
    # projection chatters to the 10000-iteration cap on many steps, as the reference's does), with the
    # FP64 and the FP32 factor at the config's tolerance (rel 1e-6), and with the FP64 factor at rel 1e-8.
    # Each step's fixed point is the same in both precisions; the rollouts differ where the stopping test
    # ends a slowly contracting stick/slip iteration at a different iterate, so the FP32-vs-FP64
    # difference is bounded by the distance of the rel-1e-6 rollout from the tighter one.
    print("[case] config 5 block 16x16x8 + moving sphere, first 40 of 200 steps forward, fp64 and fp32", flush=True)
    scene = config5()
    finals = {}
    for precision, rel in ((64, None), (32, None), (64, 1e-8)):
        s = scene if rel is None else {**scene, "rel_tol": rel, "abs_tol": 1e-13}
        model, solver, first, last, iterations = _config5_rollout(cuda, s, 40, precision)
        bits = solver.read_status()
        q = last.particle_q.detach().cpu().numpy().astype(np.float64)
        assert np.isfinite(q).all() and bits & 2 == 0
        mult = solver.contact_multipliers.cpu().numpy()
        finals[(precision, rel)] = (q, iterations)
        print(f"[result] config 5: fp{precision} rel {s['rel_tol']:.0e} iterations max {max(iterations)}, mean "
              f"{np.mean(iterations):.1f}, steps at the iteration cap {sum(i >= 10000 for i in iterations)}, status "
              f"{bits}, active contacts at the end {np.count_nonzero(mult)}", flush=True)
    # The oracle's own 200-step rollout of this scene (tests/golden/make_config5_oracle.py): the FP64 path
    # takes the oracle's iteration count on every one of the first 30 steps. (Later, stick/slip switching
    # amplifies rounding differences and the two trajectories part; the oracle too reaches the 10000
    # iteration cap on 56 of its 200 steps, mean 3669 iterations per step.)
    ref = golden("config5_oracle")
    its64 = finals[(64, None)][1]
    assert list(its64[:30]) == [int(i) for i in ref["iterations"][:30]], (its64[:30], ref["iterations"][:30])
    err = float(np.abs(finals[(32, None)][0] - finals[(64, None)][0]).max())
    spread = float(np.abs(finals[(64, 1e-8)][0] - finals[(64, None)][0]).max())
    dit = max(abs(a - b) for a, b in zip(finals[(32, None)][1], finals[(64, None)][1]))
    print(f"[result] config 5: after 40 steps fp32-factor vs fp64 position error {err:.3e} m (largest iteration "
          f"difference {dit}); fp64 at rel 1e-6 vs rel 1e-8 {spread:.3e} m", flush=True)
    assert err <= max(2.0 * spread, 1e-5 * np.abs(scene["vertices"]).max())