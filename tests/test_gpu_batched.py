"""The batched-bodies path (``independent_bodies=True``): forward and adjoint of config-4-shaped bodies.

* 8 uneven Tet4 cubes (k = 4..16, so some bodies span many 256-vertex
  chunks and they converge at different iterations), per-body E and Poisson
  ratio drawn as in the batched configuration, 3 recorded steps and the
  backward sweep of the top-face target loss: every body's forward and
  adjoint iteration counts and its log E, Poisson-ratio, q0 and v0 gradients
  against that body's oracle rollout (chained per-step oracle adjoints).
  The oracle bodies take the exact inverse factor from the host factor of
  the library, which is bitwise the restated reference factor
  (tests/test_setup_bitexact.py); the pure-Python factor is too slow at
  k = 16.
* all 256 bodies of the benchmark workload batched, against single-body
  solvers of the same bodies.
* Dirichlet gradients (explicit targets and integrated pinned velocities)
  against the oracle.
"""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from rollouts import oracle_rollout, oracle_rollout_grad
from scenes import config4_bodies, lame_parameter_gradients, packed_bodies_model, tet_model

pytestmark = pytest.mark.gpu

DT = 1e-3
SIZES = [4, 16, 6, 12, 5, 9, 14, 7]


def _solver(model, **kw):
    from diffphd_amd import SolverDiffPhD

    return SolverDiffPhD(model, density=1000.0, **kw)


def _gpu_rollout_grad(solver, model, voff, bodies, steps):
    """Recorded steps then the backward sweep; per-body gradients and per-step iteration counts."""
    dev = model.device
    idx = torch.as_tensor(np.concatenate([b["top"] + voff[i] for i, b in enumerate(bodies)]), device=dev)
    tgt = torch.as_tensor(np.concatenate([b["target"] for b in bodies]), dtype=torch.float32, device=dev)
    solver.iteration_log = []
    state = model.state(requires_grad=True)
    q0, qd0 = state.particle_q, state.particle_qd
    for _ in range(steps):
        out = model.state()
        solver.step(state, out, None, None, DT)
        state = out
    loss = ((state.particle_q[idx] - tgt) ** 2).sum()
    loss.backward()
    log, solver.iteration_log = solver.iteration_log, None
    fwd = np.array([it for kind, it in log if kind == "forward"])
    adj = np.array([it for kind, it in log if kind == "adjoint"])
    gmat = model.tet_materials.grad.detach().cpu().numpy()[:, :2].astype(np.float64)
    return {
        "materials": gmat,
        "q": q0.grad.detach().cpu().numpy(),
        "qd": qd0.grad.detach().cpu().numpy(),
        "forward": fwd,
        "adjoint": adj,
        "status": solver.read_status(),
    }


def _oracle_body(b, rel_tol, abs_tol):
    from diffphd_amd import _lib

    return O.OracleBody(
        "tet4", b["t"], poses=b["p"], materials=b["m"], density=1000.0, inverse_mass=b["free"].astype(np.float32),
        gravity=(0.0, 0.0, -9.81), dt=DT, rel_tol=rel_tol, abs_tol=abs_tol,
        factor=lambda csr: _lib.Factor(*csr).as_dict(),
    )


def _body_gradients(b, gmat, gq, gqd):
    log_e, nu = lame_parameter_gradients(gmat, b["young"], b["poisson"])
    return {"log_e": log_e, "nu": nu, "q": gq, "qd": gqd}


@pytest.mark.parametrize(
    "precision,rel_tol,abs_tol,tol,slack",
    [(64, 1e-10, 1e-14, 1e-5, 1), (32, 1e-10, 1e-14, 1e-5, 2), (64, 1e-6, 1e-9, 1e-5, 1)],
    ids=["fp64-tight", "fp32-tight", "fp64-rel1e-6"],
)
def test_batched_rollout_gradients_match_per_body_oracle(cuda, precision, rel_tol, abs_tol, tol, slack):
    """Per body: iterations of every forward and adjoint solve within +-``slack`` of the oracle's, and the
    log E, nu, q0, v0 gradients within ``tol`` of the oracle's (relative to the largest entry).

    At the batched configuration's own gate (rel 1e-6, abs 1e-9) the FP64 path follows the oracle's
    iteration path (same arithmetic, same order of the gate's sums), so it agrees far below the gate."""
    steps = 3
    bodies = config4_bodies(SIZES)
    print(f"[case] batched 8 bodies k={SIZES} fp{precision} rel {rel_tol:g}", flush=True)
    model, voff, eoff = packed_bodies_model(bodies, cuda)
    solver = _solver(model, rel_tol=rel_tol, abs_tol=abs_tol, independent_bodies=True, factor_precision=precision)
    got = _gpu_rollout_grad(solver, model, voff, bodies, steps)
    assert got["status"] == 0
    for i, b in enumerate(bodies):
        body = _oracle_body(b, rel_tol, abs_tol)
        scene = {"dt": DT}
        records = oracle_rollout(body, scene, steps, b["v"], np.zeros_like(b["v"]))
        q = records[-1]["q"]
        gq = np.zeros_like(q)
        gq[b["top"]] = 2.0 * (q[b["top"]] - b["target"].astype(np.float32))
        ref = oracle_rollout_grad(body, scene, records, gq)
        assert ref["status"] == 0
        its_f = [r["iterations"] for r in records]
        assert np.all(np.abs(got["forward"][:, i] - its_f) <= slack), (i, b["k"], got["forward"][:, i], its_f)
        assert np.all(np.abs(got["adjoint"][:, i] - ref["iterations"]) <= slack), (
            i, b["k"], got["adjoint"][:, i], ref["iterations"])
        v = slice(voff[i], voff[i + 1])
        e = slice(eoff[i], eoff[i + 1])
        mine = _body_gradients(b, got["materials"][e], got["q"][v], got["qd"][v])
        theirs = _body_gradients(b, ref["materials"], ref["q"], ref["qd"])
        errs = {}
        for key in ("log_e", "nu", "q", "qd"):
            scale = np.abs(theirs[key]).max()
            assert scale > 0.0, key
            errs[key] = np.abs(np.asarray(mine[key]) - theirs[key]).max() / scale
        print(f"[result] body {i} k={b['k']} E={b['young']:.3g} nu={b['poisson']:.3f} fwd {its_f} "
              f"adj {ref['iterations']} rel err " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
        for key, err in errs.items():
            assert err <= tol, (i, b["k"], key, err)


@pytest.mark.parametrize("precision,tol,slack", [(64, 1e-6, 0), (32, 1e-5, 1)], ids=["fp64", "fp32"])
def test_full_batch_gradients_equal_single_body_solvers(cuda, precision, tol, slack):
    """All 256 bodies of the benchmark workload in one batched solver (rel 1e-6) against single-body solvers
    of a sample of them. The batched solver's S^T runs over whole column groups while most bodies are
    active and over row slices of them after, a single body's always over slices, so the products add in
    different orders. With FP64 products the iteration counts are equal and the gradients differ only by
    that rounding as the Anderson solves carry it (~3e-8, as against the oracle); FP32 products (the FP32
    factor's defect correction) move the iterate the gate stops at by their own rounding."""
    sizes = np.random.default_rng(3).integers(4, 17, 256)
    bodies = config4_bodies(sizes, seed=1)
    steps = 2
    print(f"[case] 256 batched bodies vs single-body solvers, fp{precision}", flush=True)
    model, voff, eoff = packed_bodies_model(bodies, cuda)
    solver = _solver(model, rel_tol=1e-6, independent_bodies=True, factor_precision=precision)
    got = _gpu_rollout_grad(solver, model, voff, bodies, steps)
    assert got["status"] == 0
    order = np.argsort(sizes, kind="stable")
    sample = sorted(set(order[np.linspace(0, 255, 12).astype(int)]) | {int(np.argmax(sizes)), 0, 255})
    worst = 0.0
    for i in sample:
        b = bodies[i]
        m1, v1, _ = packed_bodies_model([b], cuda)
        s1 = _solver(m1, rel_tol=1e-6, factor_precision=precision)
        one = _gpu_rollout_grad(s1, m1, v1, [b], steps)
        assert one["status"] == 0
        assert np.all(np.abs(got["forward"][:, i] - one["forward"][:, 0]) <= slack), (i, got["forward"][:, i], one["forward"])
        assert np.all(np.abs(got["adjoint"][:, i] - one["adjoint"][:, 0]) <= slack), (i, got["adjoint"][:, i], one["adjoint"])
        v = slice(voff[i], voff[i + 1])
        e = slice(eoff[i], eoff[i + 1])
        for key, mine, theirs in (("materials", got["materials"][e], one["materials"]), ("q", got["q"][v], one["q"]),
                                  ("qd", got["qd"][v], one["qd"])):
            scale = np.abs(theirs).max()
            err = np.abs(mine - theirs).max() / scale
            worst = max(worst, err)
            assert err <= tol, (i, b["k"], key, err)
    print(f"[result] 256 bodies fp{precision}: {len(sample)} sampled bodies against their single-body solves, "
          f"largest gradient difference {worst:.2e}, adjoint iterations max {got['adjoint'].max()}", flush=True)


def _grid(cell=0.1):
    """3x2x2 cells at z = 1, mu 2e4, lambda 5e4, the x = 0 face pinned (tests/test_diff_phd_solver.py:65-92)."""
    from diffphd_amd import meshes

    vertices, tets, poses = meshes.kuhn_tet_grid(3, 2, 2, cell, origin=(0.0, 0.0, 1.0))
    materials = np.tile([2.0e4, 5.0e4, 0.0], (len(tets), 1)).astype(np.float32)
    free = vertices[:, 0] > vertices[:, 0].min()
    rng = np.random.default_rng(3)
    q = vertices.astype(np.float64)
    center = q.mean(axis=0)
    deformed = np.where(free[:, None], center + (q - center) * np.array([1.2, 0.85, 1.05]), q)
    deformed[free] += 0.01 * rng.standard_normal((free.sum(), 3))
    velocity = np.where(free[:, None], 0.3 * rng.standard_normal(q.shape), 0.0)
    return vertices.astype(np.float32), tets, poses, materials, free, deformed.astype(np.float32), velocity.astype(np.float32)


@pytest.mark.parametrize("explicit", [True, False], ids=["explicit-targets", "integrated-start"])
def test_dirichlet_gradients_match_oracle(cuda, explicit):
    """dL/d(dirichlet_positions) with explicit targets, or the pinned rows of dL/dq and dL/dqd with
    integrated pinned velocities (DP/adjoint.py:493-553), against the oracle (pinned by central
    differences and the closed form dL/dq[pinned] = -w_v / h)."""
    print(f"[case] Dirichlet gradients explicit={explicit}", flush=True)
    dt = 0.004
    vertices, tets, poses, materials, free, q, qd = _grid()
    tight = dict(max_iterations=3000, abs_tol=1e-12, rel_tol=1e-12)
    model = tet_model(cuda, tets, poses, materials, free, vertices, (0, 0, -9.81))
    solver = _solver(model, integrate_pinned_velocities=True, **tight)
    rng = np.random.default_rng(81)
    w0 = rng.standard_normal(q.shape).astype(np.float32)
    w1 = (0.01 * rng.standard_normal(q.shape)).astype(np.float32)
    s_in = model.state()
    s_in.particle_q = torch.as_tensor(q, device=cuda).clone().requires_grad_(True)
    s_in.particle_qd = torch.as_tensor(qd, device=cuda).clone().requires_grad_(True)
    s_out = model.state()
    targets = None
    dirichlet = None
    if explicit:
        targets = (q + np.array([0.002, -0.001, 0.003], dtype=np.float32)).astype(np.float32)
        dirichlet = torch.as_tensor(targets, device=cuda).clone().requires_grad_(True)
    solver.step(s_in, s_out, None, None, dt, dirichlet_positions=dirichlet)
    torch.autograd.backward([s_out.particle_q, s_out.particle_qd],
                            [torch.as_tensor(w0, device=cuda), torch.as_tensor(w1, device=cuda)])
    assert solver.read_status() == 0
    body = O.OracleBody("tet4", tets, poses=poses, materials=materials, density=1000.0,
                        inverse_mass=free.astype(np.float32), gravity=(0, 0, -9.81), dt=dt, integrate_pinned=True,
                        **tight)
    fw = body.forward(dt, q, qd, dirichlet=targets)
    bw = body.backward(dt, fw["x"], w0, w1, materials=False, explicit_dirichlet=explicit)
    assert bw["status"] == 0
    pinned = ~free
    pairs = [("q", s_in.particle_q.grad), ("qd", s_in.particle_qd.grad)]
    if explicit:
        pairs.append(("dirichlet", dirichlet.grad))
        np.testing.assert_array_equal(dirichlet.grad.cpu().numpy()[free], 0.0)
        np.testing.assert_allclose(s_in.particle_q.grad.cpu().numpy()[pinned], -w1[pinned] / dt, rtol=2e-7, atol=1e-7)
    for key, grad in pairs:
        ref = bw[key]
        assert np.abs(ref[pinned]).max() > 0.0 or key == "qd", key
        np.testing.assert_allclose(grad.cpu().numpy(), ref, rtol=0, atol=1e-6 * np.abs(ref).max(), err_msg=key)
