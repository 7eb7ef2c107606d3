"""Stream ordering and repeatable builds of the native solver.

* ``apply_inverse`` on a side stream (a torch stream of default priority)
  right after a step, with no synchronization in between: its products share
  the solver's buffers, so it must run after the step's queued work, and the
  next step after it. Its result, and the next step's, are bitwise those of
  the same calls made with a synchronization between them.
* Two builds with the product shapes fixed (``slice_rows > 0``) give bitwise
  the same forward and backward results (the tuned shapes set the products'
  summation order, the fixed ones do not depend on timing).
"""

import numpy as np
import pytest
import torch

from scenes import cube_body, tet_model

pytestmark = pytest.mark.gpu

DT = 1e-3


def _cube(cuda, k=12, precision=32, log_e=(np.log(1e5), np.log(1e7)), **kw):
    from diffphd_amd import SolverDiffPhD

    vertices, tets, poses, materials, free = cube_body(k, log_e=log_e)
    model = tet_model(cuda, tets, poses, materials, free, vertices, (0.0, 0.0, -9.81), requires_grad=True)
    solver = SolverDiffPhD(model, density=1000.0, rel_tol=1e-8, abs_tol=1e-12, factor_precision=precision, **kw)
    return model, solver


def _two_steps(model, solver, rhs, side=None):
    """Step, apply_inverse(rhs), step again: with synchronizations between the calls, or (`side`) with the
    inverse queued on the side stream and no ordering between the streams other than the solver's own."""
    s0, s1, s2 = model.state(), model.state(), model.state()
    torch.cuda.synchronize()
    solver.step(s0, s1, None, None, DT)
    if side is None:
        torch.cuda.synchronize()
        inv = solver.apply_inverse(rhs)
        torch.cuda.synchronize()
    else:
        with torch.cuda.stream(side):
            inv = solver.apply_inverse(rhs)
    solver.step(s1, s2, None, None, DT)
    torch.cuda.synchronize()
    return inv.cpu().numpy(), s2.particle_q.detach().cpu().numpy()


@pytest.mark.parametrize("precision", [64, 32])
def test_apply_inverse_on_a_side_stream_is_ordered_with_the_steps(cuda, precision):
    print(f"[case] apply_inverse on a side stream between two steps, fp{precision}", flush=True)
    model, solver = _cube(cuda, precision=precision, slice_rows=128)
    n = model.particle_count
    rhs = torch.randn((n, 3), dtype=torch.float64, device=cuda, generator=torch.Generator(cuda).manual_seed(1))
    inv_ref, q_ref = _two_steps(model, solver, rhs)
    model2, solver2 = _cube(cuda, precision=precision, slice_rows=128)
    side = torch.cuda.Stream(device=cuda)
    inv, q = _two_steps(model2, solver2, rhs, side)
    assert solver.read_status() == 0 and solver2.read_status() == 0
    assert inv.tobytes() == inv_ref.tobytes()
    assert q.tobytes() == q_ref.tobytes()


def _rollout_grad(model, solver, top, target, steps=3):
    solver.iteration_log = []
    state = model.state(requires_grad=True)
    q0 = state.particle_q
    for _ in range(steps):
        out = model.state()
        solver.step(state, out, None, None, DT)
        state = out
    ((state.particle_q[top].to(torch.float64) - target) ** 2).sum().backward()
    log = [(kind, it.tolist()) for kind, it in solver.iteration_log]
    solver.iteration_log = None
    return log, state.particle_q.detach().cpu().numpy(), q0.grad.cpu().numpy(), model.tet_materials.grad.cpu().numpy()


def test_builds_with_fixed_product_shapes_repeat_bitwise(cuda):
    print("[case] two builds with slice_rows = 128, fp32 factor, k = 14 cube of E 1e6, 3 steps + backward", flush=True)
    runs = []
    for _ in range(2):
        model, solver = _cube(cuda, k=14, precision=32, log_e=(np.log(1e6), np.log(1e6)), slice_rows=128)
        q = model.particle_q.cpu().numpy()
        top = torch.as_tensor(np.flatnonzero(q[:, 2] == q[:, 2].max()), device=cuda)
        target = model.particle_q[top].to(torch.float64) + torch.tensor([0.0, 0.0, 1e-3], device=cuda, dtype=torch.float64)
        runs.append(_rollout_grad(model, solver, top, target))
        assert solver.info.slice_rows == 128 and solver.info.tune_seconds == 0.0
        assert solver.read_status() == 0
    (log_a, q_a, gq_a, gm_a), (log_b, q_b, gq_b, gm_b) = runs
    assert log_a == log_b
    assert q_a.tobytes() == q_b.tobytes()
    assert gq_a.tobytes() == gq_b.tobytes()
    assert gm_a.tobytes() == gm_b.tobytes()
