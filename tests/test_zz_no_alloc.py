"""dphd_backward_static (no device allocation) against dphd_backward, through the C-ABI.

Two static calls in a row, the second on a smaller solver than the first
(the shared scratch keeps the first call's contents), each compared with the
allocating entry point's results on the same checkpoint; and a solver whose
adjoint does not fit the static buffer is refused with an error.
"""

import numpy as np
import pytest
import torch

from scenes import config4_bodies, cube_body, packed_bodies_model, tet_model

pytestmark = pytest.mark.gpu

DT = 1e-3


def _lib():
    from diffphd_amd import _lib

    return _lib


def _stepped(model, **kw):
    """A solver stepped once from a perturbed rest state; returns it, its FP64 checkpoint and seeds."""
    from diffphd_amd import SolverDiffPhD

    solver = SolverDiffPhD(model, density=1000.0, rel_tol=1e-8, jacobian_mode="element", **kw)
    rng = np.random.default_rng(5)
    free = model.particle_inv_mass.cpu().numpy() > 0
    q = model.particle_q.cpu().numpy() + 1e-4 * rng.standard_normal((model.particle_count, 3)) * free[:, None]
    s_in = model.state()
    s_in.particle_q = torch.as_tensor(q.astype(np.float32), device=model.device)
    s_out = model.state()
    with torch.no_grad():
        solver.step(s_in, s_out, None, None, DT)
    checkpoint = solver.solution().clone()
    gq = torch.as_tensor(rng.standard_normal(q.shape).astype(np.float32), device=model.device)
    gqd = torch.as_tensor((0.01 * rng.standard_normal(q.shape)).astype(np.float32), device=model.device)
    return solver, checkpoint, gq, gqd


def _backward(solver, entry, checkpoint, gq_out, gqd_out):
    L = _lib()
    n, e = solver._n, solver._element_count
    dev = checkpoint.device
    outs = [torch.zeros((n, 3), dtype=torch.float32, device=dev) for _ in range(3)]
    gmat = torch.zeros((e, 2), dtype=torch.float64, device=dev)
    code = getattr(L.lib(), entry)(
        solver._h, L.ptr(checkpoint), None, L.ptr(gq_out), L.ptr(gqd_out), *[L.ptr(o) for o in outs], L.ptr(gmat),
        None, None, None,
    )
    L.check(code, entry)
    torch.cuda.synchronize()
    result = {k: v.cpu().numpy() for k, v in zip(("q", "qd", "f", "materials"), (*outs, gmat))}
    result["iterations"] = solver.adjoint_iterations.numpy()
    result["status"] = solver.read_status()
    return result


def test_static_backward_matches_allocating_backward_on_consecutive_calls(cuda):
    print("[case] dphd_backward_static: 3 batched bodies, then one smaller body", flush=True)
    bodies = config4_bodies([5, 6, 4], seed=4)
    big_model, _, _ = packed_bodies_model(bodies, cuda, requires_grad=False)
    big = _stepped(big_model, independent_bodies=True)
    vertices, tets, poses, materials, free = cube_body(3, seed=2)
    small_model = tet_model(cuda, tets, poses, materials, free, vertices, (0, 0, -9.81))
    small = _stepped(small_model)
    reference = [_backward(s, "dphd_backward", c, g, gd) for s, c, g, gd in (big, small)]
    # Two static calls in a row, the second on the smaller solver.
    static = [_backward(s, "dphd_backward_static", c, g, gd) for s, c, g, gd in (big, small)]
    for which, ref, got in zip(("batched", "small"), reference, static):
        assert ref["status"] == 0 and got["status"] == 0, (which, ref["status"], got["status"])
        np.testing.assert_array_equal(got["iterations"], ref["iterations"], err_msg=which)
        for key in ("q", "qd", "f", "materials"):
            assert np.abs(ref[key]).max() > 0.0, (which, key)
            np.testing.assert_array_equal(got[key], ref[key], err_msg=f"{which} {key}")
    # And once more on the larger solver after the smaller one.
    again = _backward(big[0], "dphd_backward_static", *big[1:])
    for key in ("q", "qd", "f", "materials"):
        np.testing.assert_array_equal(again[key], reference[0][key], err_msg=f"repeat {key}")


def test_static_backward_refuses_a_solver_beyond_its_buffer(cuda):
    print("[case] dphd_backward_static: 20^3 cube beyond the static buffer", flush=True)
    L = _lib()
    vertices, tets, poses, materials, free = cube_body(20, seed=1)
    model = tet_model(cuda, tets, poses, materials, free, vertices, (0, 0, -9.81))
    solver, checkpoint, gq, gqd = _stepped(model)
    n = solver._n
    outs = [torch.zeros((n, 3), dtype=torch.float32, device=cuda) for _ in range(3)]
    code = L.lib().dphd_backward_static(solver._h, L.ptr(checkpoint), None, L.ptr(gq), L.ptr(gqd),
                                        *[L.ptr(o) for o in outs], None, None, None, None)
    assert code == 1
    message = L.lib().dphd_last_error().decode()
    assert "bytes of scratch" in message and "static buffer" in message, message
    torch.cuda.synchronize()
    for o in outs:
        assert float(o.abs().max()) == 0.0
