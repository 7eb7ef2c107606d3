"""Rollouts of the benchmark configurations on the oracle and on the GPU solver, with their gradients.

The oracle backward of a rollout chains the per-step adjoints in reverse
(dL/dq_out of step t is dL/dq_in of step t+1); parameter gradients add up
over the steps, as the reference's Tape accumulates them.
"""

import numpy as np
import torch

from oracle import oracle as O


def hex_oracle(scene, **kw):
    return O.OracleBody(
        "hex8", scene["hexes"], rest=scene["vertices"], materials=scene["materials"], density=1000.0,
        inverse_mass=scene["free"].astype(np.float32), dt=scene["dt"], abs_tol=scene["abs_tol"],
        rel_tol=scene["rel_tol"], **kw,
    )


def tet_oracle(scene, **kw):
    return O.OracleBody(
        "tet4", scene["tets"], poses=scene["poses"], materials=scene["materials"],
        density=scene.get("density", 1000.0), inverse_mass=scene["free"].astype(np.float32),
        gravity=scene.get("gravity", (0.0, 0.0, -9.81)), dt=scene["dt"], abs_tol=scene.get("abs_tol", 1e-9),
        rel_tol=scene["rel_tol"], **kw,
    )


def oracle_rollout(body, scene, steps, q0, qd0, force=None, dirichlet=None, before=None):
    """Forward rollout; returns the per-step records (x, q_out, iterations, status). ``before(t)`` runs
    ahead of step t (e.g. to move a frictional boundary)."""
    q, qd = q0, qd0
    records = []
    for t in range(steps):
        if before is not None:
            before(t)
        f = None if force is None else force(t)
        d = None if dirichlet is None else dirichlet(t)
        out = body.forward(scene["dt"], q, qd, f, dirichlet=d)
        records.append(out)
        q, qd = out["q"], out["qd"]
    return records


def oracle_rollout_grad(body, scene, records, gq_final, gqd_final=None, explicit_dirichlet=False):
    """dL/d(materials), dL/dq0, dL/dqd0 of a loss on the last step's outputs, by chained adjoints."""
    gq = np.asarray(gq_final, np.float32)
    gqd = np.zeros_like(gq) if gqd_final is None else np.asarray(gqd_final, np.float32)
    gmat = 0.0
    status = 0
    iterations = []
    for out in reversed(records):
        bw = body.backward(scene["dt"], out["x"], gq, gqd, explicit_dirichlet=explicit_dirichlet)
        gmat = gmat + bw["materials"]
        gq, gqd = bw["q"], bw["qd"]
        status |= bw["status"]
        iterations.append(bw["iterations"])
    # Adjoint iterations per step, in backward (last step first) order.
    return {"materials": gmat, "q": gq, "qd": gqd, "status": status, "iterations": iterations}


def gpu_rollout(solver, model, dt, steps, q0, qd0, force=None, dirichlet=None, track=False, before=None):
    """Step the GPU solver; with ``track`` the states stay on the autograd tape. Returns the states and
    per-step forward iteration counts."""
    dev = model.device
    state = model.state()
    state.particle_q = torch.as_tensor(np.asarray(q0, np.float32), device=dev).clone()
    state.particle_qd = torch.as_tensor(np.asarray(qd0, np.float32), device=dev).clone()
    if track:
        state.particle_q.requires_grad_(True)
        state.particle_qd.requires_grad_(True)
    first = state
    iterations = []
    for t in range(steps):
        if before is not None:
            before(t)
        if force is not None:
            state.particle_f = torch.as_tensor(force(t), device=dev)
        out = model.state()
        d = None if dirichlet is None else torch.as_tensor(dirichlet(t), device=dev)
        with torch.set_grad_enabled(track):
            solver.step(state, out, None, None, dt, dirichlet_positions=d)
        iterations.append(int(solver.iterations[0]))
        state = out
    return first, state, iterations


def log_young_gradient(g_mu_lambda, materials):
    """dL/dlog E of per-element (mu, lambda) gradients at fixed Poisson ratio (mu, lambda scale with E)."""
    m = np.asarray(materials, np.float64)[:, :2]
    return g_mu_lambda[:, 0] * m[:, 0] + g_mu_lambda[:, 1] * m[:, 1]
