"""One recorded DiffPhD step as a ``torch.autograd.Function``.

Forward runs the native step and keeps the converged iterate ``x`` (FP64) as
the checkpoint, like ``StepRecord`` (DP/adjoint.py:809-841). Backward runs
the native adjoint (``Adjoint.backward``, DP/adjoint.py:1088-1259): it
recomputes the projection Jacobian at the checkpoint, solves
``(A - K) mu = dL/dq+ + dL/dv+ / h`` with the solver's factor, and returns
the state, external-force, material, Dirichlet-target and fibre-activation
gradients (the last with respect to the contraction ``u``, DP/adjoint.py:595-620).
"""

from __future__ import annotations

import torch


class StepFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, dt, q, qd, f, materials, dirichlet, targets, activations):
        n = q.shape[0]
        q_out = torch.empty_like(q)
        qd_out = torch.empty_like(qd)
        checkpoint = torch.empty((n, 3), dtype=torch.float64, device=q.device)
        solver._native_step(
            dt,
            q.detach().contiguous(),
            qd.detach().contiguous(),
            f.detach().contiguous(),
            None if dirichlet is None else dirichlet.detach().contiguous(),
            targets,
            q_out,
            qd_out,
            checkpoint,
        )
        # The reference's solver._x holds the last step's converged iterate, taped or not.
        solver._x.copy_(checkpoint)
        ctx.solver = solver
        ctx.version = solver._version
        ctx.has_dirichlet = dirichlet is not None
        ctx.materials_shape = tuple(materials.shape)
        ctx.materials_dtype = materials.dtype
        ctx.save_for_backward(checkpoint, targets if targets is not None else torch.empty(0, device=q.device))
        ctx.has_targets = targets is not None
        ctx.activations_shape = None if activations is None else tuple(activations.shape)
        ctx.activations_dtype = None if activations is None else activations.dtype
        # The step's contact state for the Woodbury rows of its backward (StepRecord.contact).
        ctx.contact = solver.contact_state() if solver._friction is not None else None
        return q_out, qd_out

    @staticmethod
    def backward(ctx, gq_out, gqd_out):
        solver = ctx.solver
        if solver._version != ctx.version:
            raise RuntimeError("the DiffPhD factor was rebuilt after this step was recorded; recompute the rollout")
        checkpoint, targets = ctx.saved_tensors
        n = checkpoint.shape[0]
        dev = checkpoint.device
        gq = torch.zeros((n, 3), dtype=torch.float32, device=dev)
        gqd = torch.zeros_like(gq)
        gf = torch.zeros_like(gq)
        want_materials = ctx.needs_input_grad[5]
        gmat = torch.zeros((ctx.materials_shape[0], 2), dtype=torch.float64, device=dev) if want_materials else None
        gdir = torch.zeros_like(gq) if ctx.has_dirichlet else None
        want_activations = ctx.activations_shape is not None and ctx.needs_input_grad[8]
        gact = torch.zeros(ctx.activations_shape, dtype=torch.float64, device=dev) if want_activations else None
        solver._native_backward(
            checkpoint,
            targets if ctx.has_targets else None,
            None if gq_out is None else gq_out.contiguous().to(torch.float32),
            None if gqd_out is None else gqd_out.contiguous().to(torch.float32),
            gq,
            gqd,
            gf,
            gmat,
            gdir,
            gact,
            ctx.contact,
        )
        gmaterials = None
        if want_materials:
            gmaterials = torch.zeros(ctx.materials_shape, dtype=ctx.materials_dtype, device=dev)
            gmaterials[:, :2] = gmat.to(ctx.materials_dtype)
        gactivations = None if gact is None else gact.to(ctx.activations_dtype)
        return None, None, gq, gqd, gf, gmaterials, gdir, None, gactivations
