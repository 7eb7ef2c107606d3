#!/usr/bin/env python
"""Benchmark of the DiffPhD projective-dynamics path on one MI355X.

Workload: 256 uneven Tet4 cube bodies (k cells per side, k ~ U{4..16}),
6 tets per cube cell, 5 mm cells, bottom face pinned, gravity; per-body
E ~ logU[1e5, 1e7] Pa, nu ~ U[0.3, 0.45], density 1000; dt 1e-3, Anderson
m = 5, per-body relative tolerance 1e-6; loss = sum over bodies of the
squared distance of the top-face vertices to targets drawn N(rest, 1 mm).
One "step" is one forward PD step of every body plus its adjoint step.

Timed: K forward steps (autograd-recorded; K = 40 by default, the
configuration's rollout) followed by the backward sweep through them, all bodies in one batched solver (`value`, steps/s), and the
same bodies as 256 independent single-body solvers run one after another
(the plain per-body GPU path). Roofline: the factor products S and S^T
of all bodies (A^{-1} = S^T S, the dominant kernels), timed live with HIP
events; algorithmic bytes count the reference's CSR form. Gradient check:
one body of every distinct size against single-body FP64 solves of the same
rollout at a tight gate, at the metric's tolerance and at a gradient-
accurate one (relative 1e-10, absolute 0), whose batched and plain rates
are reported too. CPU baseline: the FP64 C oracle on the host cores, the
same rollout of one body per distinct size, weighted by the bodies of that
size.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "newton-solver-diff-phd_amd"))

HBM_PEAK_GBPS = 8000.0
METRIC = "fwd+adjoint PD steps/s, batched vs plain GPU path at rel-res 1e-6; HBM GB/s"


def log(*args):
    print(*args, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# Scene


def make_bodies(count, kmin, kmax, cell=0.005):
    from diffphd_amd import meshes

    sizes = np.random.default_rng(3).integers(kmin, kmax + 1, count)
    mat_rng = np.random.default_rng(1)
    young = np.exp(mat_rng.uniform(np.log(1e5), np.log(1e7), count))
    poisson = mat_rng.uniform(0.3, 0.45, count)
    target_rng = np.random.default_rng(2)
    bodies = []
    for b, k in enumerate(sizes):
        origin = (0.1 * (b % 16), 0.1 * (b // 16), 0.0)
        v, t, p = meshes.kuhn_tet_grid(int(k), int(k), int(k), cell, origin=origin)
        mu, lam = meshes.lame(young[b], poisson[b])
        mats = np.zeros((len(t), 3), np.float32)
        mats[:, 0], mats[:, 1] = mu, lam
        free = v[:, 2] > v[:, 2].min()
        top = np.flatnonzero(v[:, 2] == v[:, 2].max())
        target = v[top] + target_rng.normal(0.0, 1e-3, (len(top), 3))
        bodies.append(dict(k=int(k), v=v.astype(np.float32), t=t, p=p, m=mats, free=free, top=top, target=target,
                           young=float(young[b]), poisson=float(poisson[b])))
    return bodies


def packed_model(bodies, device):
    from diffphd_amd.model import Model

    voff, eoff = [0], [0]
    for b in bodies:
        voff.append(voff[-1] + len(b["v"]))
        eoff.append(eoff[-1] + len(b["t"]))
    model = Model(
        particle_q=np.concatenate([b["v"] for b in bodies]),
        particle_qd=np.zeros((voff[-1], 3), np.float32),
        particle_inv_mass=np.concatenate([b["free"] for b in bodies]).astype(np.float32),
        particle_world=np.zeros(voff[-1], np.int32),
        gravity=np.array([[0.0, 0.0, -9.81]]),
        tet_indices=np.concatenate([b["t"] + voff[i] for i, b in enumerate(bodies)]),
        tet_poses=np.concatenate([b["p"] for b in bodies]),
        tet_materials=np.concatenate([b["m"] for b in bodies]),
        device=device,
        requires_grad=True,
        body_vertex_offsets=voff,
        body_element_offsets=eoff,
    )
    return model, voff, eoff


def loss_index(bodies, voff, device):
    import torch

    idx = np.concatenate([b["top"] + voff[i] for i, b in enumerate(bodies)])
    tgt = np.concatenate([b["target"] for b in bodies])
    return torch.as_tensor(idx, device=device), torch.as_tensor(tgt, dtype=torch.float32, device=device)


def rollout(solver, model, steps, dt, idx, tgt):
    """K recorded forward steps, then the backward sweep through all of them."""
    import torch

    state = model.state(requires_grad=True)
    q0, qd0 = state.particle_q, state.particle_qd
    for _ in range(steps):
        out = model.state()
        solver.step(state, out, None, None, dt)
        state = out
    loss = ((state.particle_q[idx] - tgt) ** 2).sum()
    loss.backward()
    return loss, q0, qd0


# ---------------------------------------------------------------------------
# Algorithmic bytes: the compulsory traffic of the reference's algorithm at
# stored widths, counted once per logical pass (s = factor value bytes, V =
# 24-byte FP64 vertex vectors, CSR column indices 4 bytes).


def bytes_product(n, z, s=8, v_in=24, v_out=24):
    """One factor product (S or S^T) in the reference's CSR form: values and column indices, row offsets,
    the input and output vectors."""
    return z * (s + 4) + 8 * (n + 1) + n * (v_in + v_out)


def bytes_forward_iteration(n, e, z, s=8, window=5):
    """Per forward iteration per body: 2 Z (s+4) + 8 (n+1) + E (4K + G_e + 3s) + c_f n V, Tet4 (K = 4,
    G_e = 12 s), c_f = 6 + (3m + 5) + 4."""
    c_f = 6 + (3 * window + 5) + 4
    return 2 * z * (s + 4) + 8 * (n + 1) + e * (4 * 4 + 12 * s + 3 * s) + c_f * n * 24


def bytes_adjoint_iteration(n, e, z, s=8, window=5, s_j=8):
    """Per adjoint iteration per body: 2 Z (s+4) + 8 (n+1) + E (4K + 81 s_J) + c_a n V, c_a = 34."""
    return 2 * z * (s + 4) + 8 * (n + 1) + e * (4 * 4 + 81 * s_j) + 34 * n * 24


# ---------------------------------------------------------------------------


def factor_probe(solver, dev, n_total, z_total, s_bytes, args):
    """The S and S^T products of every body (the dominant kernels), timed with HIP events on the stream they
    are launched on (torch's current stream). One apply_inverse = pack + S + S^T; per launch = half."""
    import torch

    rhs = torch.randn((n_total, 3), dtype=torch.float64, device=dev)
    for _ in range(3):
        solver.apply_inverse(rhs)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(args.probe_reps):
        solver.apply_inverse(rhs)
    ev1.record()
    torch.cuda.synchronize()
    per_launch_s = ev0.elapsed_time(ev1) / 1e3 / (2 * args.probe_reps)
    v_half = 24 if s_bytes == 8 else 12  # the S-product vectors are FP32 in the FP32-factor mode
    # Algorithmic bytes of the pair (S: FP64 or FP32 in and out; S^T: in, FP64 out), per launch.
    launch_bytes = 0.5 * (
        bytes_product(n_total, z_total, s_bytes, v_half, v_half) + bytes_product(n_total, z_total, s_bytes, v_half, 24)
    )
    info = solver.info
    slots = float(info.factor_slots)
    vec4 = 4 * v_half // 3
    stored_bytes = float(info.factor_values) * s_bytes + 0.5 * (3 * slots * vec4 + n_total * 24)
    log(f"[bench] factor product probe {per_launch_s * 1e6:.1f} us/launch, {launch_bytes / per_launch_s / 1e9:.0f} "
        f"GB/s algorithmic, {stored_bytes / per_launch_s / 1e9:.0f} GB/s of stored bytes")
    return {"launch_s": per_launch_s, "launch_bytes": launch_bytes, "stored_bytes": stored_bytes}


def copy_ceiling(dev):
    """A plain 256 MiB device copy: the achievable streaming rate, and the reference kernel that calibrates
    the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    src = torch.empty(2**25, dtype=torch.float64, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(10):
        dst.copy_(src)
    ev1.record()
    torch.cuda.synchronize()
    gbps = 2 * src.numel() * 8 * 10 / (ev0.elapsed_time(ev1) / 1e3) / 1e9
    log(f"[bench] copy ceiling {gbps:.0f} GB/s")
    return gbps


def stratified_sample(bodies):
    """One body of every distinct size k (the first of each), and how many bodies each stands for."""
    first, count = {}, {}
    for i, b in enumerate(bodies):
        first.setdefault(b["k"], i)
        count[b["k"]] = count.get(b["k"], 0) + 1
    ks = sorted(first)
    return [first[k] for k in ks], [count[k] for k in ks]


def body_gradients(b, gmat, gq, gqd):
    """(dL/dlog E, dL/dnu) of a body with uniform E and Poisson ratio, and its q0 / v0 gradients."""
    g = np.asarray(gmat, np.float64)[:, :2].sum(axis=0)
    young, nu = b["young"], b["poisson"]
    mu = young / (2.0 * (1.0 + nu))
    lam = young * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    dmu = -mu / (1.0 + nu)
    dlam = young * (1.0 + 2.0 * nu**2) / ((1.0 + nu) ** 2 * (1.0 - 2.0 * nu) ** 2)
    return {"log_e": float(g[0] * mu + g[1] * lam), "nu": float(g[0] * dmu + g[1] * dlam),
            "q0": np.asarray(gq, np.float64), "v0": np.asarray(gqd, np.float64)}


def per_body_gradients(model, q0, qd0, voff, eoff, bodies, pick):
    gmat = model.tet_materials.grad.detach().cpu().numpy()
    gq = q0.grad.detach().cpu().numpy()
    gqd = qd0.grad.detach().cpu().numpy()
    return {i: body_gradients(bodies[i], gmat[eoff[i]:eoff[i + 1]], gq[voff[i]:voff[i + 1]], gqd[voff[i]:voff[i + 1]])
            for i in pick}


def gradient_errors(got, ref):
    """Relative error of each gradient against the reference over the sampled bodies: ``vector`` the L2 error
    of all sampled bodies' values together (log E and nu one entry per body), ``worst_body`` the largest
    per-body relative error (|a - b| / |b| for the scalars, which are tiny for nearly static stiff bodies)."""
    out = {}
    for key in ("log_e", "nu", "q0", "v0"):
        a = np.concatenate([np.ravel(got[i][key]) for i in ref])
        b = np.concatenate([np.ravel(ref[i][key]) for i in ref])
        worst = max(float(np.linalg.norm(np.ravel(got[i][key]) - np.ravel(ref[i][key])) /
                          max(np.linalg.norm(np.ravel(ref[i][key])), 1e-300)) for i in ref)
        out[key] = {"vector": float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)), "worst_body": worst}
    out["max_vector"] = max(v["vector"] for v in out.values())
    return out


def single_body_rollouts(bodies, pick, dev, steps, dt, **solver_kw):
    """Each picked body in a single-body solver (the plain per-body GPU path): its gradients and rollout time."""
    import torch

    from diffphd_amd import SolverDiffPhD

    grads, seconds, status = {}, {}, 0
    for i in pick:
        b = bodies[i]
        m1, v1, e1 = packed_model([b], dev)
        s1 = SolverDiffPhD(m1, density=1000.0, anderson_window=5, **solver_kw)
        s1.prepare(dt)
        i1, t1 = loss_index([b], v1, dev)
        torch.cuda.synchronize()
        start = time.perf_counter()
        _loss, q0, qd0 = rollout(s1, m1, steps, dt, i1, t1)
        torch.cuda.synchronize()
        seconds[i] = time.perf_counter() - start
        status |= s1.read_status()
        grads[i] = per_body_gradients(m1, q0, qd0, v1, e1, [b], [0])[0]
        del s1
    return grads, seconds, status


def timed_batched(model, dev, dt, idx, tgt, args, rel_tol, abs_tol):
    """Batched solver of every body: warm-up steps, then the timed K-step rollout; a further untimed rollout
    logs the iteration counts and loop counters (logging synchronizes, so it stays out of the timed one)."""
    import torch

    from diffphd_amd import SolverDiffPhD

    t0 = time.time()
    solver = SolverDiffPhD(
        model, density=1000.0, rel_tol=rel_tol, abs_tol=abs_tol, anderson_window=5, independent_bodies=True,
        factor_precision=args.precision, iteration_chunk=args.chunk,
    )
    solver.prepare(dt)
    setup = time.time() - t0
    for _ in range(args.warmup):
        rollout(solver, model, 1, dt, idx, tgt)
    torch.cuda.synchronize()
    solver.read_status()
    model.tet_materials.grad = None
    torch.cuda.synchronize()
    start = time.perf_counter()
    loss, q0, qd0 = rollout(solver, model, args.steps, dt, idx, tgt)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    status = solver.read_status()
    grads_model = model.tet_materials.grad.detach().clone()
    # The forward alone (no tape): the same K steps from rest.
    torch.cuda.synchronize()
    start = time.perf_counter()
    with torch.no_grad():
        state = model.state()
        for _ in range(args.steps):
            out = model.state()
            solver.step(state, out, None, None, dt)
            state = out
    torch.cuda.synchronize()
    forward_elapsed = time.perf_counter() - start
    before = solver.loop_stats()
    solver.iteration_log = []
    model.tet_materials.grad = None
    rollout(solver, model, args.steps, dt, idx, tgt)
    log_ = solver.iteration_log
    solver.iteration_log = None
    after = solver.loop_stats()
    model.tet_materials.grad = grads_model
    fwd = np.array([it for kind, it in log_ if kind == "forward"])
    adj = np.array([it for kind, it in log_ if kind == "adjoint"])
    loops = {m: {k: (after[m][k] - before[m][k]) / args.steps for k in after[m]} for m in after}
    return dict(solver=solver, setup=setup, elapsed=elapsed, status=status, loss=float(loss.detach()), q0=q0,
                qd0=qd0, fwd=fwd, adj=adj, loops=loops, forward_elapsed=forward_elapsed)


# ---------------------------------------------------------------------------
# The other benchmark configurations, one JSON line each (--workload):
# steps/s, iterations, microseconds per iteration, loop chunks per step and
# algorithmic GB/s. Setup (factor) is excluded from the timed region.


def bytes_iteration(n, e, z, s, corners, samples, adjoint, window=5):
    """The per-iteration byte model for one body: 2 Z (s+4) + 8 (n+1) + E (4K + G_e + 3s) + c_f n V forward
    (G_e = 12 s for Tet4, 0 for Hex8's shared table), 2 Z (s+4) + 8 (n+1) + E (4K + 81 Q s_J) + c_a n V adjoint."""
    if adjoint:
        return 2 * z * (s + 4) + 8 * (n + 1) + e * (4 * corners + 81 * samples * 8) + 34 * n * 24
    g_e = 12 * s if corners == 4 else 0
    return 2 * z * (s + 4) + 8 * (n + 1) + e * (4 * corners + g_e + 3 * s) + (6 + 3 * window + 5 + 4) * n * 24


def config_scene(name):
    """Synthetic inputs of the benchmark configurations (seeds: 1 materials, 2 targets)."""
    from diffphd_amd import meshes

    if name in ("config1", "config3"):
        nx, ny, nz, cell = (20, 5, 5, 0.01) if name == "config1" else (64, 16, 16, 0.005)
        v, hexes = meshes.hex_grid(nx, ny, nz, cell)
        xcell = np.arange(len(hexes)) // (ny * nz)
        stiff = (xcell > nx // 3) & (xcell <= (nx // 3) * 2)
        mu, lam = meshes.lame(np.where(stiff, 1.0e7, 1.0e6), 0.45)
        free = v[:, 0] != 0.0
        tip = np.flatnonzero(v[:, 0] == v[:, 0].max())
        q0 = v if name == "config1" else meshes.twisted_positions(v, np.pi / 6.0)
        return dict(kind="hex8", v=v, hexes=hexes, mats=np.column_stack([mu, lam]), free=free, tip=tip,
                    q0=q0.astype(np.float32), gravity=(0.0, 0.0, 0.0), density=1000.0,
                    dt=0.01 if name == "config1" else 1e-3, steps=25 if name == "config1" else 100,
                    rel_tol=1e-4 if name == "config1" else 1e-6, adjoint=name == "config3")
    if name == "config2":
        v, t, p = meshes.kuhn_tet_grid(20, 20, 20, 0.005)
        young = np.exp(np.random.default_rng(1).uniform(np.log(1e5), np.log(1e7), len(t)))
        mu, lam = meshes.lame(young, 0.4)
        top = np.flatnonzero(v[:, 2] == v[:, 2].max())
        target = v[top] + np.random.default_rng(2).normal(0.0, 1.0e-3, (len(top), 3))
        return dict(kind="tet4", v=v, t=t, p=p, mats=np.column_stack([mu, lam, np.zeros_like(mu)]),
                    free=v[:, 2] > v[:, 2].min(), top=top, target=target, q0=v.astype(np.float32),
                    gravity=(0.0, 0.0, -9.81), density=1000.0, dt=1e-3, steps=50, rel_tol=1e-6, adjoint=True)
    if name == "config5":
        v, t, p = meshes.kuhn_tet_grid(16, 16, 8, 0.0025)
        mu, lam = meshes.lame(np.full(len(t), 5.0e5), 0.45)
        top_z = v[:, 2].max()
        surface = np.flatnonzero(v[:, 2] == top_z).astype(np.int32)
        target = v[surface] + np.random.default_rng(2).normal(0.0, 1.0e-3, (len(surface), 3))
        center0 = np.array([0.02, 0.02, top_z + 0.005])

        def sphere(step, total=200):
            f = min(step, total) / total
            c = center0 + np.array([0.003 * f, 0.0, -0.006 * f])
            return (float(c[0]), float(c[1]), float(c[2]), 0.005)

        return dict(kind="tet4", v=v, t=t, p=p, mats=np.column_stack([mu, lam, np.zeros_like(mu)]),
                    free=v[:, 2] > v[:, 2].min(), top=surface, target=target, q0=v.astype(np.float32),
                    gravity=(0.0, 0.0, -9.81), density=1100.0, dt=1e-3, steps=40, rel_tol=1e-6, adjoint=True,
                    sphere=sphere, friction=0.5)
    raise ValueError(name)


def tip_rotation_torch(q, rest, tip):
    """Signed rotation of the tip cross-section about x (differentiable)."""
    import torch

    r0 = torch.as_tensor(rest[tip, 1:] - rest[tip, 1:].mean(axis=0), dtype=torch.float64, device=q.device)
    qt = q[torch.as_tensor(tip, device=q.device)][:, 1:].to(torch.float64)
    r = qt - qt.mean(dim=0)
    return torch.atan2((r0[:, 0] * r[:, 1] - r0[:, 1] * r[:, 0]).sum(), (r0 * r).sum())


def run_workload(name, args):
    import torch

    from diffphd_amd import Mesh, SolverDiffPhD
    from diffphd_amd.model import Model

    dev = torch.device("cuda:0")
    sc = config_scene(name)
    precision = 64 if name == "config1" else args.precision
    steps = sc["steps"] if args.workload_steps <= 0 else args.workload_steps
    n = len(sc["v"])
    kw = dict(density=sc["density"], rel_tol=sc["rel_tol"], abs_tol=1e-9, max_iterations=10000,
              factor_precision=precision, jacobian_mode="element")
    if sc["kind"] == "hex8":
        model = Model(particle_q=sc["q0"], particle_qd=np.zeros_like(sc["q0"]),
                      particle_inv_mass=sc["free"].astype(np.float32), particle_world=np.zeros(n, np.int32),
                      gravity=np.zeros((1, 3)), tet_indices=None, tet_poses=None, tet_materials=None, device=dev)
        solver = SolverDiffPhD(model, hex_mesh=Mesh("beam", sc["v"], sc["hexes"]), hex_materials=sc["mats"], **kw)
        elements, corners, samples = len(sc["hexes"]), 8, 8
        materials = solver.hex_materials
    else:
        model = Model(particle_q=sc["q0"], particle_qd=np.zeros_like(sc["q0"]),
                      particle_inv_mass=sc["free"].astype(np.float32), particle_world=np.zeros(n, np.int32),
                      gravity=np.asarray(sc["gravity"])[None], tet_indices=sc["t"], tet_poses=sc["p"],
                      tet_materials=sc["mats"], device=dev, requires_grad=sc["adjoint"])
        extra = {}
        if "sphere" in sc:
            extra = dict(frictional_boundary=("spherical", sc["sphere"](0)), frictional_vertices=sc["top"],
                         friction=sc["friction"])
        solver = SolverDiffPhD(model, **kw, **extra)
        elements, corners, samples = len(sc["t"]), 4, 1
        materials = model.tet_materials
    t0 = time.time()
    solver.prepare(sc["dt"])
    setup = time.time() - t0
    s_bytes = {64: 8, 32: 4}[precision]
    z = float(solver.body_factor_nnz()[0])
    probe = factor_probe(solver, dev, n, z, s_bytes, args)
    products = {
        "launch_us": probe["launch_s"] * 1e6,
        "algorithmic_gbps": probe["launch_bytes"] / probe["launch_s"] / 1e9,
        "frac": probe["launch_bytes"] / probe["launch_s"] / 1e9 / HBM_PEAK_GBPS,
        "stored_gbps": probe["stored_bytes"] / probe["launch_s"] / 1e9,
        "bytes_per_launch": probe["launch_bytes"],
        "stored_bytes_per_launch": probe["stored_bytes"],
    }
    if args.probe_only:
        print(json.dumps({"workload": name, "products": products, "copy_ceiling_gbps": copy_ceiling(dev)}))
        return
    if sc["adjoint"] and sc["kind"] == "hex8":
        materials.requires_grad_(True)

    def forces(t):
        if name != "config1":
            return None
        f = torch.zeros((n, 3), dtype=torch.float32, device=dev)
        f[torch.as_tensor(sc["tip"], device=dev), 0] = 80.0 * t / 25
        return f

    def one_rollout(count):
        state = model.state(requires_grad=sc["adjoint"])
        q0, qd0 = state.particle_q, state.particle_qd
        with torch.set_grad_enabled(sc["adjoint"]):
            for t in range(count):
                if "sphere" in sc:
                    solver.boundary_parameters = sc["sphere"](t)
                f = forces(t)
                if f is not None:
                    state.particle_f = f
                out = model.state()
                solver.step(state, out, None, None, sc["dt"])
                state = out
            if sc["adjoint"]:
                if name == "config3":
                    loss = tip_rotation_torch(state.particle_q, sc["v"], sc["tip"])
                else:
                    idx = torch.as_tensor(sc["top"], device=dev)
                    tgt = torch.as_tensor(sc["target"], dtype=torch.float64, device=dev)
                    loss = ((state.particle_q[idx].to(torch.float64) - tgt) ** 2).sum()
                loss.backward()
        return q0, qd0

    one_rollout(1)
    torch.cuda.synchronize()
    solver.read_status()
    before = solver.loop_stats()
    start = time.perf_counter()
    one_rollout(steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    after = solver.loop_stats()
    status = solver.read_status()
    forward_rate = None
    if sc["adjoint"]:
        with torch.no_grad():
            torch.cuda.synchronize()
            start = time.perf_counter()
            state = model.state()
            for t in range(steps):
                if "sphere" in sc:
                    solver.boundary_parameters = sc["sphere"](t)
                out = model.state()
                solver.step(state, out, None, None, sc["dt"])
                state = out
            torch.cuda.synchronize()
            forward_rate = steps / (time.perf_counter() - start)
    solver.iteration_log = []
    one_rollout(steps)
    log_ = solver.iteration_log
    solver.iteration_log = None
    fwd = np.array([int(it[0]) for kind, it in log_ if kind == "forward"])
    adj = np.array([int(it[0]) for kind, it in log_ if kind == "adjoint"])
    total_bytes = fwd.sum() * bytes_iteration(n, elements, z, s_bytes, corners, samples, False)
    if sc["adjoint"]:
        total_bytes += adj.sum() * bytes_iteration(n, elements, z, s_bytes, corners, samples, True)
    iterations = int(fwd.sum() + adj.sum())
    loops = {m: {k: (after[m][k] - before[m][k]) / steps for k in after[m]} for m in after}
    result = {
        "workload": name,
        "what": "forward" if not sc["adjoint"] else "forward+adjoint",
        "steps": steps,
        "config_steps": sc["steps"] if name != "config5" else 200,
        "steps_per_s": steps / elapsed,
        "ms_per_step": 1e3 * elapsed / steps,
        "forward_steps_per_s": forward_rate if sc["adjoint"] else steps / elapsed,
        "setup_s": setup,
        "factor_precision": precision,
        "vertices": n,
        "elements": elements,
        "factor_nnz": int(z),
        "iterations": {"forward_mean": float(fwd.mean()), "forward_max": int(fwd.max()),
                       "adjoint_mean": float(adj.mean()) if len(adj) else None,
                       "adjoint_max": int(adj.max()) if len(adj) else None},
        "us_per_iteration": 1e6 * elapsed / max(iterations, 1),
        "element_iterations_per_s": elements * iterations / elapsed,
        "algorithmic_gbps": total_bytes / elapsed / 1e9,
        "frac_of_hbm_peak": total_bytes / elapsed / 1e9 / HBM_PEAK_GBPS,
        "loops_per_step": loops,
        "product_tuning": {"slice_rows": solver.info.slice_rows, "slice_switch_groups": solver.info.slice_switch,
                           "split_quads": solver.info.split_quads, "split_switch_groups": solver.info.split_switch,
                           "tune_seconds": solver.info.tune_seconds},
        "products": products,
        "status": status,
        "cpu_baseline": None if args.no_cpu_baseline else workload_cpu_baseline(name, sc, args),
    }
    print(json.dumps(result))


# Steps of the bounded CPU sample per configuration (each about 5-60 s of oracle time on 16 cores).
WORKLOAD_CPU_STEPS = {"config1": 25, "config2": 10, "config3": 3, "config5": 20}


def workload_cpu_baseline(name, sc, args):
    """FP64 C oracle on the host cores: the configuration's first steps from its initial state, then the
    backward sweep of its loss (adjoint configurations), timed without setup."""
    import torch

    from diffphd_amd import _lib
    from oracle import oracle

    steps = args.workload_cpu_steps or WORKLOAD_CPU_STEPS[name]
    threads = min(16, os.cpu_count() or 1)
    n = len(sc["v"])
    kw = dict(materials=sc["mats"], density=sc["density"], inverse_mass=sc["free"].astype(np.float32),
              gravity=sc["gravity"], dt=sc["dt"], rel_tol=sc["rel_tol"], max_iterations=10000,
              jacobian_mode="element", threads=threads, factor=lambda csr: _lib.Factor(*csr).as_dict())
    if sc["kind"] == "hex8":
        body = oracle.OracleBody("hex8", sc["hexes"], rest=sc["v"], **kw)
    elif "sphere" in sc:
        body = oracle.OracleBody("tet4", sc["t"], poses=sc["p"], frictional_boundary=("spherical", sc["sphere"](0)),
                                 frictional_vertices=sc["top"], friction=sc["friction"], **kw)
    else:
        body = oracle.OracleBody("tet4", sc["t"], poses=sc["p"], **kw)
    q, qd = sc["q0"].astype(np.float32), np.zeros((n, 3), np.float32)
    records, iterations = [], 0
    t0 = time.perf_counter()
    for t in range(steps):
        if "sphere" in sc:
            body.contact["params"] = np.asarray(sc["sphere"](t), dtype=np.float64)
        f = None
        if name == "config1":
            f = np.zeros((n, 3), np.float32)
            f[sc["tip"], 0] = 80.0 * t / 25
        fw = body.forward(sc["dt"], q, qd, f)
        iterations += fw["iterations"]
        records.append(fw)
        q, qd = fw["q"], fw["qd"]
    if sc["adjoint"]:
        if name == "config3":
            qt = torch.tensor(q, dtype=torch.float64, requires_grad=True)
            tip_rotation_torch(qt, sc["v"], sc["tip"]).backward()
            gq = qt.grad.numpy().astype(np.float32)
        else:
            gq = np.zeros_like(q)
            gq[sc["top"]] = 2.0 * (q[sc["top"]] - sc["target"])
        gqd = np.zeros_like(q)
        for t in reversed(range(steps)):
            if "sphere" in sc:
                body.contact["params"] = np.asarray(sc["sphere"](t), dtype=np.float64)
            bw = body.backward(sc["dt"], records[t]["x"], gq, gqd, materials=True)
            iterations += bw["iterations"]
            gq, gqd = bw["q"], bw["qd"]
    seconds = time.perf_counter() - t0
    what = "forward" if not sc["adjoint"] else "forward + the backward sweep"
    return {
        "value": steps / seconds,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "nproc": os.cpu_count(),
        "cpu_model": cpu_model(),
        "sample": f"the configuration's first {steps} steps ({what}, {iterations} iterations) on the FP64 oracle; "
                  f"{seconds:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40, help="timed fwd+adj steps (the config's rollout: 40)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--bodies", type=int, default=256)
    ap.add_argument("--kmin", type=int, default=4)
    ap.add_argument("--kmax", type=int, default=16)
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64),
                    help="factor precision: 32 (the workload's FP32, FP64 defect correction) or 64")
    ap.add_argument("--chunk", type=int, default=0, help="iterations per captured graph (0 = default)")
    ap.add_argument("--plain-bodies", type=int, default=None, help="bodies timed on the plain path (default all)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-gradient-check", action="store_true")
    ap.add_argument("--accurate-rel", type=float, default=1e-11,
                    help="relative tolerance (absolute 0) of the gradient-accurate setting, the loosest gate at "
                         "which the sampled gradients meet 1e-4 (rel 1e-10 gives 1.07e-4); 0 skips it")
    ap.add_argument("--dump-iterations", default=None,
                    help="write the per-step, per-body iteration counts and factor sizes to this JSON file")
    ap.add_argument("--probe-reps", type=int, default=20)
    ap.add_argument("--workload", default="config4", choices=("config1", "config2", "config3", "config4", "config5"),
                    help="config4 (default) is the headline batched workload; the others print their own line")
    ap.add_argument("--workload-steps", type=int, default=0, help="steps of a --workload run (0 = the config's)")
    ap.add_argument("--workload-cpu-steps", type=int, default=0,
                    help="steps of a --workload run's CPU sample (0 = the configuration's default)")
    ap.add_argument("--probe-only", action="store_true",
                    help="only the factor-product probe and the copy calibration (for counter runs)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(
            f"bench.py measures one GPU: --gpus {args.gpus} is not supported (bodies are independent; "
            "a multi-GPU run would shard them with no per-step communication)",
            file=sys.stderr,
        )
        sys.exit(2)

    if args.workload != "config4":
        run_workload(args.workload, args)
        return

    import torch

    from diffphd_amd import SolverDiffPhD

    dev = torch.device("cuda:0")
    dt = 1e-3
    bodies = make_bodies(args.bodies, args.kmin, args.kmax)
    model, voff, eoff = packed_model(bodies, dev)
    idx, tgt = loss_index(bodies, voff, dev)
    n_total, e_total = voff[-1], eoff[-1]
    s_bytes = 8 if args.precision == 64 else 4
    log(f"[bench] {len(bodies)} bodies, {n_total} vertices, {e_total} tets")
    if args.probe_only:
        solver = SolverDiffPhD(model, density=1000.0, rel_tol=1e-6, anderson_window=5, independent_bodies=True,
                               factor_precision=args.precision, iteration_chunk=args.chunk)
        solver.prepare(dt)
        z_all = float(np.sum(solver.body_factor_nnz()))
        probe = factor_probe(solver, dev, n_total, z_all, s_bytes, args)
        copy_gbps = copy_ceiling(dev)
        print(json.dumps({"probe_launch_us": probe["launch_s"] * 1e6, "copy_ceiling_gbps": copy_gbps}))
        return

    run = timed_batched(model, dev, dt, idx, tgt, args, 1e-6, 1e-9)
    solver, elapsed = run["solver"], run["elapsed"]
    tuning = {"slice_rows": solver.info.slice_rows, "slice_switch_groups": solver.info.slice_switch,
              "split_quads": solver.info.split_quads, "split_switch_groups": solver.info.split_switch,
              "tune_seconds": solver.info.tune_seconds}
    log(f"[bench] batched setup {run['setup']:.1f}s, nnz(S) {solver.info.factor_nnz}; rollout {elapsed:.3f}s for "
        f"{args.steps} fwd+adj steps, loss {run['loss']:.6e}, status {run['status']}, loops/step {run['loops']}")
    fwd_iters, adj_iters = run["fwd"], run["adj"]
    grad_ok = bool(torch.isfinite(model.tet_materials.grad).all()) and bool(torch.isfinite(run["qd0"].grad).all())
    pick, weight = stratified_sample(bodies)

    # Algorithmic bytes of the rollout from the per-body iteration counts.
    n_b = np.diff(voff)
    e_b = np.diff(eoff)
    z_b = np.array(solver.body_factor_nnz(), dtype=np.float64)
    z_total = float(z_b.sum())
    per_fwd = bytes_forward_iteration(n_b, e_b, z_b, s=s_bytes)
    per_adj = bytes_adjoint_iteration(n_b, e_b, z_b, s=s_bytes)
    total_bytes = float((fwd_iters * per_fwd).sum() + (adj_iters * per_adj).sum())
    elem_iters = float((fwd_iters * e_b).sum() + (adj_iters * e_b).sum())
    rollout_gbps = total_bytes / elapsed / 1e9

    if args.dump_iterations:
        Path(args.dump_iterations).write_text(json.dumps({
            "forward": fwd_iters.tolist(), "adjoint": adj_iters.tolist(), "nnz": z_b.tolist(),
            "vertices": n_b.tolist(), "elements": e_b.tolist(), "factor_values": int(solver.info.factor_values)}))
    probe = factor_probe(solver, dev, n_total, z_total, s_bytes, args)
    copy_gbps = copy_ceiling(dev)

    # Gradient accuracy: the sampled bodies' gradients against single-body FP64 solves at a tight gate
    # (relative 1e-12, absolute 0) of the same rollout.
    grad_check = None
    accurate = None
    if not args.no_gradient_check:
        batched_grads = per_body_gradients(model, run["q0"], run["qd0"], voff, eoff, bodies, pick)
        ref, _sec, ref_status = single_body_rollouts(bodies, pick, dev, args.steps, dt, rel_tol=1e-12, abs_tol=0.0,
                                                     max_iterations=20000, factor_precision=64)
        err = gradient_errors(batched_grads, ref)
        grad_check = {
            "reference": "single-body FP64 solves of the same rollout at rel 1e-12, abs 0 (status "
                         f"{ref_status}); bodies {[bodies[i]['k'] for i in pick]} (k)",
            "rel_err_at_metric_tolerance": err,
            "target": 1e-4,
            "target_met_at_metric_tolerance": bool(err["max_vector"] <= 1e-4),
        }
        log(f"[bench] gradient error at rel 1e-6 / abs 1e-9 vs the FP64 tight solves: {err}")
        if args.accurate_rel > 0:
            del solver
            run_a = timed_batched(model, dev, dt, idx, tgt, args, args.accurate_rel, 0.0)
            acc_grads = per_body_gradients(model, run_a["q0"], run_a["qd0"], voff, eoff, bodies, pick)
            err_a = gradient_errors(acc_grads, ref)
            _g, sec_a, _st = single_body_rollouts(bodies, pick, dev, args.steps, dt, rel_tol=args.accurate_rel,
                                                 abs_tol=0.0, factor_precision=args.precision)
            plain_a = args.steps / sum(sec_a[i] * w for i, w in zip(pick, weight))
            value_a = args.steps / run_a["elapsed"]
            accurate = {
                "rel_tol": args.accurate_rel,
                "abs_tol": 0.0,
                "steps_per_s": value_a,
                "status": run_a["status"],
                "plain_steps_per_s_estimate": plain_a,
                "plain_estimate": "one single-body solve per distinct k, weighted by the bodies of that k",
                "speedup_vs_plain": value_a / plain_a,
                "gradient_rel_err": err_a,
                "target_met": bool(err_a["max_vector"] <= 1e-4),
                "iterations": {"forward_mean": float(run_a["fwd"].mean()), "adjoint_mean": float(run_a["adj"].mean()),
                               "adjoint_max": int(run_a["adj"].max())},
            }
            grad_check["accurate_setting"] = accurate
            log(f"[bench] accurate setting rel {args.accurate_rel:g} abs 0: {value_a:.2f} steps/s, plain estimate "
                f"{plain_a:.3f}, gradient error {err_a}")
            del run_a

    # Plain per-body path: one single-body solver per body, run one after another.
    plain = None
    if not args.no_plain:
        count = len(bodies) if args.plain_bodies is None else min(args.plain_bodies, len(bodies))
        solvers = []
        t0 = time.time()
        for b in bodies[:count]:
            m1, v1, _e1 = packed_model([b], dev)
            s1 = SolverDiffPhD(m1, density=1000.0, rel_tol=1e-6, anderson_window=5, factor_precision=args.precision,
                               iteration_chunk=args.chunk)
            s1.prepare(dt)
            i1, t1 = loss_index([b], v1, dev)
            solvers.append((s1, m1, i1, t1))
        log(f"[bench] plain setup {time.time() - t0:.1f}s for {count} bodies")
        for s1, m1, i1, t1 in solvers[: min(4, count)]:
            rollout(s1, m1, 1, dt, i1, t1)
        torch.cuda.synchronize()
        start = time.perf_counter()
        for s1, m1, i1, t1 in solvers:
            rollout(s1, m1, args.steps, dt, i1, t1)
        torch.cuda.synchronize()
        plain_elapsed = time.perf_counter() - start
        plain_elapsed *= len(bodies) / count
        plain = args.steps / plain_elapsed
        log(f"[bench] plain rollout {plain_elapsed:.3f}s (scaled to {len(bodies)} bodies)")
        del solvers

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(bodies, pick, weight, dt, args.steps)

    # HBM traffic per launch of the probe kernels from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE runs (calibrated against a plain copy kernel).
    traffic = None
    traffic_file = ROOT / "profiles" / f"traffic_factor_products_batched256_fp{args.precision}.json"
    if traffic_file.exists() and len(bodies) == 256 and (args.kmin, args.kmax) == (4, 16):
        traffic = json.loads(traffic_file.read_text())["traffic_bytes_per_launch"]

    value = args.steps / elapsed
    achieved = probe["launch_bytes"] / probe["launch_s"] / 1e9
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 64 else "f32-factor/f64-vectors",
        "data": "synthetic",
        "config": {
            "workload": f"{len(bodies)} uneven Tet4 cube bodies k~U{{{args.kmin}..{args.kmax}}}, 6 tets/cell, "
            "5 mm cells, bottom pinned, gravity, dt 1e-3, Anderson m=5, per-body rel 1e-6, fwd+adjoint",
            "bodies": len(bodies),
            "vertices": n_total,
            "tets": e_total,
            "factor_nnz": int(z_total),
            "parallelism": "single GPU, bodies batched per launch",
        },
        "forward_steps_per_s": args.steps / run["forward_elapsed"],
        "plain_steps_per_s": plain,
        "speedup_vs_plain": None if plain is None else value / plain,
        "element_iterations_per_s": elem_iters / elapsed,
        "iterations": {
            "forward_mean": float(fwd_iters.mean()),
            "forward_max": int(fwd_iters.max()),
            "adjoint_mean": float(adj_iters.mean()),
            "adjoint_max": int(adj_iters.max()),
        },
        "loops_per_step": run["loops"],
        "product_tuning": tuning,
        "rollout_hbm_gbps": rollout_gbps,
        "status": run["status"],
        "gradients_finite": grad_ok,
        "gradient_check": grad_check,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_factor_rows + k_factor_cols (S and S^T products of all bodies), per launch",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "frac_basis": "algorithmic bytes: the reference's CSR form (values + 4-byte column indices + offsets "
                          "+ vectors); the block layout reads no column indices",
            "traffic": traffic,
            "frac_traffic": None if traffic is None else traffic / probe["launch_s"] / 1e9 / HBM_PEAK_GBPS,
            "launch_us": probe["launch_s"] * 1e6,
            "bytes_per_launch": probe["launch_bytes"],
            "stored_bytes_per_launch": probe["stored_bytes"],
            "stored_gbps": probe["stored_bytes"] / probe["launch_s"] / 1e9,
            "frac_stored": probe["stored_bytes"] / probe["launch_s"] / 1e9 / HBM_PEAK_GBPS,
            "copy_ceiling_gbps": copy_gbps,
            "frac_of_copy": achieved / copy_gbps,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(result))


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(bodies, pick, weight, dt, steps):
    """FP64 C oracle on the host cores: the timed rollout (K steps from rest, then the backward sweep of the
    top-face loss) of one body of every distinct size, weighted by the bodies of that size."""
    from diffphd_amd import _lib
    from oracle import oracle

    threads = min(16, os.cpu_count() or 1)
    weighted = 0.0
    seconds = 0.0
    for i, w in zip(pick, weight):
        b = bodies[i]
        body = oracle.OracleBody(
            "tet4", b["t"], poses=b["p"], materials=b["m"], density=1000.0, inverse_mass=b["free"].astype(np.float32),
            gravity=(0.0, 0.0, -9.81), dt=dt, rel_tol=1e-6, threads=threads,
            factor=lambda csr: _lib.Factor(*csr).as_dict(),
        )
        q = b["v"].astype(np.float32)
        qd = np.zeros_like(q)
        t0 = time.perf_counter()
        records = []
        for _ in range(steps):
            fw = body.forward(dt, q, qd)
            records.append(fw)
            q, qd = fw["q"], fw["qd"]
        gq = np.zeros_like(q)
        gq[b["top"]] = 2.0 * (q[b["top"]] - b["target"])
        gqd = np.zeros_like(q)
        for fw in reversed(records):
            bw = body.backward(dt, fw["x"], gq, gqd)
            gq, gqd = bw["q"], bw["qd"]
        t = time.perf_counter() - t0
        seconds += t
        weighted += w * t
    return {
        "value": steps / weighted,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "nproc": os.cpu_count(),
        "cpu_model": cpu_model(),
        "sample": f"the timed rollout ({steps} fwd steps from rest + backward sweep) of one body per distinct k "
                  f"({len(pick)} bodies), each weighted by the bodies of its k; {seconds:.1f} s of oracle time",
    }


if __name__ == "__main__":
    main()
