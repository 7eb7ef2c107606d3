#!/usr/bin/env python
"""Benchmark of the DiffPhD projective-dynamics path on one MI355X.

Workload: 256 uneven Tet4 cube bodies (k cells per side, k ~ U{4..16}),
6 tets per cube cell, 5 mm cells, bottom face pinned, gravity; per-body
E ~ logU[1e5, 1e7] Pa, nu ~ U[0.3, 0.45], density 1000; dt 1e-3, Anderson
m = 5, per-body relative tolerance 1e-6; loss = sum over bodies of the
squared distance of the top-face vertices to targets drawn N(rest, 1 mm).
One "step" is one forward PD step of every body plus its adjoint step.

Timed: K forward steps (autograd-recorded) followed by the backward sweep
through them, all bodies in one batched solver (`value`, steps/s), and the
same bodies as 256 independent single-body solvers run one after another
(the plain per-body GPU path). Roofline: the factor products S and S^T
of all bodies (A^{-1} = S^T S, the dominant kernels), timed live with HIP
events; algorithmic bytes count the reference's CSR form. CPU baseline: the
FP64 C oracle, one step of every body on the host cores.
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
        bodies.append(dict(k=int(k), v=v.astype(np.float32), t=t, p=p, m=mats, free=free, top=top, target=target))
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


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed fwd+adj steps (the full config uses 40)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--bodies", type=int, default=256)
    ap.add_argument("--kmin", type=int, default=4)
    ap.add_argument("--kmax", type=int, default=16)
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64),
                    help="factor precision: 32 (the workload's FP32, FP64 defect correction) or 64")
    ap.add_argument("--chunk", type=int, default=0, help="iterations per captured graph (0 = default)")
    ap.add_argument("--plain-bodies", type=int, default=None, help="bodies timed on the plain path (default all)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-bodies", type=int, default=0, help="0 = every body")
    ap.add_argument("--probe-reps", type=int, default=20)
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

    import torch

    from diffphd_amd import SolverDiffPhD

    dev = torch.device("cuda:0")
    dt = 1e-3
    bodies = make_bodies(args.bodies, args.kmin, args.kmax)
    model, voff, eoff = packed_model(bodies, dev)
    idx, tgt = loss_index(bodies, voff, dev)
    n_total, e_total = voff[-1], eoff[-1]
    log(f"[bench] {len(bodies)} bodies, {n_total} vertices, {e_total} tets")

    t0 = time.time()
    solver = SolverDiffPhD(
        model, density=1000.0, rel_tol=1e-6, anderson_window=5, independent_bodies=True,
        factor_precision=args.precision, iteration_chunk=args.chunk,
    )
    solver.prepare(dt)
    log(f"[bench] batched setup {time.time() - t0:.1f}s, nnz(S) {solver.info.factor_nnz}")
    s_bytes = 8 if args.precision == 64 else 4
    if args.probe_only:
        z_all = float(np.sum(solver.body_factor_nnz()))
        probe = factor_probe(solver, dev, n_total, z_all, s_bytes, args)
        copy_gbps = copy_ceiling(dev)
        print(json.dumps({"probe_launch_us": probe["launch_s"] * 1e6, "copy_ceiling_gbps": copy_gbps}))
        return

    for _ in range(args.warmup):
        rollout(solver, model, 1, dt, idx, tgt)
    torch.cuda.synchronize()
    solver.read_status()
    model.tet_materials.grad = None

    solver.iteration_log = []
    torch.cuda.synchronize()
    start = time.perf_counter()
    loss, q0, qd0 = rollout(solver, model, args.steps, dt, idx, tgt)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    status = solver.read_status()
    log(f"[bench] batched rollout {elapsed:.3f}s for {args.steps} fwd+adj steps, loss {float(loss.detach()):.6e}, status {status}")
    fwd_iters = np.array([it for kind, it in solver.iteration_log if kind == "forward"])
    adj_iters = np.array([it for kind, it in solver.iteration_log if kind == "adjoint"])
    solver.iteration_log = None
    grad_ok = bool(torch.isfinite(model.tet_materials.grad).all()) and bool(torch.isfinite(qd0.grad).all())

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

    probe = factor_probe(solver, dev, n_total, z_total, s_bytes, args)
    copy_gbps = copy_ceiling(dev)

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
        cpu = cpu_baseline(bodies, z_b, dt, args.cpu_sample_bodies, solver)

    # HBM traffic per launch of the probe kernels from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE runs (calibrated against a plain copy kernel).
    traffic = None
    traffic_file = ROOT / "profiles" / f"traffic_factor_products_batched256_fp{args.precision}.json"
    if traffic_file.exists() and len(bodies) == 256 and (args.kmin, args.kmax) == (4, 16):
        traffic = json.loads(traffic_file.read_text())["traffic_bytes_per_launch"]

    value = args.steps / elapsed
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
        "plain_steps_per_s": plain,
        "speedup_vs_plain": None if plain is None else value / plain,
        "element_iterations_per_s": elem_iters / elapsed,
        "iterations": {
            "forward_mean": float(fwd_iters.mean()),
            "forward_max": int(fwd_iters.max()),
            "adjoint_mean": float(adj_iters.mean()),
            "adjoint_max": int(adj_iters.max()),
        },
        "rollout_hbm_gbps": rollout_gbps,
        "status": status,
        "gradients_finite": grad_ok,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_factor_rows + k_factor_cols (S and S^T products of all bodies), per launch",
            "achieved": probe["launch_bytes"] / probe["launch_s"] / 1e9,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": probe["launch_bytes"] / probe["launch_s"] / 1e9 / HBM_PEAK_GBPS,
            "traffic": traffic,
            "launch_us": probe["launch_s"] * 1e6,
            "bytes_per_launch": probe["launch_bytes"],
            "stored_bytes_per_launch": probe["stored_bytes"],
            "stored_gbps": probe["stored_bytes"] / probe["launch_s"] / 1e9,
            "copy_ceiling_gbps": copy_gbps,
            "frac_of_copy": probe["launch_bytes"] / probe["launch_s"] / 1e9 / copy_gbps,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(result))


def cpu_baseline(bodies, z_b, dt, sample_count, solver):
    """FP64 C oracle on the host cores: one fwd+adj step from rest of every body (or of a sample scaled by
    factor size to the whole batch)."""
    from diffphd_amd import _lib
    from oracle import oracle

    threads = min(16, os.cpu_count() or 1)
    if sample_count <= 0 or sample_count >= len(bodies):
        pick = np.arange(len(bodies))
    else:
        order = np.argsort([b["k"] for b in bodies])
        pick = order[np.linspace(0, len(order) - 1, sample_count).astype(int)]
    seconds = 0.0
    for i in pick:
        b = bodies[i]
        body = oracle.OracleBody(
            "tet4", b["t"], poses=b["p"], materials=b["m"], density=1000.0, inverse_mass=b["free"].astype(np.float32),
            gravity=(0.0, 0.0, -9.81), dt=dt, rel_tol=1e-6, threads=threads,
            factor=lambda csr: _lib.Factor(*csr).as_dict(),
        )
        q = b["v"].astype(np.float32)
        qd = np.zeros_like(q)
        t0 = time.perf_counter()
        fw = body.forward(dt, q, qd)
        gq = np.zeros_like(q)
        gq[b["top"]] = 2.0 * (fw["q"][b["top"]] - b["target"])
        body.backward(dt, fw["x"], gq, None)
        seconds += time.perf_counter() - t0
    scale = float(z_b.sum() / z_b[pick].sum())
    per_step = seconds * scale
    what = "all" if len(pick) == len(bodies) else f"{len(pick)} (k={[bodies[i]['k'] for i in pick]}), scaled by nnz(S) x{scale:.1f}"
    return {
        "value": 1.0 / per_step,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{what} of {len(bodies)} bodies, one fwd+adj step each from rest, {seconds:.1f} s of oracle time",
    }


if __name__ == "__main__":
    main()
