"""GPU parity of the native DiffPhD path against the golden vectors and the C oracle.

Every test runs the HIP kernels through the C-ABI (via SolverDiffPhD or the
dphd_eval_* entry points) and compares with (a) golden vectors written from
the reference's own NumPy oracle (tests/golden/make_golden.py) and (b) the
FP64 C restatement in oracle/ on identical inputs.
"""

import numpy as np
import pytest
import torch

from diffphd_amd import meshes
from oracle import oracle as O
from scenes import cube_body, golden, tet_model

pytestmark = pytest.mark.gpu


def _lib():
    from diffphd_amd import _lib

    return _lib


def _solver(model, **kw):
    from diffphd_amd import SolverDiffPhD

    return SolverDiffPhD(model, **kw)


def _states(model, q, qd, f=None, requires_grad=False):
    dev = model.device
    s_in = model.state()
    s_in.particle_q = torch.as_tensor(np.asarray(q, np.float32), device=dev).clone()
    s_in.particle_qd = torch.as_tensor(np.asarray(qd, np.float32), device=dev).clone()
    s_in.particle_f = torch.as_tensor(np.zeros_like(q) if f is None else np.asarray(f, np.float32), device=dev).clone()
    for t in (s_in.particle_q, s_in.particle_qd, s_in.particle_f):
        t.requires_grad_(requires_grad)
    return s_in, model.state()


# -- pointwise projections ---------------------------------------------------


def test_projections_match_golden_and_oracle(cuda):
    L = _lib()
    g = golden("projections")
    f = torch.as_tensor(g["f"], device=cuda).contiguous()
    dirs = torch.as_tensor(g["directions"], device=cuda).contiguous()
    c = len(f)
    outs = [torch.empty((c, 3), dtype=torch.float64, device=cuda)] + [
        torch.empty((c, 3, 3), dtype=torch.float64, device=cuda) for _ in range(4)
    ] + [torch.empty((c, 3), dtype=torch.float64, device=cuda)]
    mu, lame = float(g["mu"]), float(g["lame"])
    L.check(
        L.lib().dphd_eval_projections(c, L.ptr(f), L.ptr(dirs), mu, lame, 2.0 * mu, 0.8, *[L.ptr(o) for o in outs], None)
    )
    torch.cuda.synchronize()
    sigma, recon, rot, vol, nh, fib = (o.cpu().numpy() for o in outs)
    np.testing.assert_allclose(sigma, g["sigma"], rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(recon, g["f"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(rot, g["rotation"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(vol, g["volume"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(nh, g["neo_hookean"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(fib, g["fibre"], rtol=0, atol=1e-13)
    ref = O.projections(g["f"], g["directions"], mu, lame, 2.0 * mu, 0.8)
    for name, got in zip(("sigma", "rotation", "volume", "neo_hookean", "fibre"), (sigma, rot, vol, nh, fib)):
        np.testing.assert_allclose(got, ref[name], rtol=0, atol=1e-12, err_msg=name)


def test_degenerate_svd_follows_eigen_conventions(cuda):
    L = _lib()
    f = np.array([[[0.3, 0.5, 1.2], [1.0, 0.5, 0.3], [0.0, 0.0, 0.0]], np.diag([1.0, 0.0, 1.0])])
    c = len(f)
    ft = torch.as_tensor(f, device=cuda).contiguous()
    dirs = torch.as_tensor(np.eye(3)[:2], device=cuda).contiguous()
    outs = [torch.empty((c, 3), dtype=torch.float64, device=cuda)] + [
        torch.empty((c, 3, 3), dtype=torch.float64, device=cuda) for _ in range(4)
    ] + [torch.empty((c, 3), dtype=torch.float64, device=cuda)]
    L.check(L.lib().dphd_eval_projections(c, L.ptr(ft), L.ptr(dirs), 3e4, 8e4, 6e4, 0.8, *[L.ptr(o) for o in outs], None))
    torch.cuda.synchronize()
    sigma, recon, rot = outs[0].cpu().numpy(), outs[1].cpu().numpy(), outs[2].cpu().numpy()
    for i, m in enumerate(f):
        np.testing.assert_allclose(recon[i], m, atol=1e-14)
        np.testing.assert_allclose(rot[i] @ rot[i].T, np.eye(3), atol=1e-14)
        np.testing.assert_allclose(sigma[i], np.linalg.svd(m, compute_uv=False), atol=1e-14)
    # Eigen keeps U = V = [e0, e2, e1] for diag(1, 0, 1), so R = I (tests/test_diff_phd_solver.py:203-204).
    np.testing.assert_allclose(rot[1], np.eye(3), atol=1e-15)


def test_jvps_match_oracle_and_finite_differences(cuda):
    L = _lib()
    rng = np.random.default_rng(13)
    f = np.array([np.diag([1.3, 0.9, 0.7]) + 0.2 * rng.standard_normal((3, 3)) for _ in range(10)])
    df = rng.standard_normal((10, 3, 3))
    d = rng.standard_normal((10, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    mu, lame = 3.0e4, 8.0e4
    c = len(f)
    T = lambda a: torch.as_tensor(a, device=cuda).contiguous()  # noqa: E731
    outs = [torch.empty((c, 3, 3), dtype=torch.float64, device=cuda) for _ in range(3)] + [
        torch.empty((c, 3), dtype=torch.float64, device=cuda)
    ]
    # Keep the inputs referenced: a temporary would return its block to the
    # caching allocator before the kernel reads it.
    ft, dft, dt_ = T(f), T(df), T(d)
    L.check(L.lib().dphd_eval_jvps(c, L.ptr(ft), L.ptr(dft), L.ptr(dt_), mu, lame, 0.8, *[L.ptr(o) for o in outs], None))
    torch.cuda.synchronize()
    got = [o.cpu().numpy() for o in outs]
    ref = O.jvps(f, df, d, mu, lame, 0.8)
    for name, a in zip(("rotation", "volume", "neo_hookean", "fibre"), got):
        scale = max(1.0, np.abs(ref[name]).max())
        np.testing.assert_allclose(a, ref[name], rtol=0, atol=1e-10 * scale, err_msg=name)
    eps = 1e-6
    plus = O.projections(f + eps * df, d, mu, lame, 2 * mu, 0.8)
    minus = O.projections(f - eps * df, d, mu, lame, 2 * mu, 0.8)
    for name, a in zip(("rotation", "volume", "neo_hookean", "fibre"), got):
        numeric = (plus[name] - minus[name]) / (2 * eps)
        np.testing.assert_allclose(a, numeric, rtol=0, atol=2e-6 * max(1.0, np.abs(numeric).max()), err_msg=name)


# -- forward step against the reference's dense oracle ------------------------

FORWARD_CASES = [
    "tet_fwd_fixed_pd",
    "tet_fwd_fixed_nh",
    "tet_fwd_conv_c10_pd",
    "tet_fwd_conv_c10_nh",
    "tet_fwd_conv_c100_pd",
    "tet_fwd_conv_c100_nh",
    "tet_fwd_muscle_pd",
]


@pytest.mark.parametrize("name", FORWARD_CASES)
def test_forward_matches_reference_oracle(cuda, name):
    g = golden(name)
    print(f"[case] forward golden {name}", flush=True)
    material = "pd" if name.endswith("_pd") else "neo_hookean"
    model = tet_model(cuda, g["tets"], g["poses"], g["materials"], g["free"], g["q"], g["gravity"])
    kw = {}
    control = None
    if "directions" in g:
        kw = dict(muscle_directions=g["directions"], muscle_stiffness=g["stiffness"])
        control = model.control()
        control.tet_activations = torch.as_tensor((1.0 - g["targets"]).astype(np.float32), device=cuda)
    solver = _solver(
        model,
        density=float(g["density"]),
        material=material,
        max_iterations=int(g["max_iterations"]),
        abs_tol=float(g["abs_tol"]),
        rel_tol=float(g["rel_tol"]),
        **kw,
    )
    s_in, s_out = _states(model, g["q"], g["qd"], g["f"])
    solver.step(s_in, s_out, control, None, float(g["dt"]))
    x = solver.solution().cpu().numpy()
    bits = solver.read_status()
    assert bits == (0 if bool(g["converged"]) else 1)
    scale = np.abs(g["x"]).max()
    np.testing.assert_allclose(x, g["x"], rtol=0, atol=1e-10 * scale)
    assert int(solver.iterations[0]) == int(g["iterations"])
    np.testing.assert_array_equal(s_out.particle_q.cpu().numpy(), x.astype(np.float32))


@pytest.mark.parametrize("name", ["tet_bwd_pd_c2", "tet_bwd_pd_qd", "tet_bwd_nh", "tet_bwd_muscle_pd"])
def test_backward_matches_reference_oracle(cuda, name):
    g = golden(name)
    print(f"[case] backward golden {name}", flush=True)
    material = str(g["material"])
    model = tet_model(cuda, g["tets"], g["poses"], g["materials"], g["free"], g["q"], g["gravity"], requires_grad=True)
    kw, control = {}, None
    if "targets" in g:
        kw = dict(muscle_directions=g["directions"], muscle_stiffness=g["stiffness"])
        control = model.control()
        control.tet_activations = torch.as_tensor(
            (1.0 - g["targets"]).astype(np.float32), device=cuda
        ).requires_grad_(True)
    solver = _solver(
        model,
        density=float(g["density"]),
        material=material,
        max_iterations=int(g["max_iterations"]),
        abs_tol=float(g["abs_tol"]),
        rel_tol=float(g["rel_tol"]),
        **kw,
    )
    s_in, s_out = _states(model, g["q"], g["qd"], g["f"], requires_grad=True)
    solver.step(s_in, s_out, control, None, float(g["dt"]))
    if control is not None:
        scale = np.abs(g["x"]).max()
        np.testing.assert_allclose(s_out.particle_q.detach().cpu().numpy(), g["x"], rtol=0, atol=1e-6 * scale)
    torch.autograd.backward(
        [s_out.particle_q, s_out.particle_qd],
        [torch.as_tensor(g["gq"], device=cuda), torch.as_tensor(g["gqd"], device=cuda)],
    )
    bits = solver.read_status()
    assert bits & 8 == 0
    free = g["free"]
    for key, got in (("qd", s_in.particle_qd.grad), ("f", s_in.particle_f.grad)):
        ref = g[f"out_{key}"].reshape(-1, 3)
        np.testing.assert_allclose(got.cpu().numpy(), ref, rtol=0, atol=2e-5 * np.abs(ref).max(), err_msg=key)
    ref_q = g["out_q"].reshape(-1, 3)
    np.testing.assert_allclose(
        s_in.particle_q.grad.cpu().numpy()[free], ref_q[free], rtol=0, atol=2e-5 * np.abs(ref_q).max()
    )
    if material == "pd":
        ref_m = g["out_materials"]
        got_m = model.tet_materials.grad.cpu().numpy()[:, :2]
        np.testing.assert_allclose(got_m, ref_m, rtol=0, atol=2e-5 * np.abs(ref_m).max())
    if control is not None:
        # d/du of the contraction; the golden's disabled fibres are exactly zero.
        ref_a = g["out_activations"]
        got_a = control.tet_activations.grad.cpu().numpy()
        assert np.all(got_a[g["stiffness"] == 0] == 0)
        np.testing.assert_allclose(got_a, ref_a, rtol=0, atol=2e-5 * np.abs(ref_a).max())


@pytest.mark.parametrize("name", ["hex_fwd_pd", "hex_fwd_nh"])
def test_hex_forward_matches_isoparametric_oracle(cuda, name):
    from diffphd_amd import Mesh
    from diffphd_amd.model import Model

    g = golden(name)
    print(f"[case] hex golden {name}", flush=True)
    material = "pd" if name.endswith("_pd") else "neo_hookean"
    n = len(g["vertices"])
    model = Model(
        particle_q=g["q"], particle_qd=g["qd"], particle_inv_mass=g["free"].astype(np.float32),
        particle_world=np.zeros(n, np.int32), gravity=np.zeros((1, 3)), tet_indices=None, tet_poses=None,
        tet_materials=None, device=cuda,
    )
    solver = _solver(
        model, density=float(g["density"]), material=material, hex_mesh=Mesh("two-cubes", g["vertices"], g["cells"]),
        hex_materials=g["materials"], anderson_window=0, max_iterations=int(g["max_iterations"]),
        rel_tol=float(g["rel_tol"]), abs_tol=float(g["abs_tol"]),
    )
    s_in, s_out = _states(model, g["q"], g["qd"], g["f"])
    solver.step(s_in, s_out, None, None, float(g["dt"]))
    assert solver.read_status() == 0
    np.testing.assert_allclose(s_out.particle_q.cpu().numpy(), g["x"], rtol=0, atol=2e-8)
    assert int(solver.iterations[0]) == int(g["iterations"])
    rhs = torch.eye(n, dtype=torch.float64, device=cuda)
    for axis in range(0, n, 3):
        cols = rhs[:, axis : axis + 3].contiguous()
        got = solver.apply_inverse(cols).cpu().numpy()
        np.testing.assert_allclose(got, g["inverse"][:, axis : axis + 3], atol=2e-14, rtol=2e-12)


# -- mid-size scenes against the C oracle ------------------------------------


def _oracle_body(vertices, tets, poses, materials, free, dt, **kw):
    return O.OracleBody(
        "tet4", tets, poses=poses, materials=materials, density=1000.0, inverse_mass=free.astype(np.float32),
        gravity=(0.0, 0.0, -9.81), dt=dt, **kw,
    )


def _rollout_gpu(solver, model, q, steps, dt):
    s_in, s_out = _states(model, q, np.zeros_like(q))
    xs, iters = [], []
    for _ in range(steps):
        solver.step(s_in, s_out, None, None, dt)
        xs.append(solver.solution().cpu().numpy().copy())
        iters.append(solver.iterations.numpy().copy())
        s_in, s_out = s_out, s_in
    return xs, iters


@pytest.mark.parametrize("precision", [64, 32])
def test_cube_rollout_matches_oracle(cuda, precision):
    vertices, tets, poses, materials, free = cube_body(5, seed=1)
    dt, steps = 1e-3, 4
    print(f"[case] 5^3 cube rollout fp{precision}", flush=True)
    model = tet_model(cuda, tets, poses, materials, free, vertices, (0, 0, -9.81))
    solver = _solver(model, density=1000.0, rel_tol=1e-6, factor_precision=precision)
    xs, iters = _rollout_gpu(solver, model, vertices, steps, dt)
    assert solver.read_status() == 0
    body = _oracle_body(vertices, tets, poses, materials, free, dt, rel_tol=1e-6)
    q, qd = vertices.astype(np.float32), np.zeros_like(vertices, dtype=np.float32)
    for t in range(steps):
        out = body.forward(dt, q, qd)
        if precision == 64:
            # Same iteration path as the oracle: agreement far below the gate.
            atol = 1e-9 * np.abs(out["x"] - vertices).max() + 1e-12
            slack = 1
        else:
            # FP32 factor: positions within 1e-5 relative (the gate itself is
            # relative to |x|, so iterates may stop ~rel_tol*|x| apart).
            atol = 1e-5 * np.abs(out["x"]).max()
            slack = 2
        np.testing.assert_allclose(xs[t], out["x"], rtol=0, atol=atol, err_msg=f"step {t}")
        assert abs(int(iters[t][0]) - out["iterations"]) <= slack, (t, iters[t], out["iterations"])
        q, qd = out["q"], out["qd"]


def test_disconnected_tets_with_scattered_ids_match_oracle(cuda):
    """The local step stages each 256-tet chunk's distinct vertices in LDS and reads corners by 16-bit
    positions: tets that share no vertex, under a random vertex numbering, give chunks of 1024 distinct
    vertices (the largest a chunk can have) spread over the whole id range."""
    rng = np.random.default_rng(5)
    n_tets = 600
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64) * 0.01
    coords = np.zeros((4 * n_tets, 3))
    for t in range(n_tets):
        shape = base + rng.uniform(-0.002, 0.002, size=(4, 3))
        coords[4 * t:4 * t + 4] = shape + np.array([t % 25, (t // 25) % 25, t // 625], np.float64) * 0.03
    perm = rng.permutation(4 * n_tets)
    vertices = np.zeros_like(coords)
    vertices[perm] = coords
    tets = perm.reshape(n_tets, 4).astype(np.int32)
    poses = meshes.tet_poses(vertices, tets)
    young = np.exp(rng.uniform(np.log(1e5), np.log(1e7), n_tets))
    mu, lam = meshes.lame(young, np.full(n_tets, 0.4))
    materials = np.stack([mu, lam, np.zeros_like(mu)], axis=1).astype(np.float32)
    free = np.ones(4 * n_tets, bool)
    free[tets[:, 0]] = False  # one pinned corner per tet
    vertices = vertices.astype(np.float32)
    dt, steps = 1e-3, 3
    print("[case] 600 disconnected tets, scattered vertex ids, fp64", flush=True)
    model = tet_model(cuda, tets, poses, materials, free, vertices, (0, 0, -9.81))
    solver = _solver(model, density=1000.0, rel_tol=1e-6)
    xs, iters = _rollout_gpu(solver, model, vertices, steps, dt)
    assert solver.read_status() == 0
    body = _oracle_body(vertices, tets, poses, materials, free, dt, rel_tol=1e-6)
    q, qd = vertices, np.zeros_like(vertices)
    for t in range(steps):
        out = body.forward(dt, q, qd)
        disp = np.abs(out["x"] - vertices).max()
        err = np.abs(xs[t] - out["x"]).max()
        print(f"[result] step {t}: iterations {int(iters[t][0])} (oracle {out['iterations']}), position error "
              f"{err:.3e} m of a {disp:.3e} m displacement", flush=True)
        # Stiff tets of a fraction of a gram: rounding differences in the iteration reach the gate's
        # level (rel 1e-6 of |x|), not the 1e-9 of the well-conditioned cubes; a misplaced corner would
        # be off by the displacement itself.
        assert err <= 1e-5 * disp + 1e-12, (t, err, disp)
        assert abs(int(iters[t][0]) - out["iterations"]) <= 1, (t, iters[t], out["iterations"])
        q, qd = out["q"], out["qd"]


def test_batched_bodies_match_per_body_oracle(cuda):
    from diffphd_amd.model import Model

    sizes = [3, 5, 4]
    parts, voff, eoff = [], [0], [0]
    for b, k in enumerate(sizes):
        parts.append(cube_body(k, seed=10 + b, origin=(0.1 * b, 0.0, 0.0)))
        voff.append(voff[-1] + len(parts[-1][0]))
        eoff.append(eoff[-1] + len(parts[-1][1]))
    verts = np.concatenate([p[0] for p in parts])
    tets = np.concatenate([p[1] + voff[b] for b, p in enumerate(parts)])
    poses = np.concatenate([p[2] for p in parts])
    mats = np.concatenate([p[3] for p in parts])
    free = np.concatenate([p[4] for p in parts])
    dt = 1e-3
    print("[case] batched 3 bodies", flush=True)
    model = Model(
        particle_q=verts, particle_qd=np.zeros_like(verts), particle_inv_mass=free.astype(np.float32),
        particle_world=np.zeros(len(verts), np.int32), gravity=np.array([[0, 0, -9.81]]), tet_indices=tets,
        tet_poses=poses, tet_materials=mats, device=cuda, body_vertex_offsets=voff, body_element_offsets=eoff,
    )
    solver = _solver(model, density=1000.0, rel_tol=1e-6, independent_bodies=True)
    xs, iters = _rollout_gpu(solver, model, verts, 3, dt)
    assert solver.read_status() == 0
    for b, (v, t_, p, m, fr) in enumerate(parts):
        body = _oracle_body(v, t_, p, m, fr, dt, rel_tol=1e-6)
        q, qd = v.astype(np.float32), np.zeros_like(v, dtype=np.float32)
        for t in range(3):
            out = body.forward(dt, q, qd)
            got = xs[t][voff[b] : voff[b + 1]]
            scale = np.abs(out["x"] - v).max()
            np.testing.assert_allclose(got, out["x"], rtol=0, atol=1e-9 * scale + 1e-12, err_msg=f"body {b} step {t}")
            assert abs(int(iters[t][b]) - out["iterations"]) <= 1
            q, qd = out["q"], out["qd"]


HOMOGENEOUS = (np.log(1e5), np.log(1e5))  # one Young's modulus for every tet
HETEROGENEOUS = (np.log(1e5), np.log(1e7))  # per-tet log-uniform Young's modulus


def _cube_backward(cuda, k, seed, log_e, mode, precision):
    vertices, tets, poses, materials, free = cube_body(k, seed=seed, log_e=log_e)
    dt = 1e-3
    model = tet_model(cuda, tets, poses, materials, free, vertices, (0, 0, -9.81), requires_grad=True)
    solver = _solver(
        model, density=1000.0, rel_tol=1e-10, abs_tol=1e-14, jacobian_mode=mode, factor_precision=precision
    )
    rng = np.random.default_rng(3)
    q = vertices + 1e-4 * rng.standard_normal(vertices.shape) * free[:, None]
    qd = 0.05 * rng.standard_normal(vertices.shape) * free[:, None]
    s_in, s_out = _states(model, q, qd, requires_grad=True)
    solver.step(s_in, s_out, None, None, dt)
    gq = rng.standard_normal(vertices.shape).astype(np.float32)
    gqd = (0.01 * rng.standard_normal(vertices.shape)).astype(np.float32)
    torch.autograd.backward([s_out.particle_q, s_out.particle_qd], [torch.as_tensor(gq, device=cuda), torch.as_tensor(gqd, device=cuda)])
    body = _oracle_body(vertices, tets, poses, materials, free, dt, rel_tol=1e-10, abs_tol=1e-14, jacobian_mode=mode)
    fw = body.forward(dt, q.astype(np.float32), qd.astype(np.float32))
    bw = body.backward(dt, fw["x"], gq, gqd)
    return solver, model, s_in, bw


@pytest.mark.parametrize(
    "mode,precision,log_e",
    [
        ("reference", 64, HOMOGENEOUS),
        ("element", 64, HETEROGENEOUS),
        ("reference", 32, HOMOGENEOUS),
        ("element", 32, HETEROGENEOUS),
    ],
    ids=["reference-fp64", "element-fp64", "reference-fp32", "element-fp32"],
)
def test_cube_backward_matches_oracle(cuda, mode, precision, log_e):
    """Converging adjoints: status 0 on both sides, gradients within 1e-6 of the largest.

    The FP32 factor only changes the iteration path (a defect correction with
    the FP64 operators), so with a 1e-10 gate both precisions reach the same
    fixed point as the FP64 oracle.
    """
    print(f"[case] 4^3 cube backward {mode} fp{precision}", flush=True)
    solver, model, s_in, bw = _cube_backward(cuda, 4, 2, log_e, mode, precision)
    assert solver.read_status() == 0
    assert bw["status"] == 0
    assert abs(int(solver.adjoint_iterations[0]) - bw["iterations"]) <= (1 if precision == 64 else 2)
    for key, got in (("q", s_in.particle_q.grad), ("qd", s_in.particle_qd.grad), ("f", s_in.particle_f.grad)):
        ref = bw[key]
        assert np.abs(ref).max() > 0.0, key
        np.testing.assert_allclose(got.cpu().numpy(), ref, rtol=0, atol=1e-6 * np.abs(ref).max(), err_msg=key)
    ref_m = bw["materials"]
    got_m = model.tet_materials.grad.cpu().numpy()[:, :2]
    assert np.abs(ref_m).max() > 0.0
    np.testing.assert_allclose(got_m, ref_m, rtol=0, atol=1e-6 * np.abs(ref_m).max())


def test_diverging_adjoint_trips_the_same_safeguard_as_the_oracle(cuda):
    """Reference-mode Jacobian weights (per-body means) on per-tet stiffness spread over two
    decades: the adjoint fixed point diverges, and the routing guard (adjoint.py:226-261)
    must fail the body on both sides, zeroing its gradient contributions."""
    print("[case] 4^3 cube backward, diverging adjoint", flush=True)
    solver, model, s_in, bw = _cube_backward(cuda, 4, 2, HETEROGENEOUS, "reference", 64)
    bits = solver.read_status()
    assert bits & 16 and bw["status"] & 16, (bits, bw["status"])
    assert (bits & 8) == (bw["status"] & 8)
    for key, got in (("qd", s_in.particle_qd.grad), ("f", s_in.particle_f.grad)):
        np.testing.assert_allclose(got.cpu().numpy(), bw[key], rtol=0, atol=1e-12, err_msg=key)


def test_element_mode_gradients_match_central_differences(cuda):
    """Exact-mode adjoint vs central FD of the FP64 device forward (per-element weights, heterogeneous E)."""
    vertices, tets, poses, materials, free = cube_body(3, seed=4)
    dt = 1e-3
    print("[case] 3^3 cube FD check", flush=True)
    model = tet_model(cuda, tets, poses, materials, free, vertices, (0, 0, -9.81), requires_grad=True)
    solver = _solver(model, density=1000.0, rel_tol=1e-13, abs_tol=1e-16, max_iterations=5000, jacobian_mode="element")
    rng = np.random.default_rng(5)
    qd0 = (0.05 * rng.standard_normal(vertices.shape) * free[:, None]).astype(np.float32)
    weights = rng.standard_normal(vertices.shape)

    def loss_of(mats, qd):
        with torch.no_grad():
            model.tet_materials.copy_(torch.as_tensor(mats, device=cuda))
        solver.notify_model_changed()
        s_in, s_out = _states(model, vertices, qd)
        with torch.no_grad():
            solver.step(s_in, s_out, None, None, dt)
        return float((solver.solution().cpu().numpy() * weights).sum())

    s_in, s_out = _states(model, vertices, qd0, requires_grad=True)
    solver.notify_model_changed()
    solver.step(s_in, s_out, None, None, dt)
    torch.autograd.backward(s_out.particle_q, torch.as_tensor(weights.astype(np.float32), device=cuda))
    grad_m = model.tet_materials.grad.cpu().numpy().copy()
    grad_v = s_in.particle_qd.grad.cpu().numpy().copy()
    base = materials.astype(np.float64)
    for e, col, h in ((0, 0, 20.0), (7, 1, 50.0), (33, 0, 20.0)):
        mp, mm = base.copy(), base.copy()
        mp[e, col] += h
        mm[e, col] -= h
        fd = (loss_of(mp.astype(np.float32), qd0) - loss_of(mm.astype(np.float32), qd0)) / (
            np.float32(mp[e, col]) - np.float32(mm[e, col])
        )
        assert abs(grad_m[e, col] - fd) <= 1e-4 * max(abs(fd), 1e-8), (e, col, grad_m[e, col], fd)
    loss_of(base.astype(np.float32), qd0)
    node = int(np.flatnonzero(free)[5])
    for axis in (0, 2):
        qp, qm = qd0.copy(), qd0.copy()
        qp[node, axis] += 1e-3
        qm[node, axis] -= 1e-3
        fd = (loss_of(base.astype(np.float32), qp) - loss_of(base.astype(np.float32), qm)) / (qp[node, axis] - qm[node, axis])
        assert abs(grad_v[node, axis] - fd) <= 1e-4 * max(abs(fd), 1e-8), (axis, grad_v[node, axis], fd)
