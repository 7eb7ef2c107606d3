"""Parity fixtures of the reference's harness (schema ``newton.diff_phd_parity.v1``).

The reference compares a port with the executed C++ DiffPhD by exchanging NPZ
files (docs/validation.md:451-472): a fixture holds one scene's inputs (SI
units, canonical particle order, positively oriented Tet4, every value rounded
to float32 and stored as float64), each side writes its outputs under fixed
names, and tools/parity/compare.py checks them against fixed thresholds. This
module is the port side of that exchange for this solver:

* ``run(fixture)`` steps the fixture's scene on the GPU solver, back-propagates
  its linear objective, and returns the outputs under the names the reference
  port writes (tools/parity/port_cases.py:413-455): ``q_next``, ``v_next``,
  ``iterations``, ``converged``, ``status_bits``, ``dL_dq``, ``dL_dv``,
  ``dL_df_ext``, ``dL_dmaterials_per_element``, ``loss``, ``free``, and where
  they apply ``multipliers``/``active``/``gaps``, ``mesh_contact_force``/
  ``mesh_contact_pairs`` and ``dL_dactivations``;
* ``check(fixture)`` lists why a fixture cannot be run (the acceptance rules of
  compare.py:63-101, which both sides must implement);
* ``make_fixture`` and ``cases()`` write fixtures of the same form, including
  the reference harness's named scenes built on this package's grids.

Command line: ``python -m diffphd_amd.parity --out DIR`` writes
``<case>.fixture.npz`` and ``<case>.port.npz`` for the named cases, or, with
``--fixtures F.fixture.npz ...``, the port outputs of existing fixtures.
"""

from __future__ import annotations

import argparse
import hashlib
import json
import platform
import subprocess
import sys
from pathlib import Path

import numpy as np

from . import meshes

SCHEMA = "newton.diff_phd_parity.v1"
ACTIVATION_CONVENTION = "contraction u; reference target fibre length a = 1 - u"
# Tolerances of the harness's fixtures (port_cases.py:32).
TIGHT = {"abs_tol": 1.0e-12, "rel_tol": 1.0e-11, "max_iterations": 2000, "aa_window": 5}
FLOAT32_KEYS = ("rest_vertices", "mu", "lambda", "q", "v", "f_ext", "gravity", "objective_q", "objective_v")


def _f32(values):
    return np.asarray(values, dtype=np.float32).astype(np.float64)


def _value(x):
    """A 0-d array read from an NPZ file as its Python scalar (str, int, float or bool)."""
    a = np.asarray(x)
    return a.item() if a.ndim == 0 else a


def make_fixture(case, rest_vertices, tets, mu, lame, dirichlet, *, steps=1, dt=0.004, density=1000.0,
                 material="pd", q=None, v=None, f_ext=None, gravity=(0.0, 0.0, -9.81), seed=0, **extra):
    """A fixture dict with the schema's inputs; unset inputs take the harness's defaults (rest state,
    zero velocity and force, no contact, tight tolerances, a random linear objective of ``seed``).
    ``extra`` sets any other schema key (contact, muscle, Dirichlet-target, tolerance fields)."""
    rest = np.asarray(rest_vertices, dtype=np.float64)
    n = len(rest)
    rng = np.random.default_rng(seed)
    fixture = {
        "case": str(case),
        "rest_vertices": rest,
        "tets": np.asarray(tets, dtype=np.int32),
        "density": float(density),
        "mu": np.asarray(mu, dtype=np.float64),
        "lambda": np.asarray(lame, dtype=np.float64),
        "material": str(material),
        "dirichlet": np.asarray(dirichlet, dtype=bool),
        "q": rest.copy() if q is None else np.asarray(q, dtype=np.float64),
        "v": np.zeros((n, 3)) if v is None else np.asarray(v, dtype=np.float64),
        "f_ext": np.zeros((steps, n, 3)) if f_ext is None else np.asarray(f_ext, dtype=np.float64),
        "gravity": np.asarray(gravity, dtype=np.float64),
        "dt": float(dt),
        "steps": int(steps),
        **TIGHT,
        "contact_kind": "none",
        "contact_parameters": np.zeros(4),
        "contact_vertices": np.zeros(0, dtype=np.int32),
        "friction": 0.0,
        "restitution": 0.0,
        "objective_q": rng.standard_normal((n, 3)),
        "objective_v": 0.01 * rng.standard_normal((n, 3)),
        "reference_material_encoding": "per_element",
        "benchmark_steps": 0,
    }
    fixture.update(extra)
    return finalize(fixture)


PACKAGE = Path(__file__).resolve().parents[1]  # newton-solver-diff-phd_amd/


def source_hash():
    """SHA-256 over this package's sources (Python and HIP/C++), the fixtures' ``port_source_sha256``."""
    digest = hashlib.sha256()
    for path in sorted([*PACKAGE.rglob("*.py"), *PACKAGE.rglob("*.hip"), *PACKAGE.rglob("*.cpp"),
                        *PACKAGE.rglob("*.h")]):
        if "__pycache__" in path.parts or "_build" in path.parts:
            continue
        digest.update(path.relative_to(PACKAGE).as_posix().encode())
        digest.update(path.read_bytes())
    return digest.hexdigest()


def revision():
    """The repository's git revision (``+uncommitted`` if dirty), or ``unknown`` without git."""
    try:
        head = subprocess.run(["git", "-C", str(PACKAGE), "rev-parse", "HEAD"], capture_output=True, text=True,
                              check=True).stdout.strip()
        dirty = subprocess.run(["git", "-C", str(PACKAGE), "status", "--porcelain"], capture_output=True,
                               text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return "unknown"
    return head + ("+uncommitted" if dirty else "")


def finalize(fixture):
    """Round the float inputs to the float32 values Newton stores (kept as float64) and stamp the
    schema and the port's revision and source hash (which compare.py records)."""
    out = dict(fixture)
    for key in FLOAT32_KEYS + ("activations", "dirichlet_positions"):
        if key in out:
            out[key] = _f32(out[key])
    out["schema"] = SCHEMA
    out["activation_convention"] = ACTIVATION_CONVENTION
    out.setdefault("port_revision", revision())
    out.setdefault("port_source_sha256", source_hash())
    return out


def environment(device):
    """The port environment record compare.py reads from ``port_environment.json``."""
    import torch

    info = {"python": sys.version.split()[0], "platform": platform.platform(), "machine": platform.machine(),
            "torch": torch.__version__, "hip": getattr(torch.version, "hip", None), "device": str(device),
            "solver": "diffphd_amd (MI355X HIP)"}
    if torch.cuda.is_available():
        info["device_name"] = torch.cuda.get_device_name(torch.device(device))
    return info


def _exact_f32(values):
    a = np.asarray(values, dtype=np.float64)
    return np.array_equal(a, a.astype(np.float32).astype(np.float64))


def _tet_rows(fx):
    tets = np.asarray(fx["tets"])
    return tets.ndim == 2 and tets.shape[1] == 4


def _tets_in_range(fx):
    tets = np.asarray(fx["tets"])
    return len(tets) == 0 or (tets.min() >= 0 and tets.max() < len(np.asarray(fx["rest_vertices"])))


def _positively_oriented(fx):
    rest = np.asarray(fx["rest_vertices"], dtype=np.float64)
    tets = np.asarray(fx["tets"])
    edges = rest[tets[:, 1:]] - rest[tets[:, :1]]  # [E, 3 edges, 3]
    return bool(np.all(np.linalg.det(edges) > 0.0))


def _mesh_contact_faces_ok(fx):
    n = len(np.asarray(fx["q"]))
    for key in ("mesh_contact_faces_a", "mesh_contact_faces_b"):
        faces = np.asarray(fx[key])
        if faces.ndim != 2 or faces.shape[1] != 3 or len(faces) == 0 or faces.min() < 0 or faces.max() >= n:
            return False
    return True


def _per_step(fx, key):
    return np.asarray(fx[key]).shape == (int(_value(fx["steps"])), len(np.asarray(fx["q"])), 3)


def _nonnegative_finite(values, shape):
    a = np.asarray(values, dtype=np.float64)
    return a.shape == shape and bool(np.all(np.isfinite(a))) and bool(np.all(a >= 0.0))


MATERIALS = ("pd", "neo_hookean")

# What a fixture of schema newton.diff_phd_parity.v1 must satisfy before it is
# run (docs/validation.md:451-472), with what this solver runs (Tet4, "pd" or
# "neo_hookean", the four contact kinds). Each rule: the reason reported when
# it fails, the keys that make it apply (an optional input is checked only
# when present), the rules that must pass first, and the predicate.
_RULES = [
    ("schema", "schema", (), (), lambda fx: str(_value(fx.get("schema", ""))) == SCHEMA),
    ("tet_rows", "only Tet4 topology", ("tets",), (), _tet_rows),
    ("tet_range", "tet indices out of range", ("tets", "rest_vertices"), ("tet_rows",), _tets_in_range),
    ("orientation", "tetrahedra must be positively oriented", ("tets", "rest_vertices"), ("tet_range",),
     _positively_oriented),
    ("material", "material", ("material",), (), lambda fx: str(_value(fx["material"])) in MATERIALS),
    ("contact", "contact kind", ("contact_kind",), (),
     lambda fx: str(_value(fx["contact_kind"])) in ("none", "planar", "spherical", "mesh")),
    ("mesh_params", "mesh_contact parameters (contact_radius, kn_left, kn_right, kf, mu)",
     ("mesh_contact_parameters",), (), lambda fx: _nonnegative_finite(fx["mesh_contact_parameters"], (5,))),
    ("mesh_faces", "mesh_contact faces", ("mesh_contact_parameters",), (), _mesh_contact_faces_ok),
    ("convention", "activation convention", ("activations",), (),
     lambda fx: "a = 1 - u" in str(_value(fx.get("activation_convention", "")))),
    ("targets", "dirichlet_positions shape or finite values", ("dirichlet_positions",), (),
     lambda fx: _per_step(fx, "dirichlet_positions") and bool(np.all(np.isfinite(fx["dirichlet_positions"])))),
    ("targets_f32", "dirichlet_positions is not float32-representable", ("dirichlet_positions",), ("targets",),
     lambda fx: _exact_f32(fx["dirichlet_positions"])),
    ("damping", "velocity_damping must be a finite scalar in [0,1]", ("velocity_damping",), (),
     lambda fx: np.size(fx["velocity_damping"]) == 1
     and 0.0 <= float(np.asarray(fx["velocity_damping"], dtype=np.float64).reshape(-1)[0]) <= 1.0),
    ("forces", "f_ext must be (steps, N, 3)", ("f_ext",), (), lambda fx: _per_step(fx, "f_ext")),
] + [
    (f"{key}_f32", f"{key} is not float32-representable (Newton stores float32)", (key,), (),
     lambda fx, key=key: _exact_f32(fx[key]))
    for key in FLOAT32_KEYS
]


def check(fixture):
    """Reasons the fixture cannot be run (the schema's rules above, in order); empty when it can."""
    passed, reasons = set(), []
    for name, reason, keys, after, ok in _RULES:
        if not all(k in fixture for k in keys) or not all(a in passed for a in after):
            continue
        if ok(fixture):
            passed.add(name)
        else:
            reasons.append(reason)
    return reasons


def load(path):
    """A fixture or result file as a dict (no pickles: ``allow_pickle=False``)."""
    with np.load(path, allow_pickle=False) as data:
        return {key: _value(data[key]) for key in data.files}


def save(path, values):
    np.savez(path, **values)


def mesh_boundary(parameters):
    """``("mesh", ...)`` frictional-boundary parameters from the flat fixture list
    ``[radius, split_index, left_count, right_count, faces_left..., faces_right...]``."""
    parameters = np.asarray(parameters, dtype=np.float64)
    radius, split, left, right = float(parameters[0]), int(parameters[1]), int(parameters[2]), int(parameters[3])
    faces = np.asarray(parameters[4:], dtype=np.int64).reshape(-1, 3)
    if len(faces) != left + right:
        raise ValueError("mesh boundary face counts do not match")
    return radius, split, faces[:left], faces[left:]


def build(fixture, device="cuda:0", factor_precision=64):
    """The model and solver of a fixture (one body; particles of mass 0 are the Dirichlet set)."""
    import torch

    from .model import Model
    from .solver import SolverDiffPhD

    fx = fixture
    rest = np.asarray(fx["rest_vertices"], dtype=np.float64)
    tets = np.asarray(fx["tets"], dtype=np.int32)
    dirichlet = np.asarray(fx["dirichlet"], dtype=bool)
    materials = np.stack([fx["mu"], fx["lambda"], np.zeros(len(tets))], axis=1).astype(np.float32)
    model = Model(
        particle_q=rest.astype(np.float32),
        particle_qd=np.zeros((len(rest), 3), np.float32),
        particle_inv_mass=np.where(dirichlet, 0.0, 1.0).astype(np.float32),
        particle_world=np.zeros(len(rest), np.int32),
        gravity=np.asarray(fx["gravity"], dtype=np.float32)[None],
        tet_indices=tets,
        tet_poses=meshes.tet_poses(rest, tets),
        tet_materials=materials,
        device=torch.device(device),
        requires_grad=True,
    )
    options = {
        "density": float(_value(fx["density"])),
        "material": str(_value(fx["material"])),
        "max_iterations": int(_value(fx["max_iterations"])),
        "abs_tol": float(_value(fx["abs_tol"])),
        "rel_tol": float(_value(fx["rel_tol"])),
        "anderson_window": int(_value(fx["aa_window"])),
        "factor_precision": factor_precision,
    }
    if "muscle_directions" in fx:
        options["muscle_directions"] = np.asarray(fx["muscle_directions"], dtype=np.float64)
        options["muscle_stiffness"] = np.asarray(fx["muscle_stiffness"], dtype=np.float64)
    kind = str(_value(fx["contact_kind"]))
    if kind != "none":
        params = mesh_boundary(fx["contact_parameters"]) if kind == "mesh" else tuple(
            float(p) for p in np.asarray(fx["contact_parameters"], dtype=np.float64))
        options.update(
            frictional_boundary=(kind, params),
            frictional_vertices=np.asarray(fx["contact_vertices"], dtype=np.int32),
            friction=float(_value(fx["friction"])),
            restitution=float(_value(fx["restitution"])),
        )
    if "mesh_contact_parameters" in fx:
        options["mesh_contact"] = (
            *(float(p) for p in np.asarray(fx["mesh_contact_parameters"], dtype=np.float64)),
            np.asarray(fx["mesh_contact_faces_a"], dtype=np.int32),
            np.asarray(fx["mesh_contact_faces_b"], dtype=np.int32),
        )
    return model, SolverDiffPhD(model, **options)


def run(fixture, device="cuda:0", factor_precision=64):
    """Outputs of the fixture's taped rollout on this solver, named as the reference port's
    (port_cases.py:374-455): per-step converged FP64 positions, velocities (x - q_in) / dt (times
    the fixture's velocity damping), iterations, convergence and status bits (the backward's last),
    the objective's gradients with respect to the initial state, every step's external force, the
    per-element materials and, with muscles, every step's activations."""
    import torch

    reasons = check(fixture)
    if reasons:
        raise ValueError("fixture cannot be run: " + "; ".join(reasons))
    fx = fixture
    model, solver = build(fx, device, factor_precision)
    dev = model.device
    steps, dt = int(_value(fx["steps"])), float(_value(fx["dt"]))
    damping = float(_value(fx["velocity_damping"])) if "velocity_damping" in fx else None
    q = torch.tensor(np.asarray(fx["q"], np.float32), device=dev, requires_grad=True)
    v = torch.tensor(np.asarray(fx["v"], np.float32), device=dev, requires_grad=True)
    forces = [torch.tensor(np.asarray(fx["f_ext"][k], np.float32), device=dev, requires_grad=True)
              for k in range(steps)]
    acts = None
    if "activations" in fx:
        acts = [torch.tensor(np.asarray(fx["activations"][k], np.float32), device=dev, requires_grad=True)
                for k in range(steps)]
    state = model.state()
    state.particle_q, state.particle_qd = q, v
    q_in, xs, iterations, status, multipliers, mesh_force, mesh_pairs = [], [], [], [], [], [], []
    for k in range(steps):
        state.particle_f = forces[k]
        control = None
        if acts is not None:
            control = model.control()
            control.tet_activations = acts[k]
        targets = None
        if "dirichlet_positions" in fx:
            targets = torch.tensor(np.asarray(fx["dirichlet_positions"][k], np.float32), device=dev)
        out = model.state()
        q_in.append(state.particle_q.detach().to(torch.float64).cpu().numpy())
        solver.step(state, out, control, None, dt, dirichlet_positions=targets)
        xs.append(solver.solution().cpu().numpy().copy())
        iterations.append(int(solver.iterations[0]))
        status.append(solver.read_status())
        if solver.contact_multipliers is not None:
            multipliers.append(solver.contact_multipliers.cpu().numpy().copy())
        if solver.mesh_contact_pair_counts is not None:
            mesh_pairs.append(solver.mesh_contact_pair_counts)
            mesh_force.append(solver.mesh_contact_force(state.particle_q.detach(), state.particle_qd.detach())
                              .cpu().numpy().copy())
        if damping is not None:
            out.particle_qd = damping * out.particle_qd
        state = out
    loss = (state.particle_q.to(torch.float64) * torch.as_tensor(fx["objective_q"], device=dev)).sum() + (
        state.particle_qd.to(torch.float64) * torch.as_tensor(fx["objective_v"], device=dev)).sum()
    loss.backward()
    backward_status = solver.read_status()
    x = np.stack(xs)
    scale = 1.0 if damping is None else damping
    dirichlet = np.asarray(fx["dirichlet"], dtype=bool)
    result = {
        "q_next": x,
        "v_next": np.stack([(x[k] - q_in[k]) / dt * scale for k in range(steps)]),
        "iterations": np.array(iterations),
        "converged": np.array([(bits & 1) == 0 for bits in status]),
        "status_bits": np.array(status + [backward_status]),
        "dL_dq": _grad(q),
        "dL_dv": _grad(v),
        "dL_df_ext": np.stack([_grad(f) for f in forces]),
        "dL_dmaterials_per_element": _grad(model.tet_materials)[:, :2],
        "loss": float(loss.detach().cpu()),
        "free": ~dirichlet,
    }
    if multipliers:
        lam = np.stack(multipliers)
        result["multipliers"] = lam
        result["active"] = lam > 0
        if str(_value(fx["contact_kind"])) == "planar":
            params = np.asarray(fx["contact_parameters"], dtype=np.float64)
            nodes = np.asarray(fx["contact_vertices"])
            result["gaps"] = (x[:, nodes] @ params[:3] + params[3]) / np.linalg.norm(params[:3])
    if mesh_pairs:
        result["mesh_contact_force"] = np.stack(mesh_force)
        result["mesh_contact_pairs"] = np.array(mesh_pairs)
    if acts is not None:
        result["dL_dactivations"] = np.stack([_grad(a) for a in acts])
    return result


def _grad(t):
    g = t.grad
    return np.zeros(tuple(t.shape)) if g is None else g.detach().to("cpu").numpy().astype(np.float64)


# ---------------------------------------------------------------------------
# The harness's named scenes (port_cases.py:240-282) on this package's grids:
# 4 x 2 x 2 cells of 0.125 m (dyadic, so the float32 rest poses are exact),
# mu 2e4, lambda 5e4, density 1000, dt 0.004.

DT = 0.004
MU, LAME = 2.0e4, 5.0e4


def grid(fix_left=False, contrast=1.0, height=0.5, dims=(4, 2, 2), cell=0.125, origin=(0.0, 0.0, 0.0),
         fix_bottom=False):
    """Rest vertices, tets, (mu, lambda) and Dirichlet flags of a Kuhn grid; tets whose centroid lies
    beyond the mean x are ``contrast`` times stiffer."""
    rest, tets, _ = meshes.kuhn_tet_grid(*dims, cell, origin=(origin[0], origin[1], origin[2] + height))
    rest = rest.astype(np.float64)
    mu, lame = np.full(len(tets), MU), np.full(len(tets), LAME)
    stiff = rest[tets].mean(axis=1)[:, 0] > rest[:, 0].mean()
    mu[stiff] *= contrast
    lame[stiff] *= contrast
    dirichlet = np.zeros(len(rest), dtype=bool)
    if fix_left:
        dirichlet |= rest[:, 0] == rest[:, 0].min()
    if fix_bottom:
        dirichlet |= rest[:, 2] == rest[:, 2].min()
    return rest, tets, mu, lame, dirichlet


def _deformed(stretch, speed, force, seed, rest, dirichlet, steps):
    """Initial state stretched about the centroid with noise, random velocity and forces (free particles)."""
    rng = np.random.default_rng(seed)
    free = ~dirichlet
    center = rest.mean(axis=0)
    q = np.where(free[:, None], center + (rest - center) * np.asarray(stretch), rest)
    q[free] += 0.01 * rng.standard_normal((free.sum(), 3))
    v = np.where(free[:, None], speed * rng.standard_normal(q.shape), 0.0)
    f = force * rng.uniform(-1.0, 1.0, (steps, len(rest), 3))
    return q, v, f


def _scene(case, fix_left=False, contrast=1.0, *, steps=1, material="pd", seed=0, deform=None, dt=DT, **extra):
    rest, tets, mu, lame, dirichlet = grid(fix_left, contrast)
    kw = {}
    if deform is not None:
        stretch, speed, force, dseed = deform
        kw["q"], kw["v"], kw["f_ext"] = _deformed(stretch, speed, force, dseed, rest, dirichlet, steps)
    return make_fixture(case, rest, tets, mu, lame, dirichlet, steps=steps, dt=dt, material=material, seed=seed,
                        **kw, **extra)


def _plane(case, friction, velocity, restitution=0.0):
    """The grid on the plane z = 0, lifted 1 mm, moving at ``velocity``; its surface vertices registered."""
    rest, tets, mu, lame, dirichlet = grid(False, height=0.0)
    q = rest + np.array([0.0, 0.0, 0.001])
    v = np.tile(np.asarray(velocity, dtype=np.float64), (len(rest), 1))
    return make_fixture(
        case, rest, tets, mu, lame, dirichlet, steps=3, q=q, v=v, seed=5, contact_kind="planar",
        contact_parameters=np.array([0.0, 0.0, 1.0, 0.0]), contact_vertices=meshes.surface_vertices(tets),
        friction=friction, restitution=restitution,
    )


def boundary_faces(tets):
    """Faces of ``tets`` that belong to exactly one tetrahedron, outward (first occurrence order)."""
    local = np.array(((1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1)))
    faces = np.asarray(tets)[:, local].reshape(-1, 3)
    _, first, counts = np.unique(np.sort(faces, axis=1), axis=0, return_index=True, return_counts=True)
    return faces[np.sort(first[counts == 1])].astype(np.int32)


def _two_blocks():
    """A bottom-pinned 4 x 2 x 2 grid (A) and a free one-cell grid (B) 1/32 m above its top."""
    ra, ta, _, _, da = grid(height=0.0, fix_bottom=True)
    rb, tb, _, _, db = grid(height=0.0, dims=(1, 1, 1), origin=(0.3125, 0.0625, 0.28125))
    rest = np.concatenate([ra, rb])
    tets = np.concatenate([ta, tb + len(ra)])
    dirichlet = np.concatenate([da, db])
    return rest, tets, dirichlet, len(ra)


def _mesh_contact_poke():
    rest, tets, dirichlet, count_a = _two_blocks()
    steps = 3
    rng = np.random.default_rng(10)
    free = ~dirichlet
    body_b = np.arange(len(rest)) >= count_a
    q = rest + np.where(free[:, None], 0.002 * rng.standard_normal(rest.shape), 0.0)
    q[body_b, 2] -= 0.01
    v = np.where(free[:, None], 0.2 * rng.standard_normal(rest.shape), 0.0)
    v[body_b] += np.array([0.5, 0.2, -0.5])
    f = 0.5 * rng.uniform(-1.0, 1.0, (steps, len(rest), 3))
    faces = boundary_faces(tets)
    in_b = faces >= count_a
    return make_fixture(
        "mesh_contact_poke", rest, tets, np.full(len(tets), MU), np.full(len(tets), LAME), dirichlet, steps=steps,
        q=q, v=v, f_ext=f, seed=9, mesh_contact_parameters=np.array([0.08, 60.0, 150.0, 3.0, 0.5]),
        mesh_contact_faces_a=faces[~in_b.any(axis=1)], mesh_contact_faces_b=faces[in_b.all(axis=1)],
    )


def _mesh_boundary_poke():
    rest, tets, dirichlet, count_a = _two_blocks()
    rng = np.random.default_rng(11)
    free = ~dirichlet
    body_b = np.arange(len(rest)) >= count_a
    q = rest + np.where(free[:, None], 0.001 * rng.standard_normal(rest.shape), 0.0)
    q[body_b, 2] -= 0.03
    v = np.where(free[:, None], 0.05 * rng.standard_normal(rest.shape), 0.0)
    v[body_b] += np.array([0.0, 0.0, -0.5])
    faces = boundary_faces(tets)
    in_b = faces >= count_a
    left, right = faces[~in_b.any(axis=1)], faces[in_b.all(axis=1)]
    candidates = np.unique(left)
    parameters = np.concatenate([[0.01, count_a, len(left), len(right)], left.ravel(), right.ravel()])
    return make_fixture(
        "mesh_boundary_poke", rest, tets, np.full(len(tets), MU), np.full(len(tets), LAME), dirichlet, steps=1,
        q=q, v=v, seed=9, contact_kind="mesh", contact_parameters=parameters.astype(np.float64),
        contact_vertices=candidates[free[candidates]].astype(np.int32),
    )


def _mesh_contact_prescribed():
    fixture = dict(_mesh_contact_poke())
    probe = np.unique(fixture["mesh_contact_faces_b"])
    rest = fixture["rest_vertices"]
    root = probe[rest[probe, 2] == rest[probe, 2].max()]
    dirichlet = fixture["dirichlet"].copy()
    dirichlet[root] = True
    steps = int(fixture["steps"])
    targets = np.repeat(fixture["q"][None], steps, axis=0)
    for k in range(steps):
        targets[k, root] += (k + 1) * np.array([0.001, 0.0, -0.002])
    fixture.update(case="mesh_contact_prescribed", material="neo_hookean", dirichlet=dirichlet,
                   dirichlet_positions=targets, velocity_damping=0.30)
    return finalize(fixture)


def cases():
    """The harness's named scenes as fixtures (port_cases.py:240-282 lists them)."""
    out = [_scene("homogeneous_free", deform=((1.2, 0.85, 1.05), 0.3, 1.0, 1))]
    out.append(finalize({**out[0], "case": "zero_volume_weight", "lambda": np.zeros_like(out[0]["lambda"])}))
    for contrast in (10.0, 100.0):
        for encoding in ("per_element", "grouped"):
            fx = _scene(f"pinned_contrast_{contrast:g}x_{encoding}", True, contrast, seed=2,
                        deform=((1.25, 0.9, 1.1), 0.3, 1.0, 3), reference_material_encoding=encoding)
            out.append(fx)
            if contrast == 10.0 and encoding == "per_element":
                out.append(finalize({**fx, "case": "pinned_contrast_10x_per_element_small_dt", "dt": 0.001}))
    out.append(_scene("neo_hookean_large_deformation", material="neo_hookean", seed=4,
                      deform=((1.5, 0.6, 1.25), 0.5, 1.0, 4)))
    out.append(_plane("plane_impact", 0.0, (0.0, 0.0, -1.0)))
    out.append(_plane("sticking_friction", 2.0, (0.05, 0.0, -1.0)))
    out.append(_plane("sliding_friction", 0.1, (2.0, 0.0, -1.0)))
    out.append(_scene("multistep_pinned", True, steps=4, seed=6, deform=((1.2, 0.85, 1.05), 0.3, 1.0, 6)))
    muscle = _scene("muscle_activation", True, steps=2, seed=7, deform=((1.1, 0.95, 1.0), 0.2, 0.5, 7))
    rng = np.random.default_rng(8)
    count = len(muscle["tets"])
    out.append(finalize({**muscle, "muscle_directions": rng.standard_normal((count, 3)),
                         "muscle_stiffness": np.full(count, 4.0e4),
                         "activations": rng.uniform(-0.2, 0.3, (2, count))}))
    out.append(_mesh_contact_poke())
    out.append(_mesh_boundary_poke())
    out.append(_mesh_contact_prescribed())
    return out


def main(argv=None):
    parser = argparse.ArgumentParser(description="Write parity fixtures and this solver's port outputs.")
    parser.add_argument("--out", type=Path, required=True)
    parser.add_argument("--device", default="cuda:0")
    parser.add_argument("--cases", nargs="*", help="named scenes to write and run (default: all)")
    parser.add_argument("--fixtures", nargs="*", type=Path, help="existing fixture files to run instead")
    parser.add_argument("--precision", type=int, default=64, choices=(32, 64), help="factor precision")
    args = parser.parse_args(argv)
    args.out.mkdir(parents=True, exist_ok=True)
    (args.out / "port_environment.json").write_text(json.dumps(environment(args.device), indent=2) + "\n")
    if args.fixtures:
        todo = [(Path(p).name.replace(".fixture.npz", ""), load(p)) for p in args.fixtures]
    else:
        named = cases()
        known = {f["case"] for f in named}
        unknown = set(args.cases or ()) - known
        if unknown:
            parser.error(f"unknown parity cases: {', '.join(sorted(unknown))}")
        todo = [(f["case"], f) for f in named if not args.cases or f["case"] in args.cases]
        for name, fx in todo:
            save(args.out / f"{name}.fixture.npz", fx)
    for name, fx in todo:
        print(f"{name}: port", flush=True)
        save(args.out / f"{name}.port.npz", run(fx, args.device, args.precision))


if __name__ == "__main__":
    main()
