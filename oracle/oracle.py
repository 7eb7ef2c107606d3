"""ctypes front end of the C oracle (TEST ORACLE ONLY).

Test infrastructure: only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg may use it, and only as the checker.

``OracleBody`` describes one body with the reference's per-solver data
(``SolverDiffPhD.__init__`` / ``_build``, ``DP/solver.py:157-552``), built
by the NumPy restatements in ``setup_ref.py``; the factor comes either from
``sparse_ref.py`` (pure Python, small bodies) or from a caller-supplied
function returning the same CSR arrays.
"""

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

from . import setup_ref, sparse_ref

HERE = Path(__file__).resolve().parent
BUILD = HERE / "_build"
LIB = BUILD / "liboracle.so"


def build(force=False):
    """Compile the C oracle with gcc (-ffp-contract=off, OpenMP)."""
    sources = [HERE / "diffphd_oracle.c", HERE / "diffphd_contact.c"]
    newest = max(f.stat().st_mtime for f in [*sources, HERE / "diffphd_oracle.h", HERE / "contact_oracle.h"])
    if LIB.exists() and not force and LIB.stat().st_mtime >= newest:
        return LIB
    BUILD.mkdir(exist_ok=True)
    tmp = BUILD / f"liboracle.so.tmp{os.getpid()}"
    cmd = ["gcc", "-O2", "-std=c99", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", *map(str, sources), "-o",
           str(tmp), "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


class _Contact(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("params", ctypes.c_double * 4),
        ("count", ctypes.c_int),
        ("nodes", ctypes.c_void_p),
        ("friction", ctypes.c_double),
        ("restitution", ctypes.c_double),
        ("multiplier", ctypes.c_void_p),
        ("omega", ctypes.c_void_p),
        ("normal", ctypes.c_void_p),
        ("candidate", ctypes.c_void_p),
        ("radius", ctypes.c_double),
        ("split", ctypes.c_int),
        ("left_count", ctypes.c_int),
        ("face_count", ctypes.c_int),
        ("faces", ctypes.c_void_p),
    ]


class _Problem(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("elements", ctypes.c_int),
        ("corners", ctypes.c_int),
        ("samples", ctypes.c_int),
        ("indices", ctypes.c_void_p),
        ("grads", ctypes.c_void_p),
        ("wc", ctypes.c_void_p),
        ("wv", ctypes.c_void_p),
        ("wn", ctypes.c_void_p),
        ("jwc", ctypes.c_void_p),
        ("jwv", ctypes.c_void_p),
        ("jwn", ctypes.c_void_p),
        ("mdir", ctypes.c_void_p),
        ("mw", ctypes.c_void_p),
        ("material", ctypes.c_int),
        ("nh_mu", ctypes.c_double),
        ("nh_lame", ctypes.c_double),
        ("nh_penalty", ctypes.c_double),
        ("free", ctypes.c_void_p),
        ("mass", ctypes.c_double),
        ("sample_volume", ctypes.c_double),
        ("gravity", ctypes.c_double * 3),
        ("inc_off", ctypes.c_void_p),
        ("inc_ent", ctypes.c_void_p),
        ("s_off", ctypes.c_void_p),
        ("s_col", ctypes.c_void_p),
        ("s_val", ctypes.c_void_p),
        ("t_off", ctypes.c_void_p),
        ("t_col", ctypes.c_void_p),
        ("t_val", ctypes.c_void_p),
        ("window", ctypes.c_int),
        ("max_iterations", ctypes.c_int),
        ("abs_tol", ctypes.c_double),
        ("rel_tol", ctypes.c_double),
        ("integrate_pinned", ctypes.c_int),
        ("threads", ctypes.c_int),
        ("planar_count", ctypes.c_int),
        ("planar_idx", ctypes.c_void_p),
        ("planar_w", ctypes.c_void_p),
        ("plane", ctypes.c_double * 4),
        ("state_force", ctypes.c_void_p),
        ("contact", ctypes.POINTER(_Contact)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ctypes.CDLL(str(LIB))
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


class OracleBody:
    """One body of the reference solver, set up like ``SolverDiffPhD._build``."""

    def __init__(
        self,
        kind,
        indices,
        *,
        poses=None,
        rest=None,
        materials,
        density,
        inverse_mass,
        gravity=(0.0, 0.0, 0.0),
        dt,
        material="pd",
        window=5,
        max_iterations=10000,
        abs_tol=1.0e-9,
        rel_tol=1.0e-6,
        integrate_pinned=False,
        jacobian_mode="reference",
        muscle_directions=None,
        muscle_stiffness=None,
        planar_energies=None,
        mesh_contact=None,
        frictional_boundary=None,
        frictional_vertices=None,
        friction=0.0,
        restitution=0.0,
        factor=None,
        threads=0,
    ):
        self.kind = kind
        self.indices = np.ascontiguousarray(indices, dtype=np.int32)
        self.n = len(inverse_mass)
        self.free = (np.asarray(inverse_mass) > 0).astype(np.uint8)
        materials = np.asarray(materials, dtype=np.float64)[:, :2]
        E = len(self.indices)
        if kind == "tet4":
            grads, volumes = setup_ref.tet_gradients(poses)
            self.mean_volume = float(volumes.mean())
            self.samples, self.corners, shared = 1, 4, False
            self.inc_off, self.inc_ent = setup_ref.tet_incidence(self.indices, self.n)
        else:
            h = setup_ref.hex_setup(rest, self.indices)
            grads, self.mean_volume = h["grads"], h["mean_volume"]
            self.samples, self.corners, shared = 8, 8, True
            self.inc_off, self.inc_ent = h["inc_off"], h["inc_ent"]
        self.grads = np.ascontiguousarray(grads, dtype=np.float64)
        svol = self.mean_volume / self.samples
        self.sample_volume = svol
        self.mass = float(density) * self.mean_volume
        self.wc, self.wv, self.wn = setup_ref.element_weights(materials, material, svol)
        mu_bar, lam_bar = float(materials[:, 0].mean()), float(materials[:, 1].mean())
        self.nh = (mu_bar, lam_bar, 2.0 * mu_bar)
        if jacobian_mode == "reference":
            ones = np.ones(E)
            pd = material == "pd"
            self.jwc = ones * (2.0 * mu_bar * svol if pd else 0.0)
            self.jwv = ones * (lam_bar * svol if pd else 0.0)
            self.jwn = ones * (0.0 if pd else 2.0 * mu_bar * svol)
        else:
            self.jwc, self.jwv, self.jwn = self.wc.copy(), self.wv.copy(), self.wn.copy()
        self.mdir = self.mw = None
        if muscle_directions is not None:
            d = np.array(muscle_directions, dtype=np.float64)
            k = np.broadcast_to(np.asarray(muscle_stiffness, dtype=np.float64), (E,)).copy()
            lengths = np.linalg.norm(d, axis=1)
            active = k > 0
            d[active] /= lengths[active, None]
            d[~active] = 0.0
            self.mdir, self.mw = d, k * svol
        self.material = 0 if material == "pd" else 1
        self.inertia = self.mass / dt**2
        # PlanarEnergies (hex.py:204-243): distinct vertices, one plane with its
        # normal scaled to unit length (the offset with it).
        self.planar = None
        if planar_energies is not None:
            idx, stiffness, plane = planar_energies
            idx = np.ascontiguousarray(idx, dtype=np.int32)
            stiffness = np.ascontiguousarray(np.broadcast_to(np.asarray(stiffness, np.float64), idx.shape))
            plane = np.asarray(plane, dtype=np.float64)
            self.planar = (idx, stiffness, plane / np.linalg.norm(plane[:3]))
        # MeshContact (mesh_contact.py:180-243): evaluated on the start state.
        self.mesh_contact = None
        if mesh_contact is not None:
            radius, kn_left, kn_right, kf, mu, faces_a, faces_b = mesh_contact
            self.mesh_contact = (
                tuple(float(x) for x in (radius, kn_left, kn_right, kf, mu)),
                np.ascontiguousarray(faces_a, dtype=np.int32).reshape(-1, 3),
                np.ascontiguousarray(faces_b, dtype=np.int32).reshape(-1, 3),
            )
        self._state_force = None
        # FrictionalContact (contact.py:686-780); the multipliers persist between steps.
        self.contact = None
        if frictional_boundary is not None:
            kind, params = frictional_boundary
            c = {"kind": {"planar": 0, "spherical": 1, "mesh": 2}[kind], "friction": float(friction),
                 "restitution": float(restitution)}
            if kind == "mesh":
                radius, split, left, right = params
                left = np.asarray(left, dtype=np.int32).reshape(-1, 3)
                right = np.asarray(right, dtype=np.int32).reshape(-1, 3)
                c.update(radius=float(radius), split=int(split), left_count=len(left),
                         faces=np.ascontiguousarray(np.concatenate([left, right])), params=np.zeros(4))
                if frictional_vertices is None:
                    frictional_vertices = np.unique(c["faces"])
            else:
                c.update(radius=0.0, split=0, left_count=0, faces=None, params=np.asarray(params, np.float64))
                if frictional_vertices is None:
                    frictional_vertices = setup_ref.surface_vertices(self.indices)
            nodes = np.ascontiguousarray(frictional_vertices, dtype=np.int32)
            k = len(nodes)
            c.update(nodes=nodes, multiplier=np.zeros(k), omega=np.zeros(k), normal=np.zeros((k, 3)),
                     candidate=np.zeros(k, np.int32))
            self.contact = c
        weights = self.wc + self.wv + self.wn
        self.rows = setup_ref.assemble_rows(
            self.n, self.indices, self.grads, shared, weights, self.free.astype(bool), self.inertia, self.mdir, self.mw,
            None if self.planar is None else self.planar[:2],
        )
        if factor is None:
            f = sparse_ref.sparse_inverse(self.rows)
        else:
            f = factor(setup_ref.rows_to_full_csr(self.rows))
        self.factor = f
        (self.s_off, self.s_col, self.s_val) = [np.ascontiguousarray(a) for a in f["s"]]
        (self.t_off, self.t_col, self.t_val) = [np.ascontiguousarray(a) for a in f["st"]]
        self.gravity = np.asarray(gravity, dtype=np.float64)
        self.options = dict(
            window=window,
            max_iterations=max_iterations,
            abs_tol=abs_tol,
            rel_tol=rel_tol,
            integrate=int(integrate_pinned),
            threads=threads,
        )
        self._keep = None

    def _struct(self):
        o = self.options
        keep = [
            self.indices,
            self.grads,
            np.ascontiguousarray(self.wc),
            np.ascontiguousarray(self.wv),
            np.ascontiguousarray(self.wn),
            np.ascontiguousarray(self.jwc),
            np.ascontiguousarray(self.jwv),
            np.ascontiguousarray(self.jwn),
            _c(self.mdir, np.float64),
            _c(self.mw, np.float64),
            self.free,
            self.inc_off,
            self.inc_ent,
        ]
        self._keep = keep
        p = _Problem()
        p.n, p.elements, p.corners, p.samples = self.n, len(self.indices), self.corners, self.samples
        (
            p.indices,
            p.grads,
            p.wc,
            p.wv,
            p.wn,
            p.jwc,
            p.jwv,
            p.jwn,
            p.mdir,
            p.mw,
            p.free,
            p.inc_off,
            p.inc_ent,
        ) = [_ptr(a) for a in keep]
        p.material = self.material
        p.nh_mu, p.nh_lame, p.nh_penalty = self.nh
        p.mass = self.mass
        p.sample_volume = self.sample_volume
        for d in range(3):
            p.gravity[d] = self.gravity[d]
        p.s_off, p.s_col, p.s_val = _ptr(self.s_off), _ptr(self.s_col), _ptr(self.s_val)
        p.t_off, p.t_col, p.t_val = _ptr(self.t_off), _ptr(self.t_col), _ptr(self.t_val)
        p.window, p.max_iterations = o["window"], o["max_iterations"]
        p.abs_tol, p.rel_tol = o["abs_tol"], o["rel_tol"]
        p.integrate_pinned, p.threads = o["integrate"], o["threads"]
        p.state_force = _ptr(self._state_force)
        if self.contact is not None:
            c = self.contact
            self._contact_struct = _Contact()
            cs = self._contact_struct
            cs.kind, cs.count = c["kind"], len(c["nodes"])
            for d in range(4):
                cs.params[d] = c["params"][d]
            cs.nodes, cs.multiplier = _ptr(c["nodes"]), _ptr(c["multiplier"])
            cs.omega, cs.normal, cs.candidate = _ptr(c["omega"]), _ptr(c["normal"]), _ptr(c["candidate"])
            cs.friction, cs.restitution = c["friction"], c["restitution"]
            cs.radius, cs.split, cs.left_count = c["radius"], c["split"], c["left_count"]
            cs.face_count = 0 if c["faces"] is None else len(c["faces"])
            cs.faces = _ptr(c["faces"])
            p.contact = ctypes.pointer(cs)
        if self.planar is not None:
            p.planar_count = len(self.planar[0])
            p.planar_idx, p.planar_w = _ptr(self.planar[0]), _ptr(self.planar[1])
            for d in range(4):
                p.plane[d] = self.planar[2][d]
        return p

    def forward(self, dt, q, qd, f=None, dirichlet=None, targets=None):
        n = self.n
        q = _c(q, np.float32)
        qd = _c(qd, np.float32)
        self._state_force = None
        if self.mesh_contact is not None:
            params, faces_a, faces_b = self.mesh_contact
            self._state_force = mesh_contact_force(q, qd, faces_a, faces_b, *params)[0]
        f = np.zeros((n, 3), np.float32) if f is None else _c(f, np.float32)
        dirichlet = _c(dirichlet, np.float32)
        targets = _c(targets, np.float64)
        x = np.zeros((n, 3))
        qo = np.zeros((n, 3), np.float32)
        qdo = np.zeros((n, 3), np.float32)
        it, st = ctypes.c_int(), ctypes.c_int()
        p = self._struct()
        lib().oracle_forward(
            ctypes.byref(p),
            ctypes.c_double(dt),
            _ptr(q),
            _ptr(qd),
            _ptr(f),
            _ptr(dirichlet),
            _ptr(targets),
            _ptr(x),
            _ptr(qo),
            _ptr(qdo),
            ctypes.byref(it),
            ctypes.byref(st),
        )
        return {"x": x, "q": qo, "qd": qdo, "iterations": it.value, "status": st.value}

    def backward(
        self, dt, x, gq_out=None, gqd_out=None, materials=True, explicit_dirichlet=False, targets=None, activations=False
    ):
        n, E = self.n, len(self.indices)
        gq_in = np.zeros((n, 3), np.float32)
        gqd_in = np.zeros((n, 3), np.float32)
        gf_in = np.zeros((n, 3), np.float32)
        gmat = np.zeros((E, 2)) if materials else None
        gdir = np.zeros((n, 3), np.float32) if explicit_dirichlet else None
        gact = np.zeros(E) if activations else None
        mu = np.zeros((n, 3))
        it, st = ctypes.c_int(), ctypes.c_int()
        p = self._struct()
        code = lib().oracle_backward(
            ctypes.byref(p),
            ctypes.c_double(dt),
            _ptr(_c(x, np.float64)),
            _ptr(_c(targets, np.float64)),
            _ptr(_c(gq_out, np.float32)),
            _ptr(_c(gqd_out, np.float32)),
            _ptr(gq_in),
            _ptr(gqd_in),
            _ptr(gf_in),
            _ptr(gmat),
            _ptr(gdir),
            _ptr(gact),
            _ptr(mu),
            ctypes.byref(it),
            ctypes.byref(st),
        )
        if code != 0:
            raise NotImplementedError("the oracle has no adjoint with planar energies")
        return {
            "q": gq_in,
            "qd": gqd_in,
            "f": gf_in,
            "materials": gmat,
            "dirichlet": gdir,
            "activations": gact,
            "mu": mu,
            "iterations": it.value,
            "status": st.value,
        }

    def apply_inverse(self, rhs):
        rhs = _c(rhs, np.float64)
        out = np.zeros_like(rhs)
        p = self._struct()
        lib().oracle_apply_inverse(ctypes.byref(p), _ptr(rhs), _ptr(out))
        return out


def safe_clip(values, free, limit):
    """The backward's safe clip of one [n,3] vector with norm limit ``limit``; returns (values, status)."""
    v = np.array(values, dtype=np.float64, order="C").reshape(-1, 3)
    free = np.ascontiguousarray(np.asarray(free).astype(np.uint8))
    status = lib().oracle_safe_clip(len(v), _ptr(free), _ptr(v), ctypes.c_double(limit))
    return v, status


def routing_guard(rhs_sq, mu_sq, inertia, clip_norm):
    """The backward's routing guard from squared norms; returns (scale, diverged, status)."""
    scale, diverged = ctypes.c_double(), ctypes.c_int()
    status = lib().oracle_routing_guard(
        ctypes.c_double(rhs_sq), ctypes.c_double(mu_sq), ctypes.c_double(inertia), ctypes.c_double(clip_norm),
        ctypes.byref(scale), ctypes.byref(diverged),
    )
    return scale.value, diverged.value, status


def projections(f, directions, mu, lame, penalty, target):
    f = _c(f, np.float64)
    d = _c(directions, np.float64)
    c = len(f)
    out = [np.zeros((c, 3)), np.zeros((c, 3, 3)), np.zeros((c, 3, 3)), np.zeros((c, 3, 3)), np.zeros((c, 3, 3)), np.zeros((c, 3))]
    lib().oracle_projections(
        c, _ptr(f), _ptr(d), ctypes.c_double(mu), ctypes.c_double(lame), ctypes.c_double(penalty),
        ctypes.c_double(target), *[_ptr(a) for a in out]
    )
    return dict(zip(("sigma", "reconstruction", "rotation", "volume", "neo_hookean", "fibre"), out))


def jvps(f, df, directions, mu, lame, target):
    f, df, d = _c(f, np.float64), _c(df, np.float64), _c(directions, np.float64)
    c = len(f)
    out = [np.zeros((c, 3, 3)), np.zeros((c, 3, 3)), np.zeros((c, 3, 3)), np.zeros((c, 3))]
    lib().oracle_jvps(
        c, _ptr(f), _ptr(df), _ptr(d), ctypes.c_double(mu), ctypes.c_double(lame), ctypes.c_double(target),
        *[_ptr(a) for a in out]
    )
    return dict(zip(("rotation", "volume", "neo_hookean", "fibre"), out))


def mesh_contact_force(q, v, faces_a, faces_b, radius, kn_left, kn_right, kf, mu):
    """Mesh contact state force [n,3] and the pair count (mesh_contact.py:90-170)."""
    q = _c(np.asarray(q, dtype=np.float64), np.float64)
    v = _c(np.asarray(v, dtype=np.float64), np.float64)
    faces_a = np.ascontiguousarray(faces_a, dtype=np.int32).reshape(-1, 3)
    faces_b = np.ascontiguousarray(faces_b, dtype=np.int32).reshape(-1, 3)
    out = np.zeros_like(q)
    pairs = lib().oracle_mesh_contact_force(
        len(q), _ptr(q), _ptr(v), len(faces_a), _ptr(faces_a), len(faces_b), _ptr(faces_b), ctypes.c_double(radius),
        ctypes.c_double(kn_left), ctypes.c_double(kn_right), ctypes.c_double(kf), ctypes.c_double(mu), _ptr(out),
    )
    return out, pairs
