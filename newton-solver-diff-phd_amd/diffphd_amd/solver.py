"""``SolverDiffPhD``: the reference plugin surface over the native gfx950 path.

Mirrors ``SolverDiffPhD`` of ``src/newton_solver_diff_phd/_src/solvers/diff_phd/solver.py``:
constructor options (:157-179), ``step(state_in, state_out, control,
contacts, dt, *, observables=None, dirichlet_positions=None)`` (:689-786),
``notify_model_changed`` (:419-427), ``read_status`` / ``check_status``
(:380-417), ``iterations`` (:355-358) and ``deformation_determinants``.
Gradients: where the reference records a ``wp.Tape`` closure
(:801-827), this solver's steps are ``torch.autograd`` nodes whose backward
runs the native adjoint (``autograd.StepFunction``); gradients reach
``state_in.particle_q/qd/f``, ``model.tet_materials[:, :2]``,
``dirichlet_positions`` and the fibre activations (``control.tet_activations``
or ``model.tet_activations``, whichever the step read).

Additions, not in the reference: ``independent_bodies`` solves every
builder body as its own system with its own factor, convergence gate and
Anderson history, all in one launch per kernel; ``factor_precision=32``
stores S in FP32 and corrects the defect in FP64; ``jacobian_mode="element"``
uses per-element weights in the adjoint Jacobian (the exact derivative for
heterogeneous materials; the reference uses family means).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr
from .model import Mesh

MATERIALS = {"pd": 0, "neo_hookean": 1}
STATUS_NOT_CONVERGED = 1
STATUS_SINGULAR = 2
STATUS_ACTIVATION = 4
STATUS_ADJOINT_NOT_CONVERGED = 8
STATUS_ADJOINT_SAFEGUARD = 16


def _surface_vertices(tets):
    """Vertices of faces of exactly one tetrahedron (contact.py:679-684), the default frictional vertices."""
    faces = np.concatenate([tets[:, [0, 1, 2]], tets[:, [0, 1, 3]], tets[:, [0, 2, 3]], tets[:, [1, 2, 3]]])
    unique, counts = np.unique(np.sort(faces, axis=1), axis=0, return_counts=True)
    return np.unique(unique[counts == 1]).astype(np.int32)


class ConvergenceError(RuntimeError):
    """A DiffPhD fixed point did not converge; a smaller interval may restore convergence."""


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class SolverDiffPhD:
    ConvergenceError = ConvergenceError

    def __init__(
        self,
        model,
        *,
        density,
        material: str = "pd",
        max_iterations: int = 10000,
        abs_tol: float = 1.0e-9,
        rel_tol: float = 1.0e-6,
        anderson_window: int = 5,
        muscle_directions=None,
        muscle_stiffness=None,
        integrate_pinned_velocities: bool = False,
        frictional_boundary=None,
        frictional_vertices=None,
        friction: float = 0.0,
        restitution: float = 0.0,
        mesh_contact=None,
        hex_mesh: Mesh | None = None,
        hex_materials=None,
        muscle_targets=None,
        planar_energies=None,
        independent_bodies: bool = False,
        factor_precision: int = 64,
        jacobian_mode: str = "reference",
        iteration_chunk: int = 0,
        slice_rows: int = 0,
    ):
        if material not in MATERIALS:
            raise ValueError("material must be 'pd' or 'neo_hookean'")
        if model.particle_count == 0 or (model.tet_count == 0 and hex_mesh is None):
            raise ValueError("SolverDiffPhD requires tetrahedral soft bodies or an explicit hex_mesh")
        if model.body_count or getattr(model, "hex_count", 0):
            raise ValueError("SolverDiffPhD requires particles without rigid bodies; supply Hex8 using hex_mesh")
        if (hex_mesh is None) != (hex_materials is None):
            raise ValueError("hex_mesh and hex_materials must be supplied together")
        if hex_mesh is not None and model.tet_count:
            raise ValueError("hex_mesh cannot be combined with model tetrahedra")
        if hex_mesh is None and (muscle_targets is not None or planar_energies is not None):
            raise ValueError("Explicit muscle_targets and planar_energies currently require hex_mesh")
        if hex_mesh is not None and any(x is not None for x in (frictional_boundary, mesh_contact)):
            raise NotImplementedError("Hex8 currently supports planar_energies, not frictional or mesh contact")
        if frictional_boundary is None and (frictional_vertices is not None or friction or restitution):
            raise ValueError("frictional_vertices, friction and restitution need a frictional_boundary")
        if mesh_contact is not None and len(mesh_contact) != 7:
            raise ValueError("mesh_contact is (contact_radius, kn_left, kn_right, kf, mu, faces_a, faces_b)")
        density = np.broadcast_to(np.asarray(density, dtype=np.float64), (1,) if np.ndim(density) == 0 else np.shape(density))
        if not np.all(np.isfinite(density)) or np.any(density <= 0):
            raise ValueError("density must be finite and positive")
        for name, value in (("abs_tol", abs_tol), ("rel_tol", rel_tol)):
            if not np.isfinite(value) or value < 0:
                raise ValueError(f"{name} must be finite and nonnegative")
        if not np.isfinite(max_iterations) or max_iterations < 1 or int(max_iterations) != max_iterations:
            raise ValueError("max_iterations must be a positive integer")
        if int(anderson_window) != anderson_window or not 0 <= anderson_window <= 8:
            raise ValueError("anderson_window must be an integer in [0, 8]")
        if factor_precision not in (32, 64):
            raise ValueError("factor_precision must be 32 or 64")
        if jacobian_mode not in ("reference", "element"):
            raise ValueError("jacobian_mode must be 'reference' or 'element'")
        if model.device.type != "cuda":
            raise RuntimeError("SolverDiffPhD runs on a HIP device; the model must live on cuda:N")
        self.model = model
        self.device = model.device
        self.material = material
        self._max_iterations = int(max_iterations)
        self._abs_tol = float(abs_tol)
        self._rel_tol = float(rel_tol)
        self.integrate_pinned_velocities = bool(integrate_pinned_velocities)
        self._n = model.particle_count
        self._hex = hex_mesh is not None
        if self._hex:
            cells = np.ascontiguousarray(np.asarray(hex_mesh.indices), dtype=np.int32)
            if cells.ndim != 2 or cells.shape[1] != 8 or len(cells) == 0:
                raise ValueError("hex_mesh indices must have shape (element_count,8) with eight distinct valid vertices")
            rest = np.ascontiguousarray(hex_mesh.vertices, dtype=np.float64)
            if rest.shape != (self._n, 3) or not np.isfinite(rest).all():
                raise ValueError("hex_mesh vertices must contain a finite rest position for every model particle")
            mats = np.asarray(hex_materials, dtype=np.float64)
            if mats.shape not in ((len(cells), 2), (len(cells), 3)):
                raise ValueError("hex_materials must have shape (element_count,2 or 3): mu,lambda[,damping]")
            if mats.shape[1] == 3 and np.any(mats[:, 2] != 0):
                raise ValueError("The reference DiffPhD solver has no damping; element damping must be zero")
            self.hex_materials = torch.tensor(mats[:, :2], dtype=torch.float64, device=self.device)
            self._cells, self._rest = cells, rest
            self._element_count = len(cells)
            voff, eoff = np.array([0, self._n], np.int32), np.array([0, len(cells)], np.int32)
        else:
            self._element_count = model.tet_count
            if independent_bodies:
                voff = np.asarray(model.body_vertex_offsets, np.int32)
                eoff = np.asarray(model.body_element_offsets, np.int32)
            else:
                voff = np.array([0, self._n], np.int32)
                eoff = np.array([0, model.tet_count], np.int32)
        self._voff, self._eoff = np.ascontiguousarray(voff), np.ascontiguousarray(eoff)
        self.body_count_solved = len(voff) - 1
        if len(density) not in (1, self.body_count_solved):
            raise ValueError("density must be a scalar or one value per body")
        self.density = np.ascontiguousarray(np.broadcast_to(density, (self.body_count_solved,)), dtype=np.float64)
        self._has_muscles = False
        self._muscles = None
        if muscle_directions is not None or muscle_stiffness is not None:
            if muscle_directions is None or muscle_stiffness is None:
                raise ValueError("muscle_directions and muscle_stiffness must be supplied together")
            d = np.array(muscle_directions, dtype=np.float64)
            k = np.array(muscle_stiffness, dtype=np.float64)
            if k.ndim == 0:
                k = np.full(self._element_count, float(k))
            if d.shape != (self._element_count, 3) or k.shape != (self._element_count,):
                raise ValueError("Muscle directions must have shape (element_count,3), stiffness scalar or (element_count,)")
            if not np.isfinite(d).all() or not np.isfinite(k).all() or np.any(k < 0):
                raise ValueError("Muscle directions and stiffness must be finite, with nonnegative stiffness")
            self._has_muscles = bool(np.any(k > 0))
            self._muscles = (np.ascontiguousarray(d), np.ascontiguousarray(k))
        # Hex8 fibre target lengths (DP/solver.py:245-260): fixed per element,
        # ones by default; tet_activations cannot control Hex8 fibres.
        self.muscle_targets = None
        if self._hex:
            if muscle_targets is not None and not self._has_muscles:
                raise ValueError("muscle_targets requires enabled muscle_directions and muscle_stiffness")
            targets = np.ones(self._element_count) if muscle_targets is None else muscle_targets
            if isinstance(targets, torch.Tensor):
                if targets.dtype != torch.float64 or targets.device != self.device:
                    raise ValueError("muscle_targets tensor must be float64 on the model device")
                self.muscle_targets = targets
            else:
                self.muscle_targets = torch.as_tensor(np.asarray(targets, dtype=np.float64), device=self.device)
            if tuple(self.muscle_targets.shape) != (self._element_count,):
                raise ValueError("muscle_targets must have shape (element_count,)")
            values = self.muscle_targets.detach().cpu().numpy()
            if not np.isfinite(values).all() or np.any(values < 0):
                raise ValueError("muscle_targets must be finite nonnegative native target lengths")
        # Planar vertex energies (DP/hex.py:204-243), validated like the reference.
        self._planar = None
        if planar_energies is not None:
            p_idx, p_stiff, plane = planar_energies
            p_idx = np.asarray(p_idx)
            if (
                p_idx.ndim != 1
                or not np.issubdtype(p_idx.dtype, np.integer)
                or np.any(p_idx < 0)
                or np.any(p_idx >= self._n)
                or len(np.unique(p_idx)) != len(p_idx)
            ):
                raise ValueError("planar_energies needs distinct valid particle indices")
            p_stiff = np.broadcast_to(np.asarray(p_stiff, dtype=np.float64), p_idx.shape)
            plane = np.asarray(plane, dtype=np.float64)
            normal_length = np.linalg.norm(plane[:3]) if plane.shape == (4,) else 0.0
            if plane.shape != (4,) or not np.isfinite(plane).all() or not np.isfinite(normal_length) or normal_length <= 1.0e-5:
                raise ValueError("planar_energies needs a finite (nx,ny,nz,offset) with nonzero normal")
            if not np.isfinite(p_stiff).all() or np.any(p_stiff < 0):
                raise ValueError("planar_energies stiffness must be finite and nonnegative")
            self._planar = (
                np.ascontiguousarray(p_idx, dtype=np.int32),
                np.ascontiguousarray(p_stiff, dtype=np.float64),
                np.ascontiguousarray(plane),
            )
        cfg = _lib.Config(
            element_kind=1 if self._hex else 0,
            material=MATERIALS[material],
            anderson_window=int(anderson_window),
            max_iterations=self._max_iterations,
            abs_tol=self._abs_tol,
            rel_tol=self._rel_tol,
            factor_precision=int(factor_precision),
            integrate_pinned_velocities=int(self.integrate_pinned_velocities),
            jacobian_mode=0 if jacobian_mode == "reference" else 1,
            iteration_chunk=int(iteration_chunk),
            slice_rows=int(slice_rows),
        )
        handle = ctypes.c_void_p()
        check(lib().dphd_create(ctypes.byref(cfg), self.device.index or 0, ctypes.byref(handle)), "create")
        self._h = handle
        check(lib().dphd_set_bodies(self._h, len(voff) - 1, ptr(self._voff), ptr(self._eoff)), "bodies")
        if self._hex:
            check(lib().dphd_set_hex8(self._h, ptr(self._cells), ptr(self._rest)), "hex8")
        else:
            indices = np.ascontiguousarray(model.tet_indices.detach().cpu().numpy().reshape(-1, 4), dtype=np.int32)
            poses = np.ascontiguousarray(model.tet_poses.detach().cpu().numpy(), dtype=np.float32)
            check(lib().dphd_set_tet4(self._h, ptr(indices), ptr(poses)), "tet4")
        # Mesh contact state force (DP/solver.py:321-325), validated by the library.
        self._mesh_contact = None
        if mesh_contact is not None:
            radius, kn_left, kn_right, kf, mu, faces_a, faces_b = mesh_contact
            fa = np.ascontiguousarray(np.asarray(faces_a).reshape(-1, 3), dtype=np.int32)
            fb = np.ascontiguousarray(np.asarray(faces_b).reshape(-1, 3), dtype=np.int32)
            check(
                lib().dphd_set_mesh_contact(
                    self._h, float(radius), float(kn_left), float(kn_right), float(kf), float(mu), len(fa), ptr(fa),
                    len(fb), ptr(fb),
                ),
                "mesh_contact",
            )
            self._mesh_contact = (fa, fb)
        # Frictional boundary (DP/solver.py:302-319, contact.py:686-780), validated by the library.
        self._friction = None
        if frictional_boundary is not None:
            kind, params = frictional_boundary
            kinds = {"planar": 0, "spherical": 1, "mesh": 2}
            if kind not in kinds:
                raise ValueError("frictional boundary must be 'planar', 'spherical' or 'mesh'")
            faces = np.zeros((0, 3), np.int32)
            split = left = 0
            if kind == "mesh":
                if len(params) != 4:
                    raise ValueError("mesh boundary parameters are (contact_radius, split_index, faces_left, faces_right)")
                radius, split = float(params[0]), int(params[1])
                lf, rf = (np.asarray(f, dtype=np.int32).reshape(-1, 3) for f in params[2:4])
                faces = np.ascontiguousarray(np.concatenate([lf, rf]))
                left = len(lf)
                values = np.array([radius, 0.0, 0.0, 0.0])
                if frictional_vertices is None:
                    frictional_vertices = np.unique(faces)
            else:
                values = np.asarray(params, dtype=np.float64)
                if values.shape != (4,):
                    raise ValueError("frictional boundary parameters must be four finite values")
                if frictional_vertices is None:
                    tets = model.tet_indices.detach().cpu().numpy().reshape(-1, 4)
                    frictional_vertices = _surface_vertices(tets)
            nodes = np.asarray(frictional_vertices)
            if nodes.ndim != 1 or len(np.unique(nodes)) != len(nodes):
                raise ValueError("frictional vertices must be unique particle indices")
            nodes = np.ascontiguousarray(nodes, dtype=np.int32)
            values = np.ascontiguousarray(values, dtype=np.float64)
            check(
                lib().dphd_set_frictional_boundary(
                    self._h, kinds[kind], ptr(values), len(nodes), ptr(nodes), float(friction), float(restitution),
                    split, left, len(faces), ptr(faces),
                ),
                "frictional_boundary",
            )
            self._friction = {"kind": kind, "nodes": nodes, "friction": float(friction)}
        # The converged iterate of the last step in FP64 (the reference's solver._x).
        self._x = torch.zeros((self._n, 3), dtype=torch.float64, device=self.device)
        self._status_word = 0
        self._factor_key = None
        self._model_changed = True
        self._version = 0
        self._targets = None
        self.last_setup_seconds = 0.0
        # When a list, every native forward/backward appends its per-body iteration counts.
        self.iteration_log = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.dphd_destroy(h)
            self._h = None

    # -- configuration mutable between steps (DP/solver.py:206-209) ---------
    @property
    def max_iterations(self):
        return self._max_iterations

    @max_iterations.setter
    def max_iterations(self, value):
        self._max_iterations = int(value)
        self._push_tolerances()

    @property
    def abs_tol(self):
        return self._abs_tol

    @abs_tol.setter
    def abs_tol(self, value):
        self._abs_tol = float(value)
        self._push_tolerances()

    @property
    def rel_tol(self):
        return self._rel_tol

    @rel_tol.setter
    def rel_tol(self, value):
        self._rel_tol = float(value)
        self._push_tolerances()

    def _push_tolerances(self):
        check(lib().dphd_set_tolerances(self._h, self._max_iterations, self._abs_tol, self._rel_tol), "tolerances")

    def notify_model_changed(self, flags: int = 0) -> None:
        """Rebuild weights and the inverse factor before the next step."""
        self._model_changed = True

    # -- build (DP/solver.py:491-552) ------------------------------------------
    def _materials_host(self):
        if self._hex:
            return np.ascontiguousarray(self.hex_materials.detach().cpu().numpy(), dtype=np.float64)
        mats = self.model.tet_materials.detach().cpu().numpy().astype(np.float64)
        if mats.shape[1] == 3 and np.any(mats[:, 2] != 0):
            raise ValueError("The reference DiffPhD solver has no damping; element damping must be zero")
        if not np.all(np.isfinite(mats)) or np.any(mats[:, 0] <= 0) or np.any(mats[:, 1] < 0):
            raise ValueError("DiffPhD requires positive shear modulus and nonnegative Lame lambda")
        return np.ascontiguousarray(mats[:, :2])

    def _build(self, dt):
        model = self.model
        inv_mass = np.ascontiguousarray(model.particle_inv_mass.detach().cpu().numpy(), dtype=np.float32)
        worlds = model.particle_world.detach().cpu().numpy()
        gravity_table = model.gravity.detach().cpu().numpy().astype(np.float64).reshape(-1, 3)
        if np.any(worlds < -1) or np.any(worlds >= len(gravity_table)):
            raise ValueError("particle_world must reference an existing world or global world -1")
        gravity = np.zeros((self.body_count_solved, 3))
        for b in range(self.body_count_solved):
            w = worlds[self._voff[b] : self._voff[b + 1]]
            if np.any(w != w[0]):
                raise ValueError("every vertex of a body must belong to the same world")
            gravity[b] = gravity_table[max(int(w[0]), 0)]
        # Every array handed to C is bound to a name for the duration of the call.
        gravity = np.ascontiguousarray(gravity)
        check(lib().dphd_set_particles(self._h, ptr(inv_mass), ptr(gravity), ptr(self.density)), "particles")
        materials = self._materials_host()
        check(lib().dphd_set_materials(self._h, ptr(materials)), "materials")
        if self._muscles is not None:
            check(lib().dphd_set_muscles(self._h, ptr(self._muscles[0]), ptr(self._muscles[1])), "muscles")
        if self._planar is not None:
            idx, stiff, plane = self._planar
            check(lib().dphd_set_planar(self._h, len(idx), ptr(idx), ptr(stiff), ptr(plane)), "planar")
        check(lib().dphd_build(self._h, float(dt)), "build")
        info = _lib.Info()
        check(lib().dphd_get_info(self._h, ctypes.byref(info)))
        self.info = info
        self.last_setup_seconds = info.setup_seconds
        self._factor_key = float(dt)
        self._model_changed = False
        self._version += 1
        self._pinned_any = bool(np.any(inv_mass == 0))

    def _muscle_targets(self, control):
        if not self._has_muscles:
            return None
        if self._hex:
            if control is not None and control.tet_activations is not None and control.tet_activations.numel():
                raise ValueError("Hex8 uses solver.muscle_targets; tet_activations cannot control Hex8 fibres")
            t = self.muscle_targets
            if t.dtype != torch.float64 or tuple(t.shape) != (self._element_count,):
                raise ValueError("muscle_targets must have shape (element_count,)")
            if bool((~torch.isfinite(t) | (t < 0)).any()):
                self._status_word |= STATUS_ACTIVATION
            return t.detach().contiguous(), None
        source = self.model.tet_activations
        if control is not None and control.tet_activations is not None:
            source = control.tet_activations
        if source is None or tuple(source.shape) != (self._element_count,) or source.dtype != torch.float32:
            raise ValueError("Muscle activations must be a float32 (tet_count,) array")
        u = source.detach().to(torch.float64)
        weight = torch.as_tensor(self._muscles[1], device=self.device)
        bad = ~torch.isfinite(u) | (u > 1.0) | ((weight == 0) & (u != 0))
        if bool(bad.any()):
            self._status_word |= STATUS_ACTIVATION
        return (1.0 - u).contiguous(), source

    # -- stepping ----------------------------------------------------------------
    def _check_state(self, state_in, state_out, dirichlet_positions):
        for s in (state_in, state_out):
            for name in ("particle_q", "particle_qd"):
                t = getattr(s, name)
                if t.device != self.device or tuple(t.shape) != (self._n, 3) or t.dtype != torch.float32:
                    raise ValueError(f"{name} must be float32 ({self._n}, 3) on {self.device}")
        if state_in is state_out or state_in.particle_q.data_ptr() == state_out.particle_q.data_ptr():
            raise ValueError("DiffPhD requires distinct input and output states")
        if dirichlet_positions is not None:
            if not isinstance(dirichlet_positions, torch.Tensor):
                raise TypeError("dirichlet_positions must be a torch tensor")
            if dirichlet_positions.dtype != torch.float32 or tuple(dirichlet_positions.shape) != (self._n, 3):
                raise ValueError("dirichlet_positions must have dtype float32 and shape (particle_count, 3)")
            if dirichlet_positions.device != self.device:
                raise ValueError("dirichlet_positions must be on the solver device")
            if dirichlet_positions.data_ptr() in (state_out.particle_q.data_ptr(), state_out.particle_qd.data_ptr()):
                raise ValueError("dirichlet_positions must not alias an output state")

    def step(self, state_in, state_out, control, contacts, dt, *, observables=None, dirichlet_positions=None):
        """Advance one step of length ``dt`` [s] (DP/solver.py:689-786)."""
        if contacts is not None:
            raise NotImplementedError("SolverDiffPhD does not consume Newton contacts; pass contacts=None")
        if not np.isfinite(dt) or dt <= 0:
            raise ValueError("dt must be finite and positive")
        self._check_state(state_in, state_out, dirichlet_positions)
        dt = float(dt)
        if self._model_changed or self._factor_key != dt:
            self._build(dt)
        targets = self._muscle_targets(control)
        target_values = None if targets is None else targets[0]
        activations = None if targets is None else targets[1]
        f = state_in.particle_f if state_in.particle_f is not None else torch.zeros_like(state_in.particle_q)
        materials = self.hex_materials if self._hex else self.model.tet_materials
        tracked = torch.is_grad_enabled() and any(
            t is not None and t.requires_grad
            for t in (state_in.particle_q, state_in.particle_qd, f, materials, dirichlet_positions, activations)
        )
        if tracked:
            if self._planar is not None:
                raise NotImplementedError("planar_energies are forward only: the reference records no Hex8 adjoint")
            from .autograd import StepFunction

            q_out, qd_out = StepFunction.apply(
                self, dt, state_in.particle_q, state_in.particle_qd, f, materials, dirichlet_positions, target_values,
                activations,
            )
            state_out.particle_q = q_out
            state_out.particle_qd = qd_out
            return
        with torch.no_grad():
            self._native_step(
                dt,
                state_in.particle_q.detach().contiguous(),
                state_in.particle_qd.detach().contiguous(),
                f.detach().contiguous(),
                None if dirichlet_positions is None else dirichlet_positions.detach().contiguous(),
                target_values,
                state_out.particle_q,
                state_out.particle_qd,
                self._x,
            )

    def prepare(self, dt):
        """Build weights and the inverse factor for ``dt`` now instead of at the first step."""
        self._build(float(dt))

    def body_factor_nnz(self) -> np.ndarray:
        """nnz(S) of every body (S^T has the same)."""
        out = np.zeros(self.body_count_solved, np.int64)
        check(lib().dphd_body_nnz(self._h, ptr(out)), "body_nnz")
        return out

    def loop_stats(self) -> dict:
        """Iteration chunks launched and host waits of the forward, adjoint and Woodbury loops since the last
        factor build."""
        out = np.zeros(6, np.int64)
        check(lib().dphd_loop_stats(self._h, ptr(out)), "loop_stats")
        return {mode: {"chunks": int(out[2 * k]), "host_waits": int(out[2 * k + 1])}
                for k, mode in enumerate(("forward", "adjoint", "woodbury"))}

    def solution(self) -> torch.Tensor:
        """FP64 converged iterate of the last step (the reference's ``solver._x``)."""
        return self._x

    def _native_step(self, dt, q, qd, f, dirichlet, targets, q_out, qd_out, checkpoint):
        if self._factor_key != dt or self._model_changed:
            self._build(dt)
        check(
            lib().dphd_step(
                self._h, ptr(q), ptr(qd), ptr(f), ptr(dirichlet), ptr(targets), ptr(q_out), ptr(qd_out),
                ptr(checkpoint), _stream(),
            ),
            "step",
        )
        if self.iteration_log is not None:
            self.iteration_log.append(("forward", self.iterations.numpy()))

    def _native_backward(self, checkpoint, targets, gq_out, gqd_out, gq, gqd, gf, gmat, gdir, gact=None, contact=None):
        if contact is not None:
            mult, omega, normal = contact[:3]
            check(lib().dphd_contact_record(self._h, ptr(mult), ptr(omega), ptr(normal)), "contact_record")
        check(
            lib().dphd_backward(
                self._h, ptr(checkpoint), ptr(targets), ptr(gq_out), ptr(gqd_out), ptr(gq), ptr(gqd), ptr(gf),
                ptr(gmat), ptr(gdir), ptr(gact), _stream(),
            ),
            "backward",
        )
        if self.iteration_log is not None:
            self.iteration_log.append(("adjoint", self.adjoint_iterations.numpy()))

    # -- diagnostics ---------------------------------------------------------------
    @property
    def iterations(self) -> torch.Tensor:
        """Local/global iteration count of the last forward solve, one entry per body."""
        out = np.zeros(self.body_count_solved, np.int32)
        check(lib().dphd_iterations(self._h, 0, ptr(out)))
        return torch.as_tensor(out)

    @property
    def adjoint_iterations(self) -> torch.Tensor:
        out = np.zeros(self.body_count_solved, np.int32)
        check(lib().dphd_iterations(self._h, 1, ptr(out)))
        return torch.as_tensor(out)

    def read_status(self) -> int:
        """Read and clear the status bits (one host synchronization)."""
        bits = ctypes.c_int()
        if self._factor_key is not None:
            check(lib().dphd_read_status(self._h, 1, ctypes.byref(bits)))
        value = int(bits.value) | self._status_word
        self._status_word = 0
        return value

    def check_status(self) -> None:
        bits = self.read_status()
        if bits & STATUS_SINGULAR:
            raise ValueError("DiffPhD volume projection met a singular deformation gradient (reference 'Singular F.')")
        if bits & STATUS_ACTIVATION:
            if self._hex:
                raise ValueError("Hex8 muscle target lengths must be finite and nonnegative")
            raise ValueError("Muscle activations must be finite and <= 1, and zero for disabled fibres")
        if bits & STATUS_NOT_CONVERGED:
            raise ConvergenceError(f"DiffPhD did not converge within {self._max_iterations} iterations")
        if bits & STATUS_ADJOINT_NOT_CONVERGED:
            raise ConvergenceError("A DiffPhD backward fixed point did not converge")
        if bits & STATUS_ADJOINT_SAFEGUARD:
            raise ConvergenceError("A reference DiffPhD backward safeguard modified the gradient")

    @property
    def contact_multipliers(self):
        """Normal multipliers of the registered vertices after the last step (a copy), or None."""
        if self._friction is None or self._factor_key is None:
            return None
        out = torch.zeros(len(self._friction["nodes"]), dtype=torch.float64, device=self.device)
        check(lib().dphd_contact_state(self._h, None, ptr(out), None, None, None))
        return out

    @contact_multipliers.setter
    def contact_multipliers(self, values):
        """Replace the multipliers' warm start."""
        values = torch.as_tensor(values, dtype=torch.float64, device=self.device).contiguous()
        if self._friction is None or tuple(values.shape) != (len(self._friction["nodes"]),):
            raise ValueError("contact_multipliers needs one value per registered vertex")
        check(lib().dphd_contact_state(self._h, ptr(values), None, None, None, None))

    def contact_state(self):
        """(multipliers, omega, normals, candidate) of the last step for the adjoint's contact rows."""
        k = len(self._friction["nodes"])
        mult = torch.zeros(k, dtype=torch.float64, device=self.device)
        omega = torch.zeros_like(mult)
        normal = torch.zeros((k, 3), dtype=torch.float64, device=self.device)
        candidate = torch.zeros(k, dtype=torch.int32, device=self.device)
        check(lib().dphd_contact_state(self._h, None, ptr(mult), ptr(omega), ptr(normal), ptr(candidate)))
        return mult, omega, normal, candidate

    @property
    def boundary_parameters(self):
        """The four frictional boundary parameters (a copy), or None; assign to move the obstacle."""
        if self._friction is None:
            return None
        if self._factor_key is None:
            raise RuntimeError("boundary_parameters needs a built solver; step or prepare first")
        out = np.zeros(4)
        check(lib().dphd_boundary_parameters(self._h, None, ptr(out)))
        return torch.as_tensor(out)

    @boundary_parameters.setter
    def boundary_parameters(self, values):
        values = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(4))
        check(lib().dphd_boundary_parameters(self._h, ptr(values), None), "boundary_parameters")

    @property
    def mesh_contact_pairs(self):
        """Vertex-face pairs within the contact radius in the last step, as a (1,) tensor, or None."""
        if self._mesh_contact is None:
            return None
        out = ctypes.c_int()
        check(lib().dphd_mesh_contact_pairs(self._h, ctypes.byref(out)))
        return torch.tensor([max(out.value, 0)], dtype=torch.int32)

    @property
    def mesh_contact_pair_counts(self):
        """(faces_a vertices vs faces_b, faces_b vertices vs faces_a) pairs of the last step, or None."""
        if self._mesh_contact is None:
            return None
        out = np.zeros(2, dtype=np.int32)
        check(lib().dphd_mesh_contact_pair_counts(self._h, out.ctypes.data))
        return tuple(max(int(c), 0) for c in out)

    def mesh_contact_force(self, position: torch.Tensor, velocity: torch.Tensor) -> torch.Tensor:
        """The mesh contact force [N] of float64 positions and velocities (MeshContact.evaluate)."""
        if self._mesh_contact is None:
            raise ValueError("this solver has no mesh_contact")
        if self._factor_key is None:
            raise RuntimeError("mesh_contact_force needs a built solver; step or prepare first")
        position = position.to(device=self.device, dtype=torch.float64).contiguous()
        velocity = velocity.to(device=self.device, dtype=torch.float64).contiguous()
        out = torch.empty_like(position)
        check(lib().dphd_mesh_contact_force(self._h, ptr(position), ptr(velocity), ptr(out), _stream()), "mesh_contact")
        return out

    def deformation_determinants(self, state) -> torch.Tensor:
        """det(F) at all eight Hex8 Gauss samples as ``(cells, 8)`` float64 (DP/solver.py:360-378)."""
        if not self._hex:
            raise NotImplementedError("deformation_determinants currently requires hex_mesh")
        q = state.particle_q
        if q.device != self.device or tuple(q.shape) != (self._n, 3) or q.dtype != torch.float32:
            raise ValueError("state must have this solver's particle count and device")
        if self._factor_key is None:
            raise RuntimeError("deformation_determinants needs a built solver; step or prepare first")
        q = q.detach().contiguous()
        out = torch.empty((self._element_count, 8), dtype=torch.float64, device=self.device)
        check(lib().dphd_deformation_determinants(self._h, ptr(q), ptr(out), _stream()), "deformation_determinants")
        return out

    def apply_inverse(self, rhs: torch.Tensor) -> torch.Tensor:
        """A^{-1} rhs for every body through the two device SpMVs (FP64 factor only)."""
        rhs = rhs.to(device=self.device, dtype=torch.float64).contiguous()
        out = torch.empty_like(rhs)
        check(lib().dphd_apply_inverse(self._h, ptr(rhs), ptr(out), _stream()), "apply_inverse")
        return out
