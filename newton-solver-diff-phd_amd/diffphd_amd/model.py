"""Newton-shaped ``Model`` / ``State`` / ``Control`` holders over torch tensors.

The reference solver reads these fields of ``newton.Model`` (DP/solver.py:
183-241, 464-489, 745-748): particle_count, tet_count, tet_indices,
tet_poses, tet_materials, tet_activations, particle_inv_mass,
particle_world, gravity, particle_flags, body_count, shape_count. Newton is
not available here, so this module provides the same field names with torch
tensors, and a small builder for soft grids. Every ``add_soft_grid`` call is
recorded as one body, which the solver can solve independently
(``independent_bodies=True``).
"""

from __future__ import annotations

import numpy as np
import torch

from . import meshes


class State:
    def __init__(self, q, qd, f):
        self.particle_q = q
        self.particle_qd = qd
        self.particle_f = f


class Control:
    def __init__(self, tet_activations=None):
        self.tet_activations = tet_activations


class Model:
    def __init__(
        self,
        *,
        particle_q,
        particle_qd,
        particle_inv_mass,
        particle_world,
        gravity,
        tet_indices,
        tet_poses,
        tet_materials,
        device,
        requires_grad=False,
        body_vertex_offsets=None,
        body_element_offsets=None,
    ):
        dev = torch.device(device)
        self.device = dev
        self.requires_grad = requires_grad
        self.particle_q = torch.as_tensor(particle_q, dtype=torch.float32, device=dev).reshape(-1, 3).contiguous()
        self.particle_qd = torch.as_tensor(particle_qd, dtype=torch.float32, device=dev).reshape(-1, 3).contiguous()
        self.particle_inv_mass = torch.as_tensor(particle_inv_mass, dtype=torch.float32, device=dev).contiguous()
        self.particle_world = torch.as_tensor(particle_world, dtype=torch.int32, device=dev).contiguous()
        self.gravity = torch.as_tensor(gravity, dtype=torch.float32, device=dev).reshape(-1, 3).contiguous()
        self.particle_count = int(self.particle_q.shape[0])
        if tet_indices is None:
            self.tet_indices = torch.zeros(0, dtype=torch.int32, device=dev)
            self.tet_poses = torch.zeros((0, 3, 3), dtype=torch.float32, device=dev)
            self.tet_materials = torch.zeros((0, 3), dtype=torch.float32, device=dev)
        else:
            self.tet_indices = torch.as_tensor(tet_indices, dtype=torch.int32, device=dev).reshape(-1).contiguous()
            self.tet_poses = torch.as_tensor(tet_poses, dtype=torch.float32, device=dev).reshape(-1, 3, 3).contiguous()
            self.tet_materials = torch.as_tensor(tet_materials, dtype=torch.float32, device=dev).reshape(-1, 3).clone()
        self.tet_count = int(self.tet_indices.numel() // 4)
        self.tet_materials.requires_grad_(requires_grad and self.tet_count > 0)
        self.tet_activations = None
        self.particle_flags = torch.ones(self.particle_count, dtype=torch.int32, device=dev)
        self.body_count = 0
        self.shape_count = 0
        self.hex_count = 0
        n, e = self.particle_count, self.tet_count
        self.body_vertex_offsets = np.asarray(body_vertex_offsets if body_vertex_offsets is not None else [0, n], np.int32)
        self.body_element_offsets = np.asarray(
            body_element_offsets if body_element_offsets is not None else [0, e], np.int32
        )

    def state(self, requires_grad=False):
        q = self.particle_q.clone()
        qd = self.particle_qd.clone()
        f = torch.zeros_like(q)
        for t in (q, qd, f):
            t.requires_grad_(requires_grad)
        return State(q, qd, f)

    def control(self):
        return Control(None if self.tet_activations is None else self.tet_activations.clone())


class ModelBuilder:
    """Soft-grid builder with the Kuhn 6-tet cube split (see ``meshes``)."""

    def __init__(self, gravity=(0.0, 0.0, -9.81)):
        self.gravity = np.asarray(gravity, dtype=np.float64)
        self._q, self._qd, self._inv_mass = [], [], []
        self._tets, self._poses, self._materials = [], [], []
        self._voff, self._eoff = [0], [0]

    def add_soft_grid(
        self,
        pos,
        vel,
        dim_x,
        dim_y,
        dim_z,
        cell_x,
        cell_y,
        cell_z,
        density,
        k_mu,
        k_lambda,
        k_damp=0.0,
        fix_left=False,
        fix_right=False,
        fix_top=False,
        fix_bottom=False,
        rot=None,
    ):
        if rot is not None and not np.allclose(np.asarray(rot, dtype=np.float64), [0.0, 0.0, 0.0, 1.0]):
            raise NotImplementedError("only the identity grid rotation is supported")
        vertices, tets, poses = meshes.kuhn_tet_grid(dim_x, dim_y, dim_z, (cell_x, cell_y, cell_z), origin=pos)
        base = self._voff[-1]
        vertices = vertices.astype(np.float32)
        lo, hi = vertices.min(axis=0), vertices.max(axis=0)
        pinned = np.zeros(len(vertices), dtype=bool)
        if fix_left:
            pinned |= vertices[:, 0] == lo[0]
        if fix_right:
            pinned |= vertices[:, 0] == hi[0]
        if fix_bottom:
            pinned |= vertices[:, 2] == lo[2]
        if fix_top:
            pinned |= vertices[:, 2] == hi[2]
        self._q.append(vertices)
        self._qd.append(np.tile(np.asarray(vel, dtype=np.float32), (len(vertices), 1)))
        self._inv_mass.append(np.where(pinned, 0.0, 1.0).astype(np.float32))
        self._tets.append(tets + base)
        self._poses.append(poses)
        mat = np.empty((len(tets), 3), dtype=np.float32)
        mat[:, 0], mat[:, 1], mat[:, 2] = k_mu, k_lambda, k_damp
        self._materials.append(mat)
        self._voff.append(base + len(vertices))
        self._eoff.append(self._eoff[-1] + len(tets))
        return base

    def add_particles(self, pos, vel, mass):
        pos = np.asarray(pos, dtype=np.float32).reshape(-1, 3)
        mass = np.asarray(mass, dtype=np.float32)
        self._q.append(pos)
        self._qd.append(np.asarray(vel, dtype=np.float32).reshape(-1, 3))
        self._inv_mass.append(np.where(mass > 0, 1.0 / np.where(mass > 0, mass, 1.0), 0.0).astype(np.float32))
        self._voff.append(self._voff[-1] + len(pos))
        self._eoff.append(self._eoff[-1])

    def finalize(self, device="cpu", requires_grad=False):
        q = np.concatenate(self._q)
        tets = np.concatenate(self._tets) if self._tets else None
        return Model(
            particle_q=q,
            particle_qd=np.concatenate(self._qd),
            particle_inv_mass=np.concatenate(self._inv_mass),
            particle_world=np.zeros(len(q), np.int32),
            gravity=self.gravity[None],
            tet_indices=tets,
            tet_poses=np.concatenate(self._poses) if self._poses else None,
            tet_materials=np.concatenate(self._materials) if self._materials else None,
            device=device,
            requires_grad=requires_grad,
            body_vertex_offsets=self._voff,
            body_element_offsets=self._eoff,
        )


class Mesh:
    """Public mesh container of the reference (``_src/geometry/usd.py:11-17``)."""

    def __init__(self, name, vertices, indices):
        self.name = name
        self.vertices = np.asarray(vertices, dtype=np.float64)
        self.indices = np.asarray(indices)
