"""Shared test scenes: the same inputs for the product solver and the oracle."""

from pathlib import Path

import numpy as np

from diffphd_amd import meshes
from diffphd_amd.model import Model

GOLDEN = Path(__file__).resolve().parent / "golden"


def golden(name):
    return dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))


def tet_model(device, tets, poses, materials, free, q, gravity, requires_grad=False, voff=None, eoff=None):
    return Model(
        particle_q=q,
        particle_qd=np.zeros_like(q),
        particle_inv_mass=np.asarray(free, dtype=np.float32),
        particle_world=np.zeros(len(q), np.int32),
        gravity=np.asarray(gravity, dtype=np.float32)[None],
        tet_indices=np.asarray(tets, dtype=np.int32),
        tet_poses=poses,
        tet_materials=np.asarray(materials, dtype=np.float32),
        device=device,
        requires_grad=requires_grad,
        body_vertex_offsets=voff,
        body_element_offsets=eoff,
    )


def cube_body(k, cell=0.005, seed=0, log_e=(np.log(1e5), np.log(1e7)), poisson=0.4, origin=(0.0, 0.0, 0.0)):
    """k^3 Kuhn cube, bottom face pinned, per-tet log-uniform E (the Tet4 benchmark family)."""
    vertices, tets, poses = meshes.kuhn_tet_grid(k, k, k, cell, origin=origin)
    rng = np.random.default_rng(seed)
    young = np.exp(rng.uniform(log_e[0], log_e[1], len(tets)))
    mu, lam = meshes.lame(young, np.broadcast_to(poisson, young.shape))
    materials = np.stack([mu, lam, np.zeros_like(mu)], axis=1).astype(np.float32)
    free = vertices[:, 2] > vertices[:, 2].min()
    return vertices.astype(np.float32), tets, poses, materials, free
