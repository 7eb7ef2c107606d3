"""The oracle side of a parity fixture: the rollout and linear objective of ``diffphd_amd.parity.run``
on the FP64 CPU oracle, with the per-step adjoints chained in reverse (dL/dq_out of step k is
dL/dq_in of step k+1; the parameter gradients add up over the steps), and the thresholds of the
reference's tools/parity/compare.py:23-46 restated for comparing two result sets.
"""

import numpy as np

from diffphd_amd import meshes, parity
from oracle import oracle as O

U32 = 2.0**-24  # unit roundoff of float32 states (compare.py:21)


def oracle_body(fx):
    rest = np.asarray(fx["rest_vertices"], dtype=np.float64)
    tets = np.asarray(fx["tets"], dtype=np.int32)
    dirichlet = np.asarray(fx["dirichlet"], dtype=bool)
    kw = {}
    if "muscle_directions" in fx:
        kw.update(muscle_directions=np.asarray(fx["muscle_directions"], np.float64),
                  muscle_stiffness=np.asarray(fx["muscle_stiffness"], np.float64))
    kind = str(fx["contact_kind"])
    if kind != "none":
        params = parity.mesh_boundary(fx["contact_parameters"]) if kind == "mesh" else tuple(
            np.asarray(fx["contact_parameters"], np.float64))
        kw.update(frictional_boundary=(kind, params), frictional_vertices=np.asarray(fx["contact_vertices"], np.int32),
                  friction=float(fx["friction"]), restitution=float(fx["restitution"]))
    if "mesh_contact_parameters" in fx:
        kw["mesh_contact"] = (*np.asarray(fx["mesh_contact_parameters"], np.float64), fx["mesh_contact_faces_a"],
                              fx["mesh_contact_faces_b"])
    return O.OracleBody(
        "tet4", tets, poses=meshes.tet_poses(rest, tets), materials=np.stack([fx["mu"], fx["lambda"]], axis=1),
        density=float(fx["density"]), inverse_mass=np.where(dirichlet, 0.0, 1.0).astype(np.float32),
        gravity=np.asarray(fx["gravity"], np.float64), dt=float(fx["dt"]), abs_tol=float(fx["abs_tol"]),
        rel_tol=float(fx["rel_tol"]), max_iterations=int(fx["max_iterations"]), window=int(fx["aa_window"]),
        material=str(fx["material"]), **kw,
    )


def run(fx):
    """The port outputs of ``fx`` computed by the oracle (names of ``diffphd_amd.parity.run``)."""
    body = oracle_body(fx)
    steps, dt = int(fx["steps"]), float(fx["dt"])
    damping = float(fx["velocity_damping"]) if "velocity_damping" in fx else None
    explicit = "dirichlet_positions" in fx
    q, v = np.asarray(fx["q"], np.float32), np.asarray(fx["v"], np.float32)
    q_in, records, contact = [], [], []
    for k in range(steps):
        targets = None if "activations" not in fx else 1.0 - np.asarray(fx["activations"][k], np.float32).astype(
            np.float64)
        d = np.asarray(fx["dirichlet_positions"][k], np.float32) if explicit else None
        q_in.append(q.astype(np.float64))
        out = body.forward(dt, q, v, np.asarray(fx["f_ext"][k], np.float32), dirichlet=d, targets=targets)
        records.append((out, targets))
        if body.contact is not None:
            contact.append({key: np.copy(body.contact[key]) for key in ("multiplier", "omega", "normal", "candidate")})
        q, v = out["q"], out["qd"]
        if damping is not None:
            v = (v * np.float32(damping)).astype(np.float32)
    loss = float(np.sum(q.astype(np.float64) * fx["objective_q"]) + np.sum(v.astype(np.float64) * fx["objective_v"]))
    gq = np.asarray(fx["objective_q"], np.float32)
    gv = np.asarray(fx["objective_v"], np.float32)
    gmat, gf, gact, bits = 0.0, [None] * steps, [None] * steps, 0
    for k in reversed(range(steps)):
        out, targets = records[k]
        if damping is not None:
            gv = (gv * np.float32(damping)).astype(np.float32)
        if contact:
            body.contact.update({key: np.copy(val) for key, val in contact[k].items()})
        bw = body.backward(dt, out["x"], gq, gv, explicit_dirichlet=explicit, targets=targets,
                           activations=targets is not None)
        gmat = gmat + bw["materials"]
        gf[k] = bw["f"].astype(np.float64)
        if targets is not None:
            gact[k] = bw["activations"]
        gq, gv = bw["q"], bw["qd"]
        bits |= bw["status"]
    x = np.stack([r[0]["x"] for r in records])
    scale = 1.0 if damping is None else damping
    result = {
        "q_next": x,
        "v_next": np.stack([(x[k] - q_in[k]) / dt * scale for k in range(steps)]),
        "iterations": np.array([r[0]["iterations"] for r in records]),
        "converged": np.array([(r[0]["status"] & 1) == 0 for r in records]),
        "status_bits": np.array([r[0]["status"] for r in records] + [bits]),
        "dL_dq": gq.astype(np.float64),
        "dL_dv": gv.astype(np.float64),
        "dL_df_ext": np.stack(gf),
        "dL_dmaterials_per_element": np.asarray(gmat, np.float64),
        "loss": loss,
        "free": ~np.asarray(fx["dirichlet"], bool),
    }
    if contact:
        lam = np.stack([c["multiplier"] for c in contact])
        result["multipliers"] = lam
        result["active"] = lam > 0
        if str(fx["contact_kind"]) == "planar":
            params = np.asarray(fx["contact_parameters"], np.float64)
            nodes = np.asarray(fx["contact_vertices"])
            result["gaps"] = (x[:, nodes] @ params[:3] + params[3]) / np.linalg.norm(params[:3])
    if "activations" in fx:
        result["dL_dactivations"] = np.stack(gact)
    return result


def compare_results(fx, port, ref):
    """Failed checks of ``port`` against ``ref`` under compare.py's thresholds (empty when all pass):
    positions and velocities per step, iterations +-1, convergence, active set and multipliers with
    contact, gradients on free particles (1e-5 relative L2 without active contact, 1e-4 with) and the
    per-element materials."""
    steps, dt = int(fx["steps"]), float(fx["dt"])
    abs_tol, rel_tol = float(fx["abs_tol"]), float(fx["rel_tol"])
    free = ~np.asarray(fx["dirichlet"], bool)
    contact = str(fx["contact_kind"]) != "none"
    any_active = contact and bool(np.asarray(ref["active"]).any())
    tol = 1.0e-4 if any_active else 1.0e-5
    failed = []
    for k in range(steps):
        q, v = ref["q_next"][k], ref["v_next"][k]
        bound = 10.0 * (rel_tol * np.linalg.norm(q) + abs_tol) + 10.0 * k * U32 * (
            np.linalg.norm(q) + dt * np.linalg.norm(v))
        err = np.linalg.norm(port["q_next"][k] - q)
        if err > bound:
            failed.append(f"step {k} positions {err:.3e} > {bound:.3e}")
        err = np.linalg.norm(port["v_next"][k] - v)
        if err > bound / dt:
            failed.append(f"step {k} velocities {err:.3e} > {bound / dt:.3e}")
        if abs(int(port["iterations"][k]) - int(ref["iterations"][k])) > 1:
            failed.append(f"step {k} iterations {port['iterations'][k]} vs {ref['iterations'][k]}")
        if bool(port["converged"][k]) != bool(ref["converged"][k]):
            failed.append(f"step {k} converged")
        if contact:
            if np.any(port["active"][k] != ref["active"][k]):
                failed.append(f"step {k} active set")
            lam_ref, lam_port = ref["multipliers"][k], port["multipliers"][k]
            allowed = 5.0e-6 * np.abs(lam_ref) + 1.0e-6 * np.abs(lam_ref).max() + 1.0e-12
            if np.any(np.abs(lam_port - lam_ref) > allowed):
                failed.append(f"step {k} multipliers {np.abs(lam_port - lam_ref).max():.3e}")
            if "gaps" in ref and np.abs(port["gaps"][k] - ref["gaps"][k]).max() > bound:
                failed.append(f"step {k} gaps")
    grads = [("dL_dq", ref["dL_dq"], port["dL_dq"]), ("dL_dv", ref["dL_dv"], port["dL_dv"])]
    grads += [(f"dL_df_ext[{k}]", ref["dL_df_ext"][k], port["dL_df_ext"][k]) for k in range(steps)]
    if "dL_dactivations" in ref:
        grads += [(f"dL_dactivations[{k}]", ref["dL_dactivations"][k][:, None], port["dL_dactivations"][k][:, None])
                  for k in range(steps)]
    for name, g_ref, g_port in grads:
        if name.startswith("dL_dactivations"):
            err, scale = np.linalg.norm(g_port - g_ref), np.linalg.norm(g_ref)
        else:
            err, scale = np.linalg.norm((g_port - g_ref)[free]), np.linalg.norm(g_ref[free])
        if err > tol * scale:
            failed.append(f"{name} {err:.3e} > {tol:.0e} x {scale:.3e}")
    gm_ref, gm_port = ref["dL_dmaterials_per_element"], port["dL_dmaterials_per_element"]
    for col, name in ((0, "mu"), (1, "lambda")):
        bound = tol * np.abs(gm_ref[:, col].sum()) + steps * U32 * np.abs(gm_port[:, col]).sum()
        err = abs(gm_port[:, col].sum() - gm_ref[:, col].sum())
        if err > bound:
            failed.append(f"materials {name} {err:.3e} > {bound:.3e}")
    return failed
