"""The benchmark configurations on the oracle (CPU), at reduced sizes: adjoint gradients against
FP64 central finite differences (the metric's gradient check: 1e-4 relative).

Per-element Jacobian weights (``jacobian_mode="element"``) make the adjoint the exact derivative of
the step for heterogeneous materials. The losses are taken on the FP64 solution of one step (float32
states between steps would hide the parameter's effect below their rounding); the chaining of steps
is covered by the GPU rollouts against the oracle.
"""

import numpy as np

from rollouts import hex_oracle, log_young_gradient, tet_oracle
from scenes import config2, config3, tip_rotation

TIGHT = dict(rel_tol=1e-13, abs_tol=1e-16)


def _fd(f, lo, hi):
    return (f(hi) - f(lo)) / (hi - lo)


def test_config2_gradients_match_central_differences():
    scene = {**config2(k=3), **TIGHT}
    rng = np.random.default_rng(3)
    v0 = (rng.normal(0.0, 0.5, scene["vertices"].shape) * scene["free"][:, None]).astype(np.float32)
    top, targets = scene["top"], scene["targets"]

    def step(materials=None, qd0=v0):
        s = dict(scene, materials=scene["materials"] if materials is None else materials)
        body = tet_oracle(s, max_iterations=5000, jacobian_mode="element")
        return body, body.forward(scene["dt"], scene["vertices"], qd0)

    def loss(out):
        return float(((out["x"][top] - targets) ** 2).sum())

    body, out = step()
    gx = np.zeros_like(out["x"])
    gx[top] = 2.0 * (out["x"][top] - targets)
    bw = body.backward(scene["dt"], out["x"], gx.astype(np.float32), None)
    assert bw["status"] == 0
    g_log_e = log_young_gradient(bw["materials"], scene["materials"])
    eps = 1e-4
    for e in (0, 17, 40, 101):
        def at(t, e=e):
            m = scene["materials"].astype(np.float64).copy()
            m[e, :2] *= np.exp(t)
            return loss(step(m)[1])

        fd = _fd(at, -eps, eps)
        assert abs(g_log_e[e] - fd) <= 1e-4 * abs(fd), (e, g_log_e[e], fd)
    free = np.flatnonzero(scene["free"])
    for i, d in ((free[5], 2), (free[20], 0), (free[33], 1)):
        lo, hi = np.float32(v0[i, d] - 1e-3), np.float32(v0[i, d] + 1e-3)

        def at(value, i=i, d=d):
            v = v0.copy()
            v[i, d] = value
            return loss(step(qd0=v)[1])

        fd = _fd(at, float(lo), float(hi))
        assert abs(bw["qd"][i, d] - fd) <= 1e-4 * abs(fd), (i, d, bw["qd"][i, d], fd)


def test_config3_tip_rotation_gradient_matches_central_differences():
    # The twist geometry with the stiffness scaled by 1e-2: at the metric's 5 mm cells and dt the
    # fixed point contracts too slowly to reach the 1e-12 gates FD needs.
    scene = {**config3(8, 2, 2), "rel_tol": 1e-12, "abs_tol": 1e-16}
    scene["materials"] = scene["materials"] * 1e-2
    rest = scene["vertices"]

    def step(materials=None):
        s = dict(scene, materials=scene["materials"] if materials is None else materials)
        body = hex_oracle(s, max_iterations=5000, jacobian_mode="element")
        return body, body.forward(scene["dt"], scene["q0"], np.zeros_like(scene["q0"]))

    body, out = step()
    x = out["x"]
    # d(tip rotation)/dx of the closed form, by central differences (it is smooth in x).
    gx = np.zeros_like(x)
    for i in scene["tip"]:
        for d in (1, 2):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += 1e-7
            xm[i, d] -= 1e-7
            gx[i, d] = (tip_rotation(xp, rest, scene["tip"]) - tip_rotation(xm, rest, scene["tip"])) / 2e-7
    bw = body.backward(scene["dt"], x, gx.astype(np.float32), None)
    assert bw["status"] == 0
    g_log_e = log_young_gradient(bw["materials"], scene["materials"])
    eps = 1e-4
    for e in (22, 26, 31):  # x-cells 5..7, next to the tip
        def at(t, e=e):
            m = scene["materials"].copy()
            m[e] *= np.exp(t)
            return tip_rotation(step(m)[1]["x"], rest, scene["tip"])

        fd = _fd(at, -eps, eps)
        assert abs(g_log_e[e] - fd) <= 1e-4 * abs(fd), (e, g_log_e[e], fd)
