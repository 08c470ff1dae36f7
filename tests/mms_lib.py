"""The manufactured-solution convergence run of apps/convergence2d.cpp:92-148
around any RHS (test infrastructure): the forced right-hand side R(U) + G,
fixed RK4 steps in rk.cpp:23-39's association, and the L2 error of h in the
quadrature of convergence2d.cpp:128-141."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "convergence2d_p4.json")


def golden_rows():
    with open(GOLDEN) as f:
        return {r["N"]: r for r in json.load(f)["rows"]}


def run(case, rhs, weights):
    """Integrate the case's manufactured state to t = 0.1 with rhs(U, t) plus
    the forcing; return the L2 error of h against the exact solution."""
    from sgnx import cases
    m = case.meta
    x, y, c2, dt = m["x"], m["y"], m["c2"], m["dt"]
    nd = case.nd

    def f(t, U):
        dU = rhs(U, t)
        dU += cases.mms_forcing(x, y, t, c2).reshape(-1)
        return dU

    U = case.U0.copy()
    t = 0.0
    for _ in range(m["steps"]):
        # rk.cpp:23-39
        k1 = f(t, U)
        k2 = f(t + 0.5 * dt, U + 0.5 * dt * k1)
        k3 = f(t + 0.5 * dt, U + 0.5 * dt * k2)
        k4 = f(t + dt, U + dt * k3)
        U += (dt / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
        t += dt
    (h, _, _, _, _), _, _, _ = cases.mms_prims(x, y, m["t_end"])
    err = (U[:nd] - h).reshape(-1, 5, 5)               # [element, b, a]
    qw = np.asarray(weights)
    J = 0.5 * (case.Lx / case.Nx)
    e2 = float(np.sum(qw[None, :, None] * qw[None, None, :] * (J * J) * err * err))
    return float(np.sqrt(e2))
