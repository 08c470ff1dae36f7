"""The tangent-linear reference of the adjoint tests: a forward-mode
directional derivative of an adaptive run's loss, built only from a plain
RHS callable (the GPU solver's sgnx_rhs in the GPU tests, the oracle's in the
CPU test that pins it against differences of whole runs)."""
import numpy as np


def tangent_linear(rhs, U0, theta, dU0, dtheta, steps, nodes, Ubar, Gbar, rel=1e-5):
    """The tangent-linear reference: the directional derivative of
    L = <Ubar, U_N> + sum_k <Gbar_k, h_k[nodes] - theta[nodes]> along
    (dU0, dtheta), by carrying the tangent through the recorded BS3 steps
    (rk.cpp:99-108, step sizes fixed) next to the primal run. Every
    Jacobian-vector product of the RHS, in U and theta together, is a
    Richardson-extrapolated central difference of the plain RHS
    rhs(U, theta, t), 2 C(e/2) - C(e), with the step scaled to the tangent,
    so the reference shares no code with the adjoint kernels and stays
    accurate along rough directions, where a difference of whole runs does
    not. (Where the two sides of an interface have equal wave speeds, the
    Rusanov flux's max(sL, sR) leaves a term linear in e in a plain central
    difference; the extrapolation removes it, leaving O(e^2).)"""
    U = U0.copy()
    d = dU0.copy()
    dL = 0.0
    scale = np.abs(U0).max()

    def jvp(X, dX, t):
        m = max(np.abs(dX).max(), np.abs(dtheta).max())
        e = rel * scale / m

        def central(h):
            return (rhs(X + h * dX, theta + h * dtheta, t) -
                    rhs(X - h * dX, theta - h * dtheta, t)) / (2.0 * h)
        return 2.0 * central(0.5 * e) - central(e)

    for k, (t, dt) in enumerate(steps):
        k1 = rhs(U, theta, t)
        d1 = jvp(U, d, t)
        y2, e2 = U + 0.5 * dt * k1, d + 0.5 * dt * d1
        k2 = rhs(y2, theta, t + 0.5 * dt)
        d2 = jvp(y2, e2, t + 0.5 * dt)
        y3, e3 = U + 0.75 * dt * k2, d + 0.75 * dt * d2
        k3 = rhs(y3, theta, t + 0.75 * dt)
        d3 = jvp(y3, e3, t + 0.75 * dt)
        U = U + dt * ((2.0 / 9.0 * k1 + 1.0 / 3.0 * k2) + 4.0 / 9.0 * k3)
        d = d + dt * ((2.0 / 9.0 * d1 + 1.0 / 3.0 * d2) + 4.0 / 9.0 * d3)
        if Gbar is not None and k < len(Gbar):
            dL += Gbar[k] @ (d[nodes] - dtheta[nodes])
    return dL + Ubar @ d
