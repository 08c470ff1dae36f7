"""Pins of the oracle by mathematics rather than by the reference's bits:
closed forms of the P = 4 GLL tables, the Bogacki-Shampine stability
polynomial, invariants of the semi-discretisation, and the side extension
(walls = the mirror-doubled periodic domain; the inflow ghost = an eigenmode of
the linearised system)."""
import math

import numpy as np
import pytest

import oracle_lib as OL


def tables(P, gauss=0):
    n = P + 1
    arrs = [np.zeros(n), np.zeros(n * n)] + [np.zeros(n) for _ in range(5)]
    OL.oracle().orc_fr_operators(P, gauss, *[OL.dp(a) for a in arrs])
    return dict(zip(["x", "D", "lL", "lR", "gLp", "gRp", "qw"], arrs))


def test_gll_p4_closed_forms():
    T = tables(4)
    s = math.sqrt(3.0 / 7.0)
    np.testing.assert_allclose(T["x"], [-1.0, -s, 0.0, s, 1.0], rtol=0, atol=2.3e-16)  # 2 ulp
    np.testing.assert_allclose(T["qw"], [0.1, 49 / 90, 32 / 45, 49 / 90, 0.1], rtol=4.5e-16)  # 2 ulp
    assert list(T["lL"]) == [1.0, 0.0, 0.0, 0.0, 0.0]
    assert list(T["lR"]) == [0.0, 0.0, 0.0, 0.0, 1.0]
    # gL'(-1) = -P(P+1)/2 - ... = -12.5 and gR'(+1) = +12.5 exactly (Huynh g_DG)
    assert T["gLp"][0] == -12.5 and T["gRp"][4] == 12.5
    # g_L(x) = (L_4 - L_5)/2, so gL' = (L4' - L5')/2 at the nodes
    lib = OL.oracle()
    for i, xi in enumerate(T["x"]):
        want = 0.5 * (lib.orc_legendre_deriv(4, xi) - lib.orc_legendre_deriv(5, xi))
        assert T["gLp"][i] == pytest.approx(want, abs=1e-13)


@pytest.mark.parametrize("P", range(1, 7))
def test_operators_exact_on_polynomials(P):
    """test/test_fr_operators.cpp:48-111: D exact for degree <= P, quadrature
    exact for degree <= 2P-1, FR derivative exact for continuous data."""
    T = tables(P)
    n = P + 1
    x = T["x"]
    D = T["D"].reshape(n, n)
    a = np.sin(0.7 * np.arange(P + 1) + 1.0)
    f = np.polyval(a[::-1], x)
    df = np.polyval(np.polyder(a[::-1]), x)
    np.testing.assert_allclose(D @ f, df, atol=1e-11)
    for q in range(2 * P):
        exact = 2.0 / (q + 1) if q % 2 == 0 else 0.0
        assert abs(np.dot(T["qw"], x ** q) - exact) < 1e-12
    # frDerivative with the exact face values is the exact derivative
    fl, fr = np.polyval(a[::-1], -1.0), np.polyval(a[::-1], 1.0)
    fr_der = D @ f + (fl - T["lL"] @ f) * T["gLp"] + (fr - T["lR"] @ f) * T["gRp"]
    np.testing.assert_allclose(fr_der, df, atol=1e-11)
    # a dropped sign in the correction would show: the lifting of a unit jump
    # integrates to -1 / +1 (g_L(-1) = 1, g_L(1) = 0)
    assert np.dot(T["qw"], T["gLp"]) == pytest.approx(-1.0, abs=1e-13)
    assert np.dot(T["qw"], T["gRp"]) == pytest.approx(1.0, abs=1e-13)


def test_bs3_step_is_the_stability_polynomial():
    """One BS3 step of u' = lam u multiplies u by 1 + z + z^2/2 + z^3/6."""
    lib = OL.oracle()
    for lam, dt in ((-3.0, 0.1), (-0.5, 0.37), (2.0, 0.05)):
        u = np.array([1.0])
        lib.orc_bs3_step_ode(2, lam, OL.dp(u), 1, 0.0, dt)
        z = lam * dt
        assert u[0] == pytest.approx(1 + z + z * z / 2 + z ** 3 / 6, rel=1e-15)
        v = np.array([1.0])
        lib.orc_rk4_step_ode(2, lam, OL.dp(v), 1, 0.0, dt)
        assert v[0] == pytest.approx(1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24, rel=1e-15)


def test_adaptive_error_norm_closed_form():
    """On u' = lam u the embedded error of BS3(2) is u0 (stability polynomial of
    the 3rd-order solution minus that of the embedded 2nd-order one); the first
    logged error norm equals |that| / (atol + rtol*max(|u0|,|u1|))."""
    lam, dt, atol, rtol = -2.0, 0.05, 1e-8, 1e-6
    z = lam * dt
    R3 = 1 + z + z * z / 2 + z ** 3 / 6
    # embedded: u0 + dt(7/24 k1 + 1/4 k2 + 1/3 k3 + 1/8 k4)
    k1 = lam
    k2 = lam * (1 + 0.5 * z)
    k3 = lam * (1 + 0.75 * z * (1 + 0.5 * z))
    k4 = lam * R3
    Remb = 1 + dt * (7 / 24 * k1 + 0.25 * k2 + 1 / 3 * k3 + 0.125 * k4)
    err = abs(R3 - Remb) / (atol + rtol * max(1.0, abs(R3)))
    u = np.array([1.0])
    st = OL.OrcStats()
    opts = OL.OrcOpts(atol=atol, rtol=rtol, max_accepted=1)
    OL.oracle().orc_integrate_ode(2, lam, OL.dp(u), 1, 0.0, 10.0, dt, OL.C.byref(opts), OL.C.byref(st))
    if err <= 1.0:
        assert st.accepted == 1 and st.rejected == 0
        assert st.maxNormLast == pytest.approx(err, rel=1e-9)
        assert u[0] == pytest.approx(R3, rel=1e-15)


def test_lake_at_rest_and_freestream():
    """test/test_sgn2d.cpp:55-77 on the oracle (walls included)."""
    from sgnx import cases
    for bc in (("periodic",) * 4, ("wall",) * 4):
        c = cases.gate(Nx=7, Ny=5, bc=bc)
        o = OL.Oracle.from_case(c)
        Hn = c.depth["Hnod"]
        z = np.zeros_like(Hn)
        R = o.rhs(cases.prim_to_state(Hn.copy(), z, z, z, z))
        assert np.abs(R).max() < 1e-12
    flat = OL.Oracle(4, 8, 6, 10.0, 8.0, (9.81, 3 * math.sqrt(9.81), 1.5, 0.0))
    nd = 8 * 6 * 25
    one, zero = np.ones(nd), np.zeros(nd)
    R = flat.rhs(cases.prim_to_state(one, 0.5 * one, -0.3 * one, zero, zero))
    assert np.abs(R).max() < 1e-11


def mirror_y(U, Nx, Ny):
    """State on the doubled (2*Ny) mesh: rows Ny..2Ny-1 are the mirror image of
    rows Ny-1..0 (node b -> 4-b) with hv negated."""
    Nd = Nx * Ny * 25
    V = U.reshape(5, Ny, Nx, 5, 5)          # var, ey, ex, b, a
    M = V[:, ::-1, :, ::-1, :].copy()
    M[2] = -M[2]
    return np.concatenate([V, M], axis=1).reshape(-1), Nd


def mirror_x(U, Nx, Ny):
    V = U.reshape(5, Ny, Nx, 5, 5)
    M = V[:, :, ::-1, :, ::-1].copy()
    M[1] = -M[1]
    return np.concatenate([V, M], axis=2).reshape(-1)


def test_walls_equal_mirror_doubled_periodic_domain():
    """A wall is the symmetry plane of a periodic domain twice as large: the
    wall-side RHS on [0,L] equals, bit for bit, the doubled periodic domain's
    RHS restricted to [0,L] (same still depth on that half)."""
    from sgnx import cases
    Nx, Ny = 6, 4
    Lx, Ly = 5.4, 4.4
    U = cases.random_state(Nx * Ny * 25, seed=11)

    def H(x, y):
        return 1.0 + 0.1 * np.cos(0.9 * x) + 0.05 * np.sin(1.3 * y)

    def H_my(x, y):                       # mirror of H about y = Ly (doubled in y)
        return H(x, np.where(y <= Ly, y, 2 * Ly - y))

    def H_mx(x, y):
        return H(np.where(x <= Lx, x, 2 * Lx - x), y)

    par = (9.81, 3 * math.sqrt(9.81), 1.5, 0.0)
    # walls in y vs doubled-in-y periodic
    bc = ("periodic", "periodic", "wall", "wall")
    dw = cases.depth_arrays(4, Nx, Ny, Lx, Ly, H, bc)
    Rw = OL.Oracle(4, Nx, Ny, Lx, Ly, par, dw, bc).rhs(U)
    U2, Nd = mirror_y(U, Nx, Ny)
    dd = cases.depth_arrays(4, Nx, 2 * Ny, Lx, 2 * Ly, H_my, ("periodic",) * 4)
    R2 = OL.Oracle(4, Nx, 2 * Ny, Lx, 2 * Ly, par, dd).rhs(U2)
    half = R2.reshape(5, 2 * Ny, Nx, 25)[:, :Ny].reshape(-1)
    assert np.array_equal(Rw.view(np.uint64), half.view(np.uint64))
    # walls in x vs doubled-in-x periodic
    bc = ("wall", "wall", "periodic", "periodic")
    dw = cases.depth_arrays(4, Nx, Ny, Lx, Ly, H, bc)
    Rw = OL.Oracle(4, Nx, Ny, Lx, Ly, par, dw, bc).rhs(U)
    U2 = mirror_x(U, Nx, Ny)
    dd = cases.depth_arrays(4, 2 * Nx, Ny, 2 * Lx, Ly, H_mx, ("periodic",) * 4)
    R2 = OL.Oracle(4, 2 * Nx, Ny, 2 * Lx, Ly, par, dd).rhs(U2)
    half = R2.reshape(5, Ny, 2 * Nx, 25)[:, :, :Nx].reshape(-1)
    assert np.array_equal(Rw.view(np.uint64), half.view(np.uint64))


def test_mass_conserved_with_walls():
    """test/test_sgn2d.cpp:159-179: adaptive run conserves mass to round-off;
    also with reflective walls (zero normal mass flux)."""
    from sgnx import cases
    for bc in (("periodic",) * 4, ("wall",) * 4):
        c = cases.gate(Nx=8, Ny=6, bc=bc)
        o = OL.Oracle.from_case(c)
        M0 = o.mass(c.U0)
        dt0 = o.cfl_time_step(c.U0, 0.5)
        U, st, _, _, _ = o.integrate(c.U0, 0.0, 0.1, dt0, OL.OrcOpts(atol=1e-9, rtol=1e-7))
        assert not st.aborted and np.all(np.isfinite(U))
        assert abs(o.mass(U) - M0) / abs(M0) < 1e-13


def test_margin_lake_at_rest_and_mass():
    """The margin workload at 24x12 on the oracle: over its rough 4000 -> 50 m
    margin the fluid at rest stays at rest (RHS zero to round-off relative to
    g H), and the uplift's adaptive run conserves mass with walls in x and a
    periodic y axis."""
    from sgnx import cases
    c = cases.margin(seed=5, Nx=24, Ny=12)
    o = OL.Oracle.from_case(c)
    Hn = c.depth["Hnod"]
    z = np.zeros_like(Hn)
    R = o.rhs(cases.prim_to_state(Hn.copy(), z, z, z, z))
    assert np.abs(R).max() < 1e-12 * 9.81 * Hn.max() ** 2
    M0 = o.mass(c.U0)
    dt0 = o.cfl_time_step(c.U0, 0.5)
    U, st, _, _, _ = o.integrate(c.U0, 0.0, 1e9, dt0, OL.OrcOpts(max_accepted=20))
    assert st.accepted == 20 and np.all(np.isfinite(U))
    assert abs(o.mass(U) - M0) / abs(M0) < 1e-13
    o.close()


def test_tangent_linear_reference_against_run_differences():
    """The tangent-linear reference of the adjoint tests (tlm_lib), driven by
    the oracle's RHS, against Richardson-extrapolated central differences of
    the oracle's whole 6-step BS3 run (2 C(5e-4) - C(1e-3)) along three
    smooth directions of the state and the still depth: within 2e-5
    (the run differences' own error)."""
    from sgnx import cases
    from tlm_lib import tangent_linear
    from test_gpu_adjoint import theta_depth
    c = cases.gate(Nx=8, Ny=6)
    rng = np.random.default_rng(11)
    theta = c.depth["Hnod"].copy()

    def orc(th):
        return OL.Oracle(c.P, c.Nx, c.Ny, c.Lx, c.Ly, c.par, theta_depth(th, c.Nx, c.Ny, c.bc),
                         c.bc, c.inflow)

    def rhs(U, th, t):
        o = orc(th)
        r = o.rhs(U, t)
        o.close()
        return r

    o0 = orc(theta)
    dt0 = o0.cfl_time_step(c.U0, 0.5)
    o0.close()
    steps = [(i * dt0, dt0) for i in range(6)]
    nodes = rng.choice(c.nd, 5, replace=False)
    Ubar = rng.normal(size=c.U0.size)
    Gbar = rng.normal(size=(6, 5))

    def loss(U0, th):
        o = orc(th)
        U = U0.copy()
        L = 0.0
        for k, (t, dt) in enumerate(steps):
            U = o.bs3_step(U, t, dt)
            L += Gbar[k] @ (U[nodes] - th[nodes])
        o.close()
        return L + Ubar @ U

    for _ in range(3):
        dth = cases.smooth_direction(c, rng, n=3, amp=1.0)
        dU = np.concatenate([cases.smooth_direction(c, rng, n=3, amp=1.0) for _ in range(5)])
        tl = tangent_linear(rhs, c.U0, theta, dU, dth, steps, nodes, Ubar, Gbar)

        def central(eps):
            return (loss(c.U0 + eps * dU, theta + eps * dth) -
                    loss(c.U0 - eps * dU, theta - eps * dth)) / (2.0 * eps)
        fd = 2.0 * central(5e-4) - central(1e-3)
        assert abs(tl - fd) <= 2e-5 * abs(tl), (tl, fd)


def test_inflow_ghost_is_a_linear_eigenmode():
    """Rows 1-3 of the linearised system fix the ghost amplitudes; row 4
    (hp equation) then holds only if k is on the slow branch (test/
    test_dispersion.cpp header). Check the residual of row 4 with the table."""
    from sgnx import cases
    H, T, A = 0.45, 1.0, 0.01
    tab, omega = cases.inflow_table(1, 4, np.full(5, H), T, A)
    tab = tab.reshape(5, 5, 3)[0]
    g, gamma = 9.81, 1.5
    c2 = (3.0 * math.sqrt(g * H)) ** 2
    k = cases.wavenumber(omega, H)
    # complex amplitudes of exp(i(kx - wt)) with eta = Re(i A e^{i phi}) = A sin(wt) at x = 0
    eta = 1j * tab[0, 1]
    u = 1j * tab[1, 1] / H
    p = 1j * tab[4, 1] / H
    w = tab[3, 2] / H          # w = W cos(wt) = Re(W e^{i phi}) at x = 0
    # row 1: -i w eta + i k H u = 0
    assert abs(-1j * omega * eta + 1j * k * H * u) < 1e-14
    # row 2: -i w u + i k g eta + i k p = 0
    assert abs(-1j * omega * u + 1j * k * g * eta + 1j * k * p) < 1e-12
    # row 3: -i w w_hat = gamma p / H
    assert abs(-1j * omega * w - gamma * p / H) < 1e-12
    # row 4 (dispersion): -i w p = -c^2 i k u - 2 c^2 w / H
    assert abs(-1j * omega * p + c2 * 1j * k * u + 2 * c2 * w / H) < 1e-9
