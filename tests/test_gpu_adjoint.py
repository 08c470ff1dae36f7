"""The discrete adjoint (sgnx_rhs_vjp, sgnx_adjoint) against central finite
differences of the oracle.

The oracle pins the forward map bit for bit (the GPU RHS and stages are
bitwise equal to it), so a derivative of the GPU's map is pinned by finite
differences of the oracle's: for a random direction d,
    <Ubar, d>  vs  (L(x + eps d) - L(x - eps d)) / (2 eps)
with L the scalar the cotangents define, and the same for single entries
(corner, edge and interior nodes, every variable), so a dropped term, a wrong
sign or a transposed operand shows. Central differences have error O(eps^2)
plus rounding O(1e-16 / eps); the bar is a relative 1e-6.

The still depth of these runs is one nodal field theta (face depths = the
mean of the two edge values meeting at a line node, one interior value on a
non-periodic side); `theta_depth` restates that rule in numpy, and the GPU's
RHS under sgnx_set_depth(theta) is checked against the oracle built from it,
bit for bit.
"""
import numpy as np
import pytest

import oracle_lib as OL
from tlm_lib import tangent_linear
from test_gpu_rhs import assert_bitwise

pytestmark = pytest.mark.gpu

EPS = 1e-6
TOL = 1e-6


def theta_depth(th, Nx, Ny, bc):
    T = np.asarray(th, dtype=np.float64).reshape(Ny, Nx, 5, 5)     # [ey, ex, b, a]
    left, right = T[:, :, :, 0], T[:, :, :, 4]                     # [ey, ex, n]
    bot, top = T[:, :, 0, :], T[:, :, 4, :]
    fx = np.empty((Ny, Nx + 1, 5))
    fx[:, 1:Nx] = 0.5 * (right[:, :-1] + left[:, 1:])
    if bc[0] == "periodic":
        fx[:, 0] = fx[:, Nx] = 0.5 * (right[:, -1] + left[:, 0])
    else:
        fx[:, 0], fx[:, Nx] = left[:, 0], right[:, -1]
    fy = np.empty((Ny + 1, Nx, 5))
    fy[1:Ny] = 0.5 * (top[:-1] + bot[1:])
    if bc[2] == "periodic":
        fy[0] = fy[Ny] = 0.5 * (top[-1] + bot[0])
    else:
        fy[0], fy[Ny] = bot[0], top[-1]
    return {"Hnod": np.ascontiguousarray(th, dtype=np.float64), "HfaceX": fx.ravel(),
            "HfaceY": fy.ravel()}


def _cases():
    from sgnx import cases
    out = [cases.gate(Nx=8, Ny=6),
           cases.gate(Nx=13, Ny=7, bc=("wall",) * 4),
           cases.gate(Nx=11, Ny=5, bc=("periodic", "periodic", "wall", "wall")),
           cases.shoal(seed=3, Nx=24, Ny=6)]
    return out


def _oracle(c, theta):
    return OL.Oracle(c.P, c.Nx, c.Ny, c.Lx, c.Ly, c.par, theta_depth(theta, c.Nx, c.Ny, c.bc),
                     c.bc, c.inflow)


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def _state(c, seed, amp=0.15):
    from sgnx import cases
    U = cases.random_state(c.nd, seed=seed, h0=0.6, amp=amp)
    U[:c.nd] += c.depth["Hnod"] - 0.6
    return U


@pytest.mark.parametrize("ci", range(4))
def test_set_depth_matches_oracle_bitwise(ci):
    import sgnx
    c = _cases()[ci]
    rng = np.random.default_rng(ci)
    theta = c.depth["Hnod"] + 0.01 * rng.uniform(-1, 1, c.nd)
    s = sgnx.SGN2D(**c.solver_kwargs())
    s.setDepth(theta)
    o = _oracle(c, theta)
    U = _state(c, 40 + ci)
    assert_bitwise(s.rhs(U, 0.3), o.rhs(U, 0.3), f"{c.name} rhs after set_depth")
    Hx, Hy = s.stillDepthGrad()
    _, oHx, oHy = o.bathymetry()
    assert_bitwise(Hx, oHx, "Hx")
    assert_bitwise(Hy, oHy, "Hy")
    s.close()


@pytest.mark.parametrize("ci", range(4))
def test_rhs_vjp_matches_finite_differences(ci):
    import sgnx
    c = _cases()[ci]
    rng = np.random.default_rng(100 + ci)
    theta = c.depth["Hnod"] + 0.01 * rng.uniform(-1, 1, c.nd)
    s = sgnx.SGN2D(**c.solver_kwargs())
    s.setDepth(theta)
    U = _state(c, 50 + ci)
    W = rng.normal(size=U.size)
    t = 0.37
    Ub, thb = s.rhs_vjp(U, W, t)
    o = _oracle(c, theta)
    dU = rng.uniform(-1.0, 1.0, U.size) * np.repeat([0.05, 0.05, 0.05, 0.02, 0.02], c.nd)
    fd = W @ (o.rhs(U + EPS * dU, t) - o.rhs(U - EPS * dU, t)) / (2 * EPS)
    assert _rel(Ub @ dU, fd) < TOL, (c.name, Ub @ dU, fd)
    # per variable: each block of U is pinned on its own
    for v in range(5):
        d = np.zeros_like(dU)
        d[v * c.nd:(v + 1) * c.nd] = dU[v * c.nd:(v + 1) * c.nd]
        fdv = W @ (o.rhs(U + EPS * d, t) - o.rhs(U - EPS * d, t)) / (2 * EPS)
        assert _rel(Ub @ d, fdv) < TOL, (c.name, v, Ub @ d, fdv)
    dth = rng.uniform(-1.0, 1.0, c.nd) * 0.05
    fdt = W @ (_oracle(c, theta + EPS * dth).rhs(U, t) -
               _oracle(c, theta - EPS * dth).rhs(U, t)) / (2 * EPS)
    assert _rel(thb @ dth, fdt) < TOL, (c.name, thb @ dth, fdt)
    # the face depths cancel between each Rusanov flux and the element's own
    # face flux (both evaluated at the same depth), so R does not depend on
    # them: differences with the face depths frozen agree to rounding
    def frozen_faces(th):
        d = theta_depth(theta, c.Nx, c.Ny, c.bc)
        d["Hnod"] = np.ascontiguousarray(th)
        return OL.Oracle(c.P, c.Nx, c.Ny, c.Lx, c.Ly, c.par, d, c.bc, c.inflow)
    fdx = W @ (frozen_faces(theta + EPS * dth).rhs(U, t) -
               frozen_faces(theta - EPS * dth).rhs(U, t)) / (2 * EPS)
    assert _rel(fdx, fdt) < 1e-8
    # single entries: corner, edge and interior nodes of a boundary element and
    # of an interior one, every variable
    scaleU, scaleT = np.abs(Ub).max(), np.abs(thb).max()
    for e in (0, c.Nx + 1, c.Nx * c.Ny - 1):
        for loc in (0, 2, 12, 24):
            i = 25 * e + loc
            for v in range(5):
                d = np.zeros_like(U)
                d[v * c.nd + i] = 1.0
                fdi = W @ (o.rhs(U + EPS * d, t) - o.rhs(U - EPS * d, t)) / (2 * EPS)
                got = Ub[v * c.nd + i]
                assert abs(got - fdi) <= TOL * abs(fdi) + 1e-9 * scaleU, (c.name, e, loc, v, got, fdi)
            d = np.zeros(c.nd)
            d[i] = 1.0
            fdi = W @ (_oracle(c, theta + EPS * d).rhs(U, t) -
                       _oracle(c, theta - EPS * d).rhs(U, t)) / (2 * EPS)
            assert abs(thb[i] - fdi) <= TOL * abs(fdi) + 1e-9 * scaleT, (c.name, e, loc, thb[i], fdi)
    s.close()


def _taped_run(c, theta, U0, nodes, steps, max_snap, keep=False):
    import sgnx
    s = sgnx.SGN2D(**c.solver_kwargs())
    s.setDepth(theta)
    s.upload(U0)
    s.setGauges(nodes, 4 * steps)
    dt0 = s.cflTimeStep(None, 0.5)
    s.tapeBegin(max_snap)
    s.tapeKeepStages(keep)
    st = s.integrateAdaptive(0.0, 1e9, dt0, sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=steps))
    s.tapeEnd()
    assert st.accepted == steps
    return s, s.download(), s.gaugeLog()[1]


@pytest.mark.parametrize("ci", range(4))
def test_adjoint_of_adaptive_run_matches_finite_differences(ci):
    c = _cases()[ci]
    rng = np.random.default_rng(200 + ci)
    theta = c.depth["Hnod"] + 0.002 * rng.uniform(-1, 1, c.nd)
    U0 = c.U0 + 1e-3 * rng.uniform(-1, 1, c.U0.size) * np.repeat([1, 0.5, 0.5, 0.2, 0.2], c.nd)
    nodes = rng.choice(c.nd, size=7, replace=False).astype(np.int64)
    steps = 6
    s, UN, G = _taped_run(c, theta, U0, nodes, steps, 1 << 20)
    Ubar = rng.normal(size=UN.size)
    Gbar = rng.normal(size=G.shape)
    U0b, thb = s.adjoint(Ubar, Gbar)
    tt, dts = s.tapeSteps()
    assert tt.size == steps
    o0 = _oracle(c, theta)

    def loss(U0x, thx, o=None):
        o = o or _oracle(c, thx)
        U = U0x.copy()
        rows = []
        for t, dt in zip(tt, dts):
            U = o.bs3_step(U, t, dt)
            rows.append(U[nodes] - thx[nodes])
        return U, np.array(rows), Ubar @ U + np.sum(Gbar * np.array(rows))

    # the oracle's replay of the taped steps is the recorded run, bit for bit
    Ur, Gr, _ = loss(U0, theta, o0)
    assert_bitwise(UN, Ur, f"{c.name}: oracle replay of the tape")
    assert_bitwise(G, Gr, f"{c.name}: oracle replay gauges")
    # and so is the device replay
    s.upload(U0)
    s.setGauges(nodes, 4 * steps)
    s.tapeReplay()
    assert_bitwise(s.download(), UN, f"{c.name}: device replay")
    assert_bitwise(s.gaugeLog()[1], G, f"{c.name}: device replay gauges")

    dU = rng.uniform(-1.0, 1.0, U0.size) * np.repeat([0.05, 0.05, 0.05, 0.02, 0.02], c.nd)
    fd = (loss(U0 + EPS * dU, theta, o0)[2] - loss(U0 - EPS * dU, theta, o0)[2]) / (2 * EPS)
    assert _rel(U0b @ dU, fd) < TOL, (c.name, U0b @ dU, fd)
    dth = rng.uniform(-1.0, 1.0, c.nd) * 0.05
    fdt = (loss(U0, theta + EPS * dth)[2] - loss(U0, theta - EPS * dth)[2]) / (2 * EPS)
    assert _rel(thb @ dth, fdt) < TOL, (c.name, thb @ dth, fdt)
    # the gauges' direct theta term (eta = h - theta) is part of it
    direct = -np.bincount(nodes, weights=Gbar.sum(axis=0), minlength=c.nd)
    assert _rel((thb - direct) @ dth, fdt) > 1e3 * TOL
    s.close()


def test_repeated_gauge_nodes_in_the_adjoint():
    """Gauges may share a node: each gauge's cotangent reaches that node once,
    in gauge order (the gauge kernel sums a node's gauges on one thread). The
    gradient matches central differences of the oracle's replay with every
    gauge counted, duplicates included, and one more step through the
    batched sweep gives the same bits."""
    import sgnx
    c = _cases()[2]
    rng = np.random.default_rng(515)
    theta = c.depth["Hnod"] + 0.002 * rng.uniform(-1, 1, c.nd)
    U0 = c.U0 + 1e-3 * rng.uniform(-1, 1, c.U0.size) * np.repeat([1, 0.5, 0.5, 0.2, 0.2], c.nd)
    a, b_, d = rng.choice(c.nd, size=3, replace=False)
    nodes = np.array([a, b_, a, d, a, b_, d, a], dtype=np.int64)
    steps = 4
    s, UN, G = _taped_run(c, theta, U0, nodes, steps, 1 << 20)
    Ubar = rng.normal(size=UN.size)
    Gbar = rng.normal(size=G.shape)
    U0b, thb = s.adjoint(Ubar, Gbar)
    tt, dts = s.tapeSteps()

    def loss(U0x, thx):
        o = _oracle(c, thx)
        U = U0x.copy()
        rows = []
        for t, dt in zip(tt, dts):
            U = o.bs3_step(U, t, dt)
            rows.append(U[nodes] - thx[nodes])
        o.close()
        return Ubar @ U + np.sum(Gbar * np.array(rows))

    dU = rng.uniform(-1.0, 1.0, U0.size) * np.repeat([0.05, 0.05, 0.05, 0.02, 0.02], c.nd)
    fd = (loss(U0 + EPS * dU, theta) - loss(U0 - EPS * dU, theta)) / (2 * EPS)
    assert _rel(U0b @ dU, fd) < TOL, (U0b @ dU, fd)
    dth = rng.uniform(-1.0, 1.0, c.nd) * 0.05
    fdt = (loss(U0, theta + EPS * dth) - loss(U0, theta - EPS * dth)) / (2 * EPS)
    assert _rel(thb @ dth, fdt) < TOL, (thb @ dth, fdt)
    bt = sgnx.Batch([s])
    (U0b2, thb2), = bt.adjoint([Ubar], [Gbar])
    assert_bitwise(U0b2, U0b, "U0bar, batched sweep, repeated gauge nodes")
    assert_bitwise(thb2, thb, "thetaBar, batched sweep, repeated gauge nodes")
    bt.close()
    s.close()


def _recompute_cost(L, c):
    """Forward steps an optimal schedule needs to reverse L steps from a held
    state with c free slots (dynamic programming over the first checkpoint)."""
    from functools import lru_cache

    @lru_cache(None)
    def F(L, c):
        if L <= 1:
            return 0
        if c <= 0:
            return L * (L - 1) // 2
        return min(m + F(L - m, c - 1) + F(m, c) for m in range(1, L))
    return F(L, c)


@pytest.mark.parametrize("budget,slots", [(2, -1), (1, -1), (2, 0), (2, 1), (3, 2), (1, 3)])
def test_checkpoint_budgets_give_the_same_adjoint(budget, slots):
    """Snapshot budgets below the step count (stride grows, segments
    recomputed) and slot budgets below a segment's length (binomial schedule)
    give the store-all cotangents bit for bit, recomputing exactly as many
    steps as an optimal schedule needs."""
    c = _cases()[2]
    theta = c.depth["Hnod"].copy()
    nodes = np.array([5, 77, 300], dtype=np.int64)
    res = []
    for bud, sl in ((1 << 20, -1), (budget, slots)):
        s, UN, G = _taped_run(c, theta, c.U0, nodes, 13, bud)
        s.tapeSlots(sl)
        steps, stride, held = s.tapeInfo()
        assert steps == 13 and held <= bud
        r = np.random.default_rng(9)
        res.append(s.adjoint(r.normal(size=UN.size), r.normal(size=G.shape)))
        recomputed, peak = s.adjointStats()
        if sl >= 0:
            assert peak <= sl
        segs = [min(stride, 13 - j * stride) for j in range((13 + stride - 1) // stride)]
        want = sum(_recompute_cost(L, L - 1 if sl < 0 else sl) for L in segs)
        assert recomputed == want, (bud, sl, stride, recomputed, want)
        s.close()
    assert_bitwise(res[1][0], res[0][0], "U0bar under a checkpoint budget")
    assert_bitwise(res[1][1], res[0][1], "thetaBar under a checkpoint budget")


@pytest.mark.parametrize("plan,states,run", [(13, 3, 13), (13, 4, 13), (13, 6, 13), (13, 13, 13),
                                              (13, 4, 9), (9, 3, 13)])
def test_planned_tape_recomputes_the_fewest_steps(plan, states, run):
    """A tape planned for `plan` steps with `states` held states (snapshots at
    the binomial schedule's first sweep) gives the store-all cotangents bit for
    bit. When the run takes the planned steps, the sweep recomputes exactly
    F(L, states - 1) - (L - 1) steps, F the optimum of the dynamic programme
    over the first checkpoint (the taped run is the schedule's first sweep);
    a shorter or longer run is still reversed exactly."""
    import sgnx
    c = _cases()[2]
    theta = c.depth["Hnod"].copy()
    nodes = np.array([5, 77, 300], dtype=np.int64)
    r = np.random.default_rng(11)
    res = []
    for planned in (False, True):
        s = sgnx.SGN2D(**c.solver_kwargs())
        s.setDepth(theta)
        s.upload(c.U0)
        s.setGauges(nodes, 4 * run)
        dt0 = s.cflTimeStep(None, 0.5)
        if planned:
            s.tapeBeginPlan(plan, states)
        else:
            s.tapeBegin(1 << 20)
        st = s.integrateAdaptive(0.0, 1e9, dt0,
                                 sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=run))
        s.tapeEnd()
        assert st.accepted == run
        UN, G = s.download(), s.gaugeLog()[1]
        steps, stride, held = s.tapeInfo()
        if planned:
            assert steps == run and stride == 0 and held <= states
        rr = np.random.default_rng(9)
        res.append(s.adjoint(rr.normal(size=UN.size), rr.normal(size=G.shape)))
        recomputed, peak = s.adjointStats()
        if planned:
            assert peak <= states - 1, (peak, states)
            if run == plan:
                want = _recompute_cost(plan, states - 1) - (plan - 1)
                assert recomputed == want, (plan, states, recomputed, want)
        s.close()
    assert_bitwise(res[1][0], res[0][0], "U0bar of the planned tape")
    assert_bitwise(res[1][1], res[0][1], "thetaBar of the planned tape")


@pytest.mark.parametrize("ci", range(4))
def test_kept_stages_give_the_same_adjoint(ci):
    """A taped run that keeps its stage inputs (written by the forward's own
    stage launches, rejected attempts overwritten by the accepted one) runs
    the same steps to the same state and gauge log, and its reverse sweep,
    which recomputes no stage, gives the recomputing sweep's cotangents bit
    for bit; under a snapshot budget the run drops them and still agrees."""
    c = _cases()[ci]
    rng = np.random.default_rng(600 + ci)
    theta = c.depth["Hnod"] + 0.002 * rng.uniform(-1, 1, c.nd)
    nodes = rng.choice(c.nd, size=5, replace=False).astype(np.int64)
    steps = 9
    out = []
    for bud, keep in ((1 << 20, False), (1 << 20, True), (3, True)):
        s, UN, G = _taped_run(c, theta, c.U0, nodes, steps, bud, keep)
        assert s.tapeStages() == (steps if keep and bud > steps else 0)
        r = np.random.default_rng(1)
        out.append((UN, G, s.tapeSteps(), s.adjoint(r.normal(size=UN.size), r.normal(size=G.shape))))
        s.close()
    (U0, G0, (t0, d0), (ub0, tb0)) = out[0]
    for tag, (UN, G, (tt, dts), (ub, tb)) in zip(("kept", "dropped"), out[1:]):
        assert_bitwise(UN, U0, f"{tag}: final state")
        assert_bitwise(G, G0, f"{tag}: gauge log")
        assert_bitwise(tt, t0, f"{tag}: step times")
        assert_bitwise(dts, d0, f"{tag}: step sizes")
        assert_bitwise(ub, ub0, f"{tag}: U0bar")
        assert_bitwise(tb, tb0, f"{tag}: thetaBar")


@pytest.mark.parametrize("ci,keep", [(0, True), (3, True), (3, False)])
def test_chained_step_equals_single_products_bitwise(ci, keep):
    """The reverse step chains its three products: the first two launches also
    write the face coefficients of the next product's cotangent, and the next
    launch skips the coefficient kernel. One reversed step equals the same
    three products evaluated one at a time through sgnx_rhs_vjp, with the
    cotangents and sums formed on the host in the kernels' order, bit for bit
    (stage inputs kept by the taped run, or recomputed by the sweep)."""
    c = _cases()[ci]
    rng = np.random.default_rng(900 + ci)
    theta = c.depth["Hnod"] + 0.002 * rng.uniform(-1, 1, c.nd)
    U0 = c.U0 + 1e-3 * rng.uniform(-1, 1, c.U0.size) * np.repeat([1, 0.5, 0.5, 0.2, 0.2], c.nd)
    nodes = rng.choice(c.nd, size=3, replace=False).astype(np.int64)
    s, UN, _ = _taped_run(c, theta, U0, nodes, 1, 1 << 20, keep)
    lam = rng.normal(size=UN.size)
    U0b, thb = s.adjoint(lam, None)
    (t,), (dt,) = s.tapeSteps()
    # the step's stage inputs (rk.cpp:99-104), formed as the fused launches form them
    k1 = s.rhs(U0, t)
    y2 = U0 + (0.5 * dt) * k1
    k2 = s.rhs(y2, t + 0.5 * dt)
    y3 = U0 + (0.75 * dt) * k2
    y3b, th3 = s.rhs_vjp(y3, (dt * (4.0 / 9.0)) * lam, t + 0.75 * dt)
    y2b, th2 = s.rhs_vjp(y2, (dt * (1.0 / 3.0)) * lam + (0.75 * dt) * y3b, t + 0.5 * dt)
    p1, th1 = s.rhs_vjp(U0, (dt * (2.0 / 9.0)) * lam + (0.5 * dt) * y2b, t)
    s.close()
    assert_bitwise(U0b, ((lam + y3b) + y2b) + p1, "U0bar of one chained step")
    assert_bitwise(thb, ((0.0 + th3) + th2) + th1, "thetaBar of one chained step")


@pytest.mark.parametrize("budget,keep", [(1 << 20, False), (3, False), (1 << 20, True)])
def test_taped_run_leaves_the_solver_intact(budget, keep):
    """Snapshots take over the buffer of each accepted step's input instead of
    copying it: a taped run, its reverse sweep and a further untaped run give
    the states, step counts and gauge logs of the same calls on an untaped
    solver, bit for bit."""
    import sgnx
    c = _cases()[3]
    nodes = np.array([3, 140, 611], dtype=np.int64)
    opts = sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=9)
    out = []
    for taped in (True, False):
        s = sgnx.SGN2D(**c.solver_kwargs())
        s.setDepth(c.depth["Hnod"])
        s.upload(c.U0)
        s.setGauges(nodes, 64)
        dt0 = s.cflTimeStep(None, 0.5)
        if taped:
            s.tapeKeepStages(keep)
            s.tapeBegin(budget)
        st1 = s.integrateAdaptive(0.0, 1e9, dt0, opts)
        U1 = s.download()
        if taped:
            s.tapeEnd()
            s.adjoint(np.ones(U1.size), None)
        st2 = s.integrateAdaptive(st1.tReached, 1e9, st1.dtLast, opts)
        out.append((st1.as_dict(), st2.as_dict(), U1, s.download(), s.gaugeLog()[1]))
        s.close()
    a, b = out
    assert a[0] == b[0] and a[1] == b[1]
    for k, what in ((2, "state after the taped run"), (3, "state after the next run"),
                    (4, "gauge log")):
        assert_bitwise(a[k], b[k], what)


def _plain_rhs(s):
    def rhs(U, th, t):
        s.setDepth(th)
        return s.rhs(U, t)
    return rhs


@pytest.mark.parametrize("ci", range(5))
def test_adjoint_matches_tangent_linear_model(ci):
    """The adjoint's directional derivatives <U0bar, dU0> + <thetaBar, dtheta>
    against the tangent-linear reference computed with the plain solver's RHS,
    in three seeded random (rough) directions of the initial state and the
    still depth: the four small cases over 6 adaptive steps, and the
    gauge-inversion workload at 48x24 over 60 steps with its 20 gauges."""
    import sgnx
    from sgnx import cases
    if ci < 4:
        c = _cases()[ci]
        steps = 6
        rng = np.random.default_rng(800 + ci)
        theta = c.depth["Hnod"] + 0.002 * rng.uniform(-1, 1, c.nd)
        U0 = c.U0 + 1e-3 * rng.uniform(-1, 1, c.U0.size) * np.repeat([1, 0.5, 0.5, 0.2, 0.2], c.nd)
        nodes = rng.choice(c.nd, size=6, replace=False).astype(np.int64)
    else:
        c = cases.inversion(seed=6, Nx=48, Ny=24)
        steps = 60
        rng = np.random.default_rng(806)
        theta, U0, nodes = c.meta["theta"], c.U0, c.meta["gauges"]
    s, UN, G = _taped_run(c, theta, U0, nodes, steps, 1 << 20, keep=True)
    Ubar = rng.normal(size=UN.size)
    Gbar = rng.normal(size=G.shape)
    U0b, thb = s.adjoint(Ubar, Gbar)
    tt, dts = s.tapeSteps()
    rhs = _plain_rhs(s)
    for trial in range(3):
        dU = rng.normal(size=U0.size)
        dth = rng.normal(size=c.nd)
        ad = U0b @ dU + thb @ dth
        tl = tangent_linear(rhs, U0, theta, dU, dth, list(zip(tt, dts)), nodes, Ubar, Gbar)
        err = _rel(ad, tl)
        print(f"{c.name} direction {trial}: adjoint {ad:.12e} tangent-linear {tl:.12e} rel {err:.1e}")
        assert err <= TOL, (ci, trial, ad, tl)
    s.close()


def test_inversion_gradient_32_directions():
    """The gauge-inversion workload (smaller grid and run): the gradient of
    the gauge misfit 0.5 |G(eta0, theta) - G_obs|^2 with respect to the initial
    surface and the still depth, through 150 adaptive steps under 8 snapshots
    and 8 slots, against central differences of the plain solver's replay of
    the taped steps (fourth-order, eps 3e-3) in 32 seeded smooth directions near the
    gauges, and in one of them against the oracle's replay. (With the step sizes held fixed, a rough direction
    excites grid modes the recorded steps do not resolve stably; the
    controller's rejections guarded the recorded run against them.)"""
    import sgnx
    from sgnx import cases
    c = cases.inversion(seed=4, Nx=64, Ny=32)
    theta, eta0, nodes = c.meta["theta"], c.meta["eta0"], c.meta["gauges"]
    nd = c.nd
    rest = c.U0[nd:]
    steps = 150
    s = sgnx.SGN2D(**c.solver_kwargs())
    s.setDepth(theta)
    s.upload(c.U0)
    dt0 = s.cflTimeStep(None, 0.5)
    opts = sgnx.RkOpts(atol=c.meta["atol"], rtol=c.meta["rtol"], max_accepted=steps)
    # observations: the same run from a perturbed surface
    s.upload(np.concatenate([theta + 1.1 * eta0, rest]))
    s.setGauges(nodes, steps)
    s.integrateAdaptive(0.0, 1e9, dt0, opts)
    G_obs = s.gaugeLog()[1]
    s.upload(c.U0)
    s.setGauges(nodes, steps)
    s.tapeBegin(8)
    st = s.integrateAdaptive(0.0, 1e9, dt0, opts)
    s.tapeEnd()
    s.tapeSlots(8)
    assert st.accepted == steps
    G = s.gaugeLog()[1]
    U0b, thb = s.adjoint(None, G - G_obs)
    assert s.tapeInfo()[2] <= 8 and s.adjointStats()[1] <= 8
    g_eta = U0b[:nd]                      # u0 = (theta + eta0, 0, ...)
    g_theta = thb + U0b[:nd]

    def misfit(e, th):
        s.setDepth(th)
        s.upload(np.concatenate([th + e, rest]))
        s.setGauges(nodes, steps)
        s.tapeReplay()
        return 0.5 * np.sum((s.gaugeLog()[1] - G_obs) ** 2)

    # The differenced runs carry rounding noise (about 1e-14 in the misfit's
    # derivative) that the held step sizes amplify; a fourth-order difference
    # at eps 3e-3 keeps both that noise and the truncation small. The error is
    # taken against |eta part| + |theta part| of the directional derivative:
    # the two parts can cancel to a net value near the noise.
    rng = np.random.default_rng(32)
    eps = 3e-3
    worst = 0.0
    for k in range(32):
        de = cases.smooth_direction(c, rng, near=nodes)
        dth = cases.smooth_direction(c, rng, near=nodes)
        f = {k: misfit(eta0 + k * eps * de, theta + k * eps * dth) for k in (-2, -1, 1, 2)}
        fd = (8.0 * (f[1] - f[-1]) - (f[2] - f[-2])) / (12.0 * eps)
        pe, pt = g_eta @ de, g_theta @ dth
        worst = max(worst, abs(pe + pt - fd) / (abs(pe) + abs(pt)))
    assert worst < TOL, worst
    # one direction against the oracle's replay of the same steps
    tt, dts = s.tapeSteps()

    def misfit_oracle(e, th):
        o = _oracle(c, th)
        U = np.concatenate([th + e, rest])
        rows = []
        for t, dt in zip(tt, dts):
            U = o.bs3_step(U, t, dt)
            rows.append(U[nodes] - th[nodes])
        o.close()
        return 0.5 * np.sum((np.array(rows) - G_obs) ** 2)
    fdo = (misfit_oracle(eta0 + eps * de, theta + eps * dth) -
           misfit_oracle(eta0 - eps * de, theta - eps * dth)) / (2 * eps)
    assert _rel(g_eta @ de + g_theta @ dth, fdo) < TOL
    s.close()


def test_autograd_function_matches_finite_differences():
    """torch.autograd over (eta0, theta) -> (final state, gauges), with
    u0 = (theta + eta0, hu, hv, hw, hp) composed in torch: the gradients of a
    loss of both outputs equal central differences of the taped step sequence
    (replay), and a second backward through a replay gives the same gradients."""
    torch = pytest.importorskip("torch")
    import sgnx
    from sgnx.autograd import replay, run_adaptive
    c = _cases()[0]
    rng = np.random.default_rng(11)
    theta0 = c.depth["Hnod"].copy()
    eta0 = c.U0[:c.nd] - theta0
    rest = torch.tensor(c.U0[c.nd:].copy())
    nodes = rng.choice(c.nd, size=5, replace=False).astype(np.int64)
    s = sgnx.SGN2D(**c.solver_kwargs())
    s.setDepth(theta0)
    s.upload(c.U0)
    dt0 = s.cflTimeStep(None, 0.5)
    opts = sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=5)
    wg = torch.tensor(rng.normal(size=(5, nodes.size)))

    def loss(e, th, taped):
        U0 = torch.cat([th + e, rest])
        if taped:
            UN, G = run_adaptive(s, U0, th, 0.0, 1e9, dt0, opts, nodes)
        else:
            UN, G = replay(s, U0, th, nodes)
        return (wg * G).sum() + 0.5 * (UN[:c.nd] ** 2).sum()

    def grads(taped):
        e = torch.tensor(eta0, requires_grad=True)
        th = torch.tensor(theta0, requires_grad=True)
        L = loss(e, th, taped)
        L.backward()
        return L.item(), e.grad.numpy().copy(), th.grad.numpy().copy()

    L0, ge, gt = grads(True)
    assert s.tapeInfo()[0] == 5
    L1, ge1, gt1 = grads(False)
    assert L1 == L0
    assert_bitwise(ge1, ge, "eta0 gradient through a replay")
    assert_bitwise(gt1, gt, "theta gradient through a replay")
    de = rng.uniform(-1, 1, c.nd) * 0.01
    dth = rng.uniform(-1, 1, c.nd) * 0.01
    with torch.no_grad():
        T = lambda a: torch.tensor(a)
        fde = (loss(T(eta0 + EPS * de), T(theta0), False).item() -
               loss(T(eta0 - EPS * de), T(theta0), False).item()) / (2 * EPS)
        fdt = (loss(T(eta0), T(theta0 + EPS * dth), False).item() -
               loss(T(eta0), T(theta0 - EPS * dth), False).item()) / (2 * EPS)
    assert _rel(ge @ de, fde) < TOL, (ge @ de, fde)
    assert _rel(gt @ dth, fdt) < TOL, (gt @ dth, fdt)
    s.close()


def test_batched_adjoint_matches_solo_bitwise():
    """Members of uneven sizes and step counts (periodic, walls, inflow, a
    partial strip), each with its own tape and gauges: the batched reverse
    sweep gives every member's solo cotangents bit for bit, and a tape that
    lacks a step's snapshot is refused."""
    import dataclasses
    import sgnx
    from sgnx import cases
    cs = [cases.gate(Nx=8, Ny=6), cases.solitary(Nx=32, Ny=16), cases.shoal(seed=3, Nx=48, Ny=24),
          dataclasses.replace(cases.gate(Nx=37, Ny=11, bc=("periodic", "periodic", "wall", "wall")),
                              name="gate37x11")]
    steps = [7, 3, 5, 9]
    rng = np.random.default_rng(77)
    ss, cot, solo = [], [], []
    for c, k in zip(cs, steps):
        theta = c.depth["Hnod"] if c.depth is not None else np.zeros(c.nd)
        nodes = rng.choice(c.nd, size=4, replace=False).astype(np.int64)
        s, UN, G = _taped_run(c, theta, c.U0, nodes, k, 1 << 20)
        ub, gb = rng.normal(size=UN.size), rng.normal(size=G.shape)
        solo.append(s.adjoint(ub, gb))
        cot.append((ub, gb))
        ss.append(s)
    b = sgnx.Batch(ss)
    res = b.adjoint([u for u, _ in cot], [g for _, g in cot])
    for i, ((U0b, thb), (sU, sT)) in enumerate(zip(res, solo)):
        assert_bitwise(U0b, sU, f"member {i} ({cs[i].name}) U0bar")
        assert_bitwise(thb, sT, f"member {i} ({cs[i].name}) thetaBar")
    # a member whose tape was thinned is refused
    c = cs[0]
    ss[0].upload(c.U0)
    ss[0].tapeBegin(2)
    ss[0].integrateAdaptive(0.0, 1e9, ss[0].cflTimeStep(None, 0.5),
                            sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=6))
    ss[0].tapeEnd()
    with pytest.raises(sgnx.SgnxError, match="every step"):
        b.adjoint()
    b.close()
    for s in ss:
        s.close()


def test_batched_adjoint_of_sweep_members_matches_oracle_differences():
    """The batched sweep on members of the production sweep (seeded shoal,
    per-point bathymetry noise, inflow and walls, uneven sizes and step
    counts), one tape without kept stages and one with them: each member's
    gradient of <Ubar, U_N> + <Gbar, G> against central differences of the
    oracle's replay of its own taped steps, along a random direction in the
    initial state and one in the still depth."""
    import sgnx
    from sgnx import cases
    cs = [cases.sweep_member(11, 20), cases.sweep_member(12, 32)]
    steps = [5, 4]
    keeps = [False, True]
    rng = np.random.default_rng(404)
    ss, runs = [], []
    for c, k, keep in zip(cs, steps, keeps):
        theta = c.depth["Hnod"].copy()
        U0 = c.U0 + 1e-3 * rng.uniform(-1, 1, c.U0.size) * np.repeat([1, 0.5, 0.5, 0.2, 0.2], c.nd)
        nodes = rng.choice(c.nd, size=5, replace=False).astype(np.int64)
        s, UN, G = _taped_run(c, theta, U0, nodes, k, 1 << 20, keep)
        assert s.tapeStages() == (k if keep else 0)
        ub, gb = rng.normal(size=UN.size), rng.normal(size=G.shape)
        runs.append((c, theta, U0, nodes, s.tapeSteps(), ub, gb))
        ss.append(s)
    b = sgnx.Batch(ss)
    res = b.adjoint([r[5] for r in runs], [r[6] for r in runs])
    for (U0b, thb), (c, theta, U0, nodes, (tt, dts), ub, gb) in zip(res, runs):
        def loss(U0x, thx):
            o = _oracle(c, thx)
            U = U0x.copy()
            rows = []
            for t, dt in zip(tt, dts):
                U = o.bs3_step(U, t, dt)
                rows.append(U[nodes] - thx[nodes])
            o.close()
            return ub @ U + np.sum(gb * np.array(rows))

        dU = rng.uniform(-1.0, 1.0, U0.size) * np.repeat([0.05, 0.05, 0.05, 0.02, 0.02], c.nd)
        fd = (loss(U0 + EPS * dU, theta) - loss(U0 - EPS * dU, theta)) / (2 * EPS)
        assert _rel(U0b @ dU, fd) < TOL, (c.name, U0b @ dU, fd)
        dth = rng.uniform(-1.0, 1.0, c.nd) * 0.01
        fdt = (loss(U0, theta + EPS * dth) - loss(U0, theta - EPS * dth)) / (2 * EPS)
        assert _rel(thb @ dth, fdt) < TOL, (c.name, thb @ dth, fdt)
    b.close()
    for s in ss:
        s.close()


@pytest.mark.parametrize("keep", [(), (0, 2)])
def test_batched_forward_records_the_solo_tapes(keep):
    """Members recording tapes inside a batched adaptive run (graph on) get the
    tapes of their solo runs: the same steps, and batched and solo reverse
    sweeps give the same cotangents bit for bit -- also where some members
    keep their stage inputs (the batched stage launches write them; the
    batched sweep recomputes stages for the others only)."""
    import sgnx
    from sgnx import cases
    cs = [cases.gate(Nx=8, Ny=6), cases.solitary(Nx=32, Ny=16), cases.shoal(seed=3, Nx=48, Ny=24)]
    steps = [6, 4, 5]
    rng = np.random.default_rng(5)
    solo, ss, cot = [], [], []
    for c, k in zip(cs, steps):
        nodes = rng.choice(c.nd, size=3, replace=False).astype(np.int64)
        s, UN, G = _taped_run(c, c.depth["Hnod"], c.U0, nodes, k, 1 << 20)
        ub, gb = rng.normal(size=UN.size), rng.normal(size=G.shape)
        solo.append((s.tapeSteps(), s.adjoint(ub, gb)))
        cot.append((ub, gb, nodes))
        ss.append(s)
    for i, (c, s, (_, _, nodes)) in enumerate(zip(cs, ss, cot)):
        s.upload(c.U0)
        s.setGauges(nodes, 64)
        s.tapeBegin(1 << 20)
        s.tapeKeepStages(i in keep)
    b = sgnx.Batch(ss)
    dt0 = [s.cflTimeStep(None, 0.5) for s in ss]
    for c, s in zip(cs, ss):
        s.upload(c.U0)
    opts = [sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=k) for k in steps]
    b.integrateAdaptive(0.0, 1e9, dt0, opts, graph=True)
    for s in ss:
        s.tapeEnd()
    for i, s in enumerate(ss):
        tt, dts = s.tapeSteps()
        assert s.tapeStages() == (steps[i] if i in keep else 0)
        assert_bitwise(tt, solo[i][0][0], f"member {i} step times")
        assert_bitwise(dts, solo[i][0][1], f"member {i} step sizes")
    res = b.adjoint([u for u, _, _ in cot], [g for _, g, _ in cot])
    for i, (U0b, thb) in enumerate(res):
        assert_bitwise(U0b, solo[i][1][0], f"member {i} U0bar")
        assert_bitwise(thb, solo[i][1][1], f"member {i} thetaBar")
    b.close()
    for s in ss:
        s.close()


def test_batched_forward_records_planned_tapes():
    """Members whose tapes are planned (binomial snapshots for their step
    counts, fewer states than steps) record them inside a batched adaptive run
    (graph on) as they would solo, and each member's own reverse sweep gives
    its store-all cotangents bit for bit."""
    import sgnx
    from sgnx import cases
    cs = [cases.gate(Nx=8, Ny=6), cases.solitary(Nx=32, Ny=16), cases.shoal(seed=3, Nx=48, Ny=24)]
    steps, states = [9, 7, 8], [3, 4, 2]
    rng = np.random.default_rng(7)
    want, ss, cot = [], [], []
    for c, k in zip(cs, steps):
        nodes = rng.choice(c.nd, size=3, replace=False).astype(np.int64)
        s, UN, G = _taped_run(c, c.depth["Hnod"], c.U0, nodes, k, 1 << 20)
        ub, gb = rng.normal(size=UN.size), rng.normal(size=G.shape)
        want.append(s.adjoint(ub, gb))
        cot.append((ub, gb, nodes))
        ss.append(s)
    for c, s, (_, _, nodes), k, m in zip(cs, ss, cot, steps, states):
        s.upload(c.U0)
        s.setGauges(nodes, 64)
        s.tapeBeginPlan(k, m)
    b = sgnx.Batch(ss)
    dt0 = [s.cflTimeStep(None, 0.5) for s in ss]
    for c, s in zip(cs, ss):
        s.upload(c.U0)
    opts = [sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=k) for k in steps]
    b.integrateAdaptive(0.0, 1e9, dt0, opts, graph=True)
    for s in ss:
        s.tapeEnd()
    b.close()
    for i, (s, (ub, gb, _)) in enumerate(zip(ss, cot)):
        n, stride, held = s.tapeInfo()
        assert n == steps[i] and stride == 0 and held <= states[i]
        U0b, thb = s.adjoint(ub, gb)
        assert s.adjointStats()[1] <= states[i] - 1
        assert_bitwise(U0b, want[i][0], f"member {i} U0bar (planned tape)")
        assert_bitwise(thb, want[i][1], f"member {i} thetaBar (planned tape)")
        s.close()


def test_batched_adjoint_of_large_members_matches_solo_bitwise():
    """Members that each fill the GPU (at least six product workgroups per
    compute unit) run their own sweeps inside the batched call: every
    member's cotangents are its own sgnx_adjoint's bit for bit, and a
    thinned tape is refused as in the lockstep."""
    import sgnx
    from sgnx import cases
    cs = [cases.shoal(seed=5, Nx=160, Ny=80), cases.shoal(seed=6, Nx=170, Ny=86)]
    steps = [4, 3]
    rng = np.random.default_rng(78)
    ss, cot, solo = [], [], []
    for c, k in zip(cs, steps):
        nodes = rng.choice(c.nd, size=4, replace=False).astype(np.int64)
        s, UN, G = _taped_run(c, c.depth["Hnod"], c.U0, nodes, k, 1 << 20)
        ub, gb = rng.normal(size=UN.size), rng.normal(size=G.shape)
        solo.append(s.adjoint(ub, gb))
        cot.append((ub, gb))
        ss.append(s)
    b = sgnx.Batch(ss)
    res = b.adjoint([u for u, _ in cot], [g for _, g in cot])
    for i, ((U0b, thb), (sU, sT)) in enumerate(zip(res, solo)):
        assert_bitwise(U0b, sU, f"large member {i} U0bar")
        assert_bitwise(thb, sT, f"large member {i} thetaBar")
    c = cs[0]
    ss[0].upload(c.U0)
    ss[0].tapeBegin(2)
    ss[0].integrateAdaptive(0.0, 1e9, ss[0].cflTimeStep(None, 0.5),
                            sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=6))
    ss[0].tapeEnd()
    with pytest.raises(sgnx.SgnxError, match="every step"):
        b.adjoint()
    b.close()
    for s in ss:
        s.close()
