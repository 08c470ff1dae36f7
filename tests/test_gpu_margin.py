"""The large single-grid margin workload (cases.margin: 2000x1000 elements on
200x100 km, a tanh continental margin 4000 -> 50 m with 2 % seeded roughness,
walls in x, periodic in y, an elliptical uplift at rest) against the oracle:
the RHS bit for bit at reduced and at full size, an adaptive trajectory at
reduced size to the trajectory bar (same step counts, relative L2 <= 1e-12),
and mass conservation of an adaptive run at full size."""
import numpy as np
import pytest

import oracle_lib as OL
from test_gpu_rhs import assert_bitwise
from test_gpu_rk import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("Nx,Ny", [(40, 20), (37, 11)])
def test_margin_rhs_bitwise(Nx, Ny):
    """The initial state and a moving state (the uplift plus seeded currents),
    both kernel variants."""
    import sgnx
    from sgnx import cases
    c = cases.margin(seed=1, Nx=Nx, Ny=Ny)
    rng = np.random.default_rng(3)
    U = c.U0.copy()
    h = U[:c.nd]
    for v, amp in ((1, 0.3), (2, 0.2), (3, 0.05), (4, 0.5)):
        U[v * c.nd:(v + 1) * c.nd] = h * amp * rng.uniform(-1, 1, c.nd)
    o = OL.Oracle.from_case(c)
    s = sgnx.SGN2D(**c.solver_kwargs())
    for tag, X in (("initial", c.U0), ("moving", U)):
        want = o.rhs(X)
        for variant, rows in ((1, 16), (1, 3), (0, 16)):
            s.set_rhs_kernel(variant, rows)
            assert_bitwise(s.rhs(X), want, f"margin {Nx}x{Ny} {tag} variant {variant}/{rows}")
    s.close()
    o.close()


def test_margin_adaptive_matches_oracle():
    """40 accepted adaptive steps at 80x40 (atol 1e-8, rtol 1e-6): the same
    accepted and rejected counts as the oracle's run and a relative L2
    distance <= 1e-12."""
    import sgnx
    from sgnx import cases
    c = cases.margin(seed=2, Nx=80, Ny=40)
    o = OL.Oracle.from_case(c)
    dt0 = o.cfl_time_step(c.U0, 0.5)
    s = sgnx.SGN2D(**c.solver_kwargs())
    s.upload(c.U0)
    st = s.integrateAdaptive(0.0, 1e9, dt0, sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=40))
    Ug = s.download()
    s.close()
    Uo, sto, _, _, _ = o.integrate(c.U0, 0.0, 1e9, dt0,
                                   OL.OrcOpts(atol=1e-8, rtol=1e-6, max_accepted=40))
    o.close()
    assert (st.accepted, st.rejected, st.rhsEvals) == (sto.accepted, sto.rejected, sto.rhsEvals)
    assert st.accepted == 40
    assert rel_l2(Ug, Uo) <= 1e-12


@pytest.fixture(scope="module")
def margin_full():
    from sgnx import cases
    return cases.margin(seed=0)


def test_margin_fullsize_rhs_bitwise(margin_full):
    """2000x1000 (50 M nodes, 2 GB per state): one RHS of the initial state
    against the oracle, bit for bit."""
    import sgnx
    c = margin_full
    s = sgnx.SGN2D(**c.solver_kwargs())
    got = s.rhs(c.U0)
    s.close()
    o = OL.Oracle.from_case(c)
    want = o.rhs(c.U0)
    o.close()
    assert_bitwise(got, want, "margin 2000x1000")


def test_margin_fullsize_mass_conserved(margin_full):
    """Walls in x and a periodic y axis conserve mass: 5 adaptive steps at
    full size change it by round-off only; the run is finite and
    reproducible bit for bit."""
    import sgnx
    c = margin_full
    s = sgnx.SGN2D(**c.solver_kwargs())
    s.upload(c.U0)
    m0 = s.mass()
    dt0 = s.cflTimeStep(None, 0.5)
    opts = sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=5)
    st = s.integrateAdaptive(0.0, 1e9, dt0, opts)
    m1 = s.mass()
    U1 = s.download()
    s.upload(c.U0)
    st2 = s.integrateAdaptive(0.0, 1e9, dt0, opts)
    U2 = s.download()
    s.close()
    assert st.accepted == 5 and st.as_dict() == st2.as_dict()
    assert np.all(np.isfinite(U1))
    assert abs(m1 - m0) <= 1e-12 * abs(m0), (m0, m1)
    assert_bitwise(U2, U1, "margin rerun")


def test_margin_source_gradient_matches_finite_differences():
    """The margin's source-inversion gradient at 48x24: the gauge misfit
    0.5 |G(eta0) - G_obs|^2 over 40 adaptive steps (walls in x, periodic in
    y, 12 gauges inside the uplift) under 4 snapshots and 4 slots, the same
    bits as a tape holding every step, and within 1e-6 of fourth-order
    central differences (eps 3e-3) of the plain solver's replay of the taped
    steps in 6 seeded smooth directions (Gaussians 10-30 km wide)."""
    import sgnx
    from sgnx import cases
    c = cases.margin(seed=3, Nx=48, Ny=24)
    nd, steps = c.nd, 40
    Hn = c.depth["Hnod"]
    eta0, rest = c.U0[:nd] - Hn, c.U0[nd:]
    nodes = cases.margin_gauges(c, 12, seed=3)
    assert nodes.size == 12
    s = sgnx.SGN2D(**c.solver_kwargs())
    s.upload(c.U0)
    dt0 = s.cflTimeStep(None, 0.5)
    opts = sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=steps)
    s.upload(np.concatenate([Hn + 1.1 * eta0, rest]))
    s.setGauges(nodes, steps)
    s.integrateAdaptive(0.0, 1e9, dt0, opts)
    G_obs = s.gaugeLog()[1]
    grads = []
    for snaps, slots in ((1 << 20, -1), (4, 4)):
        s.upload(c.U0)
        s.setGauges(nodes, steps)
        s.tapeBegin(snaps)
        st = s.integrateAdaptive(0.0, 1e9, dt0, opts)
        s.tapeEnd()
        s.tapeSlots(slots)
        assert st.accepted == steps
        G = s.gaugeLog()[1]
        grads.append(s.adjoint(None, G - G_obs, theta=False)[:nd])
    assert s.tapeInfo()[2] <= 4
    assert_bitwise(grads[1], grads[0], "margin gradient: 4 snapshots + 4 slots vs every step")
    g_eta = grads[1]

    def misfit(e):
        s.upload(np.concatenate([Hn + e, rest]))
        s.setGauges(nodes, steps)
        s.tapeReplay()
        return 0.5 * np.sum((s.gaugeLog()[1] - G_obs) ** 2)

    rng = np.random.default_rng(11)
    eps, worst = 3e-3, 0.0
    for _ in range(6):
        de = cases.smooth_direction(c, rng, n=2, near=nodes, scale=5000.0)
        f = {k: misfit(eta0 + k * eps * de) for k in (-2, -1, 1, 2)}
        fd = (8.0 * (f[1] - f[-1]) - (f[2] - f[-2])) / (12.0 * eps)
        pe = g_eta @ de
        worst = max(worst, abs(pe - fd) / max(abs(pe), abs(fd)))
    s.close()
    assert worst < 1e-6, worst
