"""The benchmark workloads against the oracle at their own sizes.

  config 2  the shoal production share, 1024x512: two accepted adaptive steps
            from the initial state (the FSAL head and 6 stage RHS), the oracle's
            step counts and relative L2 <= 1e-12
  config 3  members of the batched shoal sweep (cases.sweep_member: seeded
            round shoal on the slope, per-point relative bathymetry noise,
            inflow on the left, walls elsewhere) at Nx 128, 134 and 256 (each
            with a partial last strip): the RHS bit for bit against the oracle;
            a batched adaptive run with the captured step graph against the
            oracle's runs (step counts, relative L2 <= 1e-12) and bit for bit
            against the members' solo runs
  config 4  the gauge-inversion workload at its 256x128: the RHS bit for bit
            and 12 adaptive steps against the oracle

The batched sweep of the reference is a set of separate jobs
(scripts/shoal_sweep.sbatch:52-53); each member here must match what one of
those jobs, restated by the oracle, computes. The gate pattern (GPU RHS against
the CPU rhs(), then an adaptive run against the CPU's) follows
test/test_cuda_sgn2d.cu:66-128.
"""
import numpy as np
import pytest

import oracle_lib as OL
from test_gpu_rhs import assert_bitwise

pytestmark = pytest.mark.gpu

SWEEP = ((11, 128), (12, 134), (13, 256))      # (member seed, Nx); Ny = Nx / 2


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _perturbed(case, seed):
    """A non-trivial state over the case's still depth: random primitives of
    10 % amplitude about a 0.3 m layer, shifted onto the bathymetry."""
    from sgnx import cases
    U = cases.random_state(case.nd, seed=seed, h0=0.3, amp=0.1)
    U[:case.nd] += case.depth["Hnod"] - 0.3
    return U


@pytest.fixture(scope="module")
def sweep_members():
    from sgnx import cases
    out = []
    for seed, nx in SWEEP:
        c = cases.sweep_member(seed, nx)
        assert c.meta["shoal"]["noise"] > 0.0          # the noise path is exercised
        assert nx % 5 != 0                             # a partial last strip
        out.append(c)
    return out


def test_sweep_members_rhs_bitwise(sweep_members):
    import sgnx
    for i, c in enumerate(sweep_members):
        U = _perturbed(c, 100 + i)
        t = 0.13 + 0.2 * i                             # the inflow ghost's time
        o = OL.Oracle.from_case(c)
        want = o.rhs(U, t)
        o.close()
        s = sgnx.SGN2D(**c.solver_kwargs())
        for rows in (16, 3):
            s.set_rhs_kernel(1, rows)
            assert_bitwise(s.rhs(U, t), want, f"{c.name} {c.Nx}x{c.Ny} rows {rows}")
        s.close()


def test_sweep_batch_adaptive_matches_oracle_and_solo(sweep_members):
    """12 accepted steps of every member in one batch (captured step graph):
    the oracle's step counts and final state to 1e-12, and the member's own
    solo run bit for bit."""
    import sgnx
    steps = 12
    ss = [sgnx.SGN2D(**c.solver_kwargs()) for c in sweep_members]
    dt0 = []
    for c, s in zip(sweep_members, ss):
        s.upload(c.U0)
        dt0.append(s.cflTimeStep(None, 0.5))
    opts = [sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=steps) for _ in ss]
    b = sgnx.Batch(ss)
    sts = b.integrateAdaptive(0.0, 1e9, dt0, opts, graph=True)
    batched = [s.download() for s in ss]
    for i, (c, s, st) in enumerate(zip(sweep_members, ss, sts)):
        assert st.accepted == steps and not st.aborted
        o = OL.Oracle.from_case(c)
        Uo, sto, _, _, _ = o.integrate(c.U0, 0.0, 1e9, dt0[i],
                                       OL.OrcOpts(atol=1e-8, rtol=1e-6, max_accepted=steps))
        o.close()
        assert (st.accepted, st.rejected, st.rhsEvals) == \
            (sto.accepted, sto.rejected, sto.rhsEvals), (c.name, st.as_dict(), sto.as_dict())
        assert st.tReached == pytest.approx(sto.tReached, rel=1e-12)
        err = rel_l2(batched[i], Uo)
        assert err <= 1e-12, (c.name, err)
        # the member alone, same tile height: the same bits
        s.upload(c.U0)
        sts_solo = s.integrateAdaptive(0.0, 1e9, dt0[i], opts[i])
        assert sts_solo.as_dict() == st.as_dict()
        assert_bitwise(s.download(), batched[i], f"{c.name} batched vs solo")
    b.close()
    for s in ss:
        s.close()


def test_batch_graph_recaptures_when_a_member_changes_tile_height():
    """A captured step bakes in the launch grid; after a member's tile height
    changes (set_rhs_kernel), the next graph run must capture again. Every
    member of both runs equals its solo run under the same tile heights."""
    import sgnx
    from sgnx import cases
    cs = [cases.sweep_member(21, 40), cases.sweep_member(22, 64)]
    ss = [sgnx.SGN2D(**c.solver_kwargs()) for c in cs]
    dt0 = []
    for c, s in zip(cs, ss):
        s.upload(c.U0)
        dt0.append(s.cflTimeStep(None, 0.5))
    opts = [sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=k) for k in (5, 7)]
    b = sgnx.Batch(ss)
    for rows in ((16, 16), (2, 5)):
        for s, r in zip(ss, rows):
            s.set_rhs_kernel(1, r)
        solo = []
        for c, s, d, o in zip(cs, ss, dt0, opts):
            s.upload(c.U0)
            st = s.integrateAdaptive(0.0, 1e9, d, o)
            solo.append((s.download(), st.as_dict()))
            s.upload(c.U0)
        sts = b.integrateAdaptive(0.0, 1e9, dt0, opts, graph=True)
        for i, (s, st) in enumerate(zip(ss, sts)):
            assert st.as_dict() == solo[i][1], (rows, i)
            assert_bitwise(s.download(), solo[i][0], f"member {i} rows {rows}")
    b.close()
    for s in ss:
        s.close()


def test_shoal_fullsize_two_steps_match_oracle():
    """Config 2 at 1024x512: the adaptive run's first two accepted steps from
    the initial state against the oracle's (about 15 s of oracle time)."""
    import sgnx
    from sgnx import cases
    c = cases.shoal(seed=0)
    s = sgnx.SGN2D(**c.solver_kwargs())
    s.upload(c.U0)
    dt0 = s.cflTimeStep(None, 0.5)
    st = s.integrateAdaptive(0.0, 1e9, dt0, sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=2))
    Ug = s.download()
    s.close()
    o = OL.Oracle.from_case(c)
    assert dt0 == o.cfl_time_step(c.U0, 0.5)
    Uo, sto, _, _, _ = o.integrate(c.U0, 0.0, 1e9, dt0,
                                   OL.OrcOpts(atol=1e-8, rtol=1e-6, max_accepted=2))
    o.close()
    assert (st.accepted, st.rejected, st.rhsEvals) == (sto.accepted, sto.rejected, sto.rhsEvals)
    assert rel_l2(Ug, Uo) <= 1e-12
    # the inflow has moved the state: the check is not of a state at rest
    assert rel_l2(Ug, c.U0) > 1e-12


def test_inversion_fullsize_rhs_and_steps():
    """Config 4 at 256x128 (walls, six mounds, eight surface humps): the RHS of
    a perturbed state bit for bit, and 12 adaptive steps at the workload's
    tolerances with the oracle's counts and relative L2 <= 1e-12."""
    import sgnx
    from sgnx import cases
    c = cases.inversion(seed=0)
    assert (c.Nx, c.Ny) == (256, 128)
    s = sgnx.SGN2D(**c.solver_kwargs())
    o = OL.Oracle.from_case(c)
    U = _perturbed(c, 7)
    assert_bitwise(s.rhs(U), o.rhs(U), "inversion 256x128 RHS")
    assert_bitwise(s.rhs(c.U0), o.rhs(c.U0), "inversion 256x128 RHS of the initial state")
    s.upload(c.U0)
    dt0 = s.cflTimeStep(None, 0.5)
    kw = dict(atol=c.meta["atol"], rtol=c.meta["rtol"], max_accepted=12)
    st = s.integrateAdaptive(0.0, 1e9, dt0, sgnx.RkOpts(**kw))
    Uo, sto, _, _, _ = o.integrate(c.U0, 0.0, 1e9, dt0, OL.OrcOpts(**kw))
    assert (st.accepted, st.rejected, st.rhsEvals) == (sto.accepted, sto.rejected, sto.rhsEvals)
    assert rel_l2(s.download(), Uo) <= 1e-12
    s.close()
    o.close()


def test_gauge_cotangent_size_is_checked():
    """adjoint() reads rows x gauges values from the gauge cotangent: a buffer
    of another size is refused before the library reads it."""
    import sgnx
    from sgnx import cases
    c = cases.gate()
    s = sgnx.SGN2D(**c.solver_kwargs())
    s.setDepth(c.depth["Hnod"])
    s.upload(c.U0)
    s.setGauges(np.array([3, 40, 77], dtype=np.int64), 8)
    s.tapeBegin(8)
    s.integrateAdaptive(0.0, 1e9, s.cflTimeStep(None, 0.5), sgnx.RkOpts(max_accepted=4))
    s.tapeEnd()
    rows = s.gaugeLog()[1].shape[0]
    assert rows == 4
    with pytest.raises(ValueError, match="gauge_bar"):
        s.adjoint(None, np.ones(rows * 3 - 1))
    U0b, thb = s.adjoint(None, np.ones((rows, 3)))
    assert np.isfinite(U0b).all() and np.isfinite(thb).all()
    b = sgnx.Batch([s])
    with pytest.raises(ValueError, match="member 0"):
        b.adjoint(None, [np.ones(5)])
    b.close()
    s.close()
