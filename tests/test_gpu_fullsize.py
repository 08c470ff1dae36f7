"""Checks at the full production size (1024x512, P = 4): the RHS against the
oracle bit for bit, and size-independent properties of the discretisation.
Each test builds the shoal workload once (about a second on the host)."""
import numpy as np
import pytest

import oracle_lib as OL
from test_gpu_rhs import assert_bitwise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def shoal():
    from sgnx import cases
    return cases.shoal(seed=0)


def _walls(case):
    import dataclasses
    return dataclasses.replace(case, bc=("wall",) * 4, inflow=None)


def test_fullsize_rhs_bitwise(shoal):
    """The production workload (inflow left, walls elsewhere) on a non-trivial
    state, both kernel variants."""
    import sgnx
    from sgnx import cases
    U = cases.random_state(shoal.nd, seed=4, h0=0.3, amp=0.2)
    U[:shoal.nd] += shoal.depth["Hnod"] - 0.3
    s = sgnx.SGN2D(**shoal.solver_kwargs())
    o = OL.Oracle.from_case(shoal)
    want = o.rhs(U, 0.37)
    for variant, rows in ((1, 16), (0, 16)):
        s.set_rhs_kernel(variant, rows)
        assert_bitwise(s.rhs(U, 0.37), want, f"1024x512 variant {variant}")
    s.close()


def test_fullsize_lake_at_rest(shoal):
    """h = H, u = v = w = p = 0 over the slope and shoal with walls: RHS == 0
    to round-off (test/test_sgn2d.cpp:55-66 at full size)."""
    import sgnx
    from sgnx import cases
    case = _walls(shoal)
    Hn = case.depth["Hnod"]
    z = np.zeros_like(Hn)
    s = sgnx.SGN2D(**case.solver_kwargs())
    R = s.rhs(cases.prim_to_state(Hn.copy(), z, z, z, z))
    s.close()
    assert np.abs(R).max() < 1e-12


def test_fullsize_mass_rate_vanishes_with_walls(shoal):
    """With walls on every side the semi-discrete mass rate sum(qw_a qw_b J dh/dt)
    is a telescoping sum of interface fluxes: zero up to round-off."""
    import sgnx
    from sgnx import cases
    case = _walls(shoal)
    U = cases.random_state(case.nd, seed=8, h0=0.3, amp=0.1)
    U[:case.nd] += case.depth["Hnod"] - 0.3
    s = sgnx.SGN2D(**case.solver_kwargs())
    R = s.rhs(U)
    qw = s.ops.weights()
    s.close()
    w2 = np.outer(qw, qw).ravel()                   # node a + 5b -> qw[a] qw[b]
    dh = R[:case.nd].reshape(-1, 25)
    rate = float(np.sum(dh * w2[None, :]))
    scale = float(np.sum(np.abs(dh) * w2[None, :]))
    assert abs(rate) <= 1e-12 * scale


def test_fullsize_determinism_and_fused_step(shoal):
    """Run-to-run bitwise determinism of the RHS, and one adaptive step with
    the fused stage kernels equal to one with the separate step kernels
    (same candidate state bit for bit; error norms equal to summation order)."""
    import sgnx
    s = sgnx.SGN2D(**shoal.solver_kwargs())
    s.set_rhs_kernel(1, 16)
    a = s.rhs(shoal.U0, 0.1)
    b = s.rhs(shoal.U0, 0.1)
    assert_bitwise(a, b, "run-to-run")
    dt0 = s.cflTimeStep(shoal.U0, 0.5)
    out = {}
    for variant in (0, 1):
        s.set_rhs_kernel(variant, 16)
        s.upload(shoal.U0)
        st, log = s.integrateAdaptive(0.0, 1e9, dt0, sgnx.RkOpts(max_accepted=1), step_log_cap=8)
        out[variant] = (s.download(), st, log)
    s.close()
    (U0, st0, log0), (U1, st1, log1) = out[0], out[1]
    assert (st0.accepted, st0.rejected) == (st1.accepted, st1.rejected)
    n = st0.accepted + st0.rejected
    assert np.array_equal(log0[0, 0], log1[0, 0])
    np.testing.assert_allclose(log1[:n, 1], log0[:n, 1], rtol=1e-13, atol=0)
    if st0.rejected == 0:
        assert_bitwise(U1, U0, "state after one fused step")
    else:
        assert float(np.linalg.norm(U1 - U0) / np.linalg.norm(U0)) <= 1e-13
