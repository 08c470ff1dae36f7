"""GPU parity of the Runge-Kutta paths and diagnostics against the oracle.

Fixed-step RK4 / BS3 are bitwise (RHS and stage arithmetic are). The adaptive
integrator's error sum is a fixed-order device reduction rather than the CPU's
sequential sum, so trajectories are held to the stated bar: identical accepted
and rejected step counts and relative L2 distance <= 1e-12.
"""
import numpy as np
import pytest

import oracle_lib as OL
from test_gpu_rhs import assert_bitwise

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("variant,rows", [(0, 16), (1, 16), (1, 2)])
def test_fixed_steps_bitwise(variant, rows):
    """Fixed RK4 and BS3 steps; with the row-sweep variant the BS3 stages run
    the fused stage kernels (u + a*k and the candidate formed while loading)."""
    import sgnx
    from sgnx import cases
    case = cases.gate(Nx=13, Ny=6, bc=("periodic", "periodic", "wall", "wall"))
    s = sgnx.SGN2D(**case.solver_kwargs())
    s.set_rhs_kernel(variant, rows)
    o = OL.Oracle.from_case(case)
    dt = 0.5 * o.cfl_time_step(case.U0, 0.5)
    s.upload(case.U0)
    Uo = case.U0.copy()
    for k in range(3):
        s.rk4Step(k * dt, dt)
        Uo = o.rk4_step(Uo, k * dt, dt)
    assert_bitwise(s.download(), Uo, "3 RK4 steps")
    s.upload(case.U0)
    Uo = case.U0.copy()
    for k in range(3):
        s.bs3Step(k * dt, dt)
        Uo = o.bs3_step(Uo, k * dt, dt)
    assert_bitwise(s.download(), Uo, "3 BS3 steps")
    s.close()


def test_diagnostics():
    import sgnx
    from sgnx import cases
    case = cases.gate()
    s = sgnx.SGN2D(**case.solver_kwargs())
    o = OL.Oracle.from_case(case)
    U = case.U0
    assert s.maxWaveSpeed(U) == o.max_wave_speed(U)
    assert s.cflTimeStep(U, 0.5) == o.cfl_time_step(U, 0.5)
    assert abs(s.mass(U) - o.mass(U)) <= 1e-14 * abs(o.mass(U))
    assert abs(s.energy(U) - o.energy(U)) <= 1e-13 * abs(o.energy(U))
    s.close()


def _adaptive(case, t_end, opts_kw, max_acc=0, step_cap=4096, variant=1, rows=16):
    import sgnx
    s = sgnx.SGN2D(**case.solver_kwargs())
    s.set_rhs_kernel(variant, rows)
    o = OL.Oracle.from_case(case)
    dt0 = o.cfl_time_step(case.U0, 0.5)
    s.upload(case.U0)
    st, glog = s.integrateAdaptive(0.0, t_end, dt0, sgnx.RkOpts(max_accepted=max_acc, **opts_kw),
                                   step_log_cap=step_cap)
    Ug = s.download()
    s.close()
    Uo, sto, olog, _, _ = o.integrate(case.U0, 0.0, t_end, dt0,
                                      OL.OrcOpts(max_accepted=max_acc, **opts_kw),
                                      step_cap=step_cap)
    return st, sto, Ug, Uo, glog, olog


@pytest.mark.parametrize("variant,rows", [(0, 16), (1, 16), (1, 4)])
def test_adaptive_gate_matches_oracle(variant, rows):
    """The reference's adaptive gate (test/test_cuda_sgn2d.cu:96-128), held to
    the stricter bar: same step counts, relative L2 <= 1e-12 (unfused step
    kernels with variant 0; fused stage inputs and error with variant 1)."""
    from sgnx import cases
    case = cases.gate()
    st, sto, Ug, Uo, glog, olog = _adaptive(case, 0.3, dict(atol=1e-10, rtol=1e-8),
                                            variant=variant, rows=rows)
    assert not st.aborted and not sto.aborted
    assert (st.accepted, st.rejected, st.rhsEvals) == (sto.accepted, sto.rejected, sto.rhsEvals)
    assert rel_l2(Ug, Uo) <= 1e-12
    n = st.accepted + st.rejected
    # first step: identical stages, the error sums differ only in summation order
    assert glog[0, 0] == olog[0, 0]
    assert abs(glog[0, 1] - olog[0, 1]) <= 1e-13 * olog[0, 1]
    # later steps: the embedded error (a difference of nearly equal states)
    # magnifies last-bit state differences, so dt tracks loosely; the bar is
    # the step counts and the trajectory distance above
    np.testing.assert_allclose(glog[:n, 0], olog[:n, 0], rtol=1e-6, atol=0)


def test_adaptive_solitary_200_steps():
    """Solitary wave, 128x64, walls in y: 200 accepted steps in one call."""
    from sgnx import cases
    case = cases.solitary()
    st, sto, Ug, Uo, glog, olog = _adaptive(case, 1e9, dict(atol=1e-8, rtol=1e-6), max_acc=200)
    assert (st.accepted, st.rejected, st.rhsEvals) == (sto.accepted, sto.rejected, sto.rhsEvals)
    assert st.accepted == 200
    assert rel_l2(Ug, Uo) <= 1e-12
    assert st.tReached == pytest.approx(sto.tReached, rel=1e-12)


@pytest.mark.parametrize("walls", [True, False])
def test_solitary_first_10_steps_rhs_bitwise(walls):
    """The solitary-wave workload's gate: every RHS evaluation of the oracle's
    first 10 accepted adaptive steps (the FSAL head and three per attempt:
    y2, y3 and the candidate; 31 without rejections), evaluated by the GPU on
    the same inputs, bit for bit, with the row sweep and the tile kernel; the
    outputs' digests and times are those the oracle's own run logged, in
    order, so the inputs are its stage states. On the all-periodic grid the
    reference's own run is in the fixtures: the GPU's first 31 outputs carry
    its digests (tests/golden/golden.json, "first10")."""
    import json
    import os
    import sgnx
    from sgnx import cases
    case = cases.solitary(walls=walls)
    o = OL.Oracle.from_case(case)
    dt0 = o.cfl_time_step(case.U0, 0.5)
    _, sto, olog, ct, ch = o.integrate(case.U0, 0.0, 1e9, dt0,
                                       OL.OrcOpts(atol=1e-8, rtol=1e-6, max_accepted=10),
                                       step_cap=64, call_cap=64)
    s = sgnx.SGN2D(**case.solver_kwargs())
    u, t = case.U0.copy(), 0.0
    evals = 0
    gpu_digests = []

    def both(X, tt):
        nonlocal evals
        want = o.rhs(X, tt)
        for variant, rows in ((1, 16), (0, 16)):
            s.set_rhs_kernel(variant, rows)
            got = s.rhs(X, tt)
            assert_bitwise(got, want, f"RHS evaluation {evals}, variant {variant}")
        gpu_digests.append(OL.fnv64(got))
        assert gpu_digests[-1] == int(ch[evals]) and tt == ct[evals], f"RHS evaluation {evals}"
        evals += 1
        return want

    k1 = both(u, t)
    for i in range(sto.accepted + sto.rejected):
        dt, err = olog[i]
        # rk.cpp:99-108
        k2 = both(u + 0.5 * dt * k1, t + 0.5 * dt)
        k3 = both(u + 0.75 * dt * k2, t + 0.75 * dt)
        un = u + dt * ((2.0 / 9.0 * k1 + 1.0 / 3.0 * k2) + 4.0 / 9.0 * k3)
        k4 = both(un, t + dt)
        if err <= 1.0:
            u, t, k1 = un, t + dt, k4
    s.close()
    o.close()
    assert sto.accepted == 10 and evals == 1 + 3 * (sto.accepted + sto.rejected)
    if not walls:
        path = os.path.join(os.path.dirname(__file__), "golden", "golden.json")
        with open(path) as f:
            m = json.load(f)["solitary_128x64_periodic"]
        if OL.fnv64_py(case.U0) != m["IC_digest"]:
            pytest.skip("initial state differs from the fixture's (host libm differs)")
        want = [int(v) for v in m["first10"]["call_h"]]
        assert [int(h) for h in ch[:len(want)]] == want
        assert gpu_digests[:len(want)] == want


def test_adaptive_abort_on_dtmin():
    """dtMin above the step the controller wants -> aborted, like rk.cpp:155-160."""
    from sgnx import cases
    case = cases.gate()
    st, sto, Ug, Uo, _, _ = _adaptive(case, 0.3, dict(atol=1e-14, rtol=1e-14, dtMin=1e-2))
    assert st.aborted == 1 and sto.aborted == 1
    assert (st.accepted, st.rejected) == (sto.accepted, sto.rejected)
