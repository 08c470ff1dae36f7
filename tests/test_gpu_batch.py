"""Batches of independent runs (sgnx_batch_*): every member of a batched
launch or adaptive run is bit for bit its solo counterpart (same RHS, same
final state, same step counts), with members of uneven sizes, all three side
types, partial strips and members that stop at different steps, with and
without the captured step graph."""
import dataclasses

import numpy as np
import pytest

from test_gpu_rhs import assert_bitwise

pytestmark = pytest.mark.gpu


def _members():
    """Four uneven grids: periodic, walls in y, inflow + walls, and a width that
    leaves a partial strip (37 elements), each with a non-trivial state."""
    from sgnx import cases
    out = [cases.gate(Nx=8, Ny=6), cases.solitary(Nx=64, Ny=32), cases.shoal(seed=3, Nx=48, Ny=24)]
    g = cases.gate(Nx=37, Ny=11, bc=("periodic", "periodic", "wall", "wall"))
    out.append(dataclasses.replace(g, name="gate37x11"))
    return out


def _solvers(cs, rows):
    import sgnx
    ss = []
    for c in cs:
        s = sgnx.SGN2D(**c.solver_kwargs())
        s.set_rhs_kernel(1, rows)
        ss.append(s)
    return ss


@pytest.mark.parametrize("rows", [16, 3])
def test_batch_rhs_bitwise_vs_solo(rows):
    import sgnx
    from sgnx import cases, hip
    cs = _members()
    ss = _solvers(cs, rows)
    Us, dUs, want = [], [], []
    for i, (c, s) in enumerate(zip(cs, ss)):
        U = cases.random_state(c.nd, seed=20 + i, h0=0.6, amp=0.15)
        U[:c.nd] += c.depth["Hnod"] - 0.6
        want.append(s.rhs(U, 0.21 * i))
        Us.append(hip.DeviceArray.from_host(U))
        dUs.append(hip.DeviceArray(U.size))
    b = sgnx.Batch(ss)
    b.rhs_device([u.address for u in Us], [d.address for d in dUs], [0.21 * i for i in range(4)])
    for i, d in enumerate(dUs):
        assert_bitwise(d.download(), want[i], f"member {i} ({cs[i].name})")
    b.close()
    for s in ss:
        s.close()


@pytest.mark.parametrize("graph", [True, False])
def test_batch_integrate_bitwise_vs_solo(graph):
    """Members stop after different numbers of accepted steps; each final state
    and its statistics equal the solo run's."""
    import sgnx
    cs = _members()
    ss = _solvers(cs, 16)
    steps = [12, 5, 9, 20]
    opts = [sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=k) for k in steps]
    dt0 = []
    solo = []
    for c, s, o in zip(cs, ss, opts):
        s.upload(c.U0)
        d = s.cflTimeStep(None, 0.5)
        dt0.append(d)
        st = s.integrateAdaptive(0.0, 1e9, d, o)
        solo.append((s.download(), st.as_dict()))
        s.upload(c.U0)
    b = sgnx.Batch(ss)
    sts = b.integrateAdaptive(0.0, 1e9, dt0, opts, graph=graph)
    for i, (s, st) in enumerate(zip(ss, sts)):
        U, sd = solo[i]
        assert st.as_dict() == sd, (i, st.as_dict(), sd)
        assert st.accepted == steps[i]
        assert_bitwise(s.download(), U, f"member {i} ({cs[i].name}) after {steps[i]} steps")
    # a second call continues from the members' current states
    sts2 = b.integrateAdaptive(0.0, 1e9, dt0, opts, graph=graph)
    assert [t.accepted for t in sts2] == steps
    b.close()
    for s in ss:
        s.close()


def test_batch_rejects_bad_members():
    import sgnx
    from sgnx import cases
    c = cases.gate()
    s = sgnx.SGN2D(**c.solver_kwargs())
    with pytest.raises(sgnx.SgnxError, match="twice"):
        sgnx.Batch([s, s])
    s.set_rhs_kernel(0, 16)
    with pytest.raises(sgnx.SgnxError, match="row-sweep"):
        sgnx.Batch([s])
    s.close()


def test_autotune_leaves_state_and_members_stay_identical():
    """The autotuners write only stage scratch buffers (U is unchanged), pick a
    candidate tile height, and a batch under the tuned setting still matches
    its members' solo runs bit for bit."""
    import sgnx
    cs = _members()
    ss = _solvers(cs, 16)
    for c, s in zip(cs, ss):
        s.upload(c.U0)
    rows, waves = ss[1].autotune(reps=2)
    assert rows in (2, 4, 8, 12, 16, 24, 32) and 0 <= waves <= 0xf
    assert_bitwise(ss[1].download(), cs[1].U0, "state after autotune")
    b = sgnx.Batch(ss)
    rows, waves = b.autotune(reps=2)
    for c, s in zip(cs, ss):
        assert_bitwise(s.download(), c.U0, f"{c.name} state after batch autotune")
    opts = [sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=k) for k in (4, 7, 3, 6)]
    dt0 = [s.cflTimeStep(None, 0.5) for s in ss]
    solo = []
    for c, s, o, d in zip(cs, ss, opts, dt0):
        st = s.integrateAdaptive(0.0, 1e9, d, o)
        solo.append((s.download(), st.as_dict()))
        s.upload(c.U0)
    sts = b.integrateAdaptive(0.0, 1e9, dt0, opts, graph=True)
    for i, (s, st) in enumerate(zip(ss, sts)):
        assert st.as_dict() == solo[i][1]
        assert_bitwise(s.download(), solo[i][0], f"member {i} under rows {rows}")
    b.close()
    for s in ss:
        s.close()
