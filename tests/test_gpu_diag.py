"""eta diagnostics on the device against the reference applications' host
loops restated in numpy: max eta with its node in a window (a strict ">" scan
keeps the first index, apps/shoal_focus_mpi.cu:199-204), gauges
(apps/tsunami_margin_mpi.cu:274-278), element means (src/io/vtk2d.h:39-81, the
same summation order, so bit for bit), the legacy VTK frame, and the gauge log
recorded during adaptive runs (solo and batched)."""
import math

import numpy as np
import pytest

from test_gpu_rhs import assert_bitwise

pytestmark = pytest.mark.gpu


def _state(case, seed):
    from sgnx import cases
    U = cases.random_state(case.nd, seed=seed, h0=0.5, amp=0.2)
    U[:case.nd] += case.depth["Hnod"] - 0.5
    return U


def _scan_max(eta, mask):
    best, idx = -1e300, -1
    for i in np.flatnonzero(mask):           # ascending node order
        if eta[i] > best:
            best, idx = eta[i], int(i)
    return best, idx


def test_eta_max_and_gauges_match_host_scan():
    import sgnx
    from sgnx import cases
    c = cases.gate(Nx=13, Ny=7)
    s = sgnx.SGN2D(**c.solver_kwargs())
    U = _state(c, 3)
    s.upload(U)
    eta = U[:c.nd] - c.depth["Hnod"]
    x, y = s.nodeCoords() if hasattr(s, "nodeCoords") else sgnx.node_coords(c.P, c.Nx, c.Ny, c.Lx, c.Ly)
    for win in [None, (2.0, 7.5, 1.0, 6.0), (4.999, 5.001, -1.0, 9.0), (20.0, 30.0, 0.0, 1.0)]:
        x0, x1, y0, y1 = win if win else (-math.inf, math.inf, -math.inf, math.inf)
        mask = (x > x0) & (x < x1) & (y > y0) & (y < y1)
        got = s.etaMax(win)
        want = _scan_max(eta, mask)
        assert got == want, (win, got, want)
    # ties resolve to the lowest node: a plateau planted at three nodes
    U2 = U.copy()
    top = float(eta.max()) + 1.0
    for i in (700, 91, 2000):
        U2[i] = c.depth["Hnod"][i] + top
    s.upload(U2)
    assert s.etaMax() == (U2[91] - c.depth["Hnod"][91], 91)
    nodes = np.array([0, 5, 91, 700, c.nd - 1])
    assert_bitwise(s.gauges(nodes), U2[nodes] - c.depth["Hnod"][nodes], "gauges")
    s.close()


def test_element_means_bitwise_and_vtk(tmp_path):
    import sgnx
    from sgnx import cases, io
    c = cases.gate(Nx=9, Ny=4)
    s = sgnx.SGN2D(**c.solver_kwargs())
    U = _state(c, 5)
    s.upload(U)
    nd, ne = c.nd, c.Nx * c.Ny
    Hn = c.depth["Hnod"]
    qw = s.ops.weights()
    fields = [U[:nd] - Hn, U[nd:2 * nd] / U[:nd], U[2 * nd:3 * nd] / U[:nd],
              U[4 * nd:5 * nd] / U[:nd], Hn]
    want = np.zeros((5, ne))
    for k, f in enumerate(fields):
        fe = f.reshape(ne, 25)
        acc = np.zeros(ne)
        for b in range(5):
            for a in range(5):
                acc = acc + qw[a] * qw[b] * fe[:, a + 5 * b]
        want[k] = 0.25 * acc
    got = s.elementMeans()
    assert_bitwise(got.reshape(5, -1), want, "element means")
    path = tmp_path / "frame.vtk"
    io.write_vtk(s, str(path))
    lines = path.read_text().splitlines()
    assert lines[0] == "# vtk DataFile Version 3.0" and lines[4] == f"DIMENSIONS {c.Nx} {c.Ny} 1"
    body = [ln for ln in lines if ln and ln[0] in "-0123456789" and " " in ln.strip()]
    vals = np.array([float(v) for ln in body for v in ln.split()])
    assert vals.size == 5 * ne
    np.testing.assert_allclose(vals, want.reshape(-1), rtol=1e-5, atol=1e-12)
    s.close()


def test_gauge_log_during_integration_solo_and_batched():
    """The log holds one row per accepted step, its last row is the final
    state's gauges, its times end at tReached, and a batch member's log equals
    its solo run's bit for bit."""
    import sgnx
    from sgnx import cases
    cs = [cases.solitary(Nx=32, Ny=16), cases.gate(Nx=11, Ny=8)]
    nodes = [np.array([3, 400, 5000, 12000]), np.array([0, 77, 1500])]
    opts = [sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=k) for k in (9, 14)]
    solo = []
    ss = []
    for c, nd_, o in zip(cs, nodes, opts):
        s = sgnx.SGN2D(**c.solver_kwargs())
        s.upload(c.U0)
        dt0 = s.cflTimeStep(None, 0.5)
        s.setGauges(nd_, 64)
        st = s.integrateAdaptive(0.0, 1e9, dt0, o)
        t, eta = s.gaugeLog()
        assert eta.shape == (st.accepted, nd_.size)
        assert t[-1] == st.tReached and np.all(np.diff(t) > 0)
        assert_bitwise(eta[-1], s.gauges(nd_), "last log row")
        solo.append((t, eta, dt0))
        s.upload(c.U0)
        s.setGauges(nd_, 64)
        ss.append(s)
    b = sgnx.Batch(ss)
    b.integrateAdaptive(0.0, 1e9, [d for _, _, d in solo], opts)
    for s, (t, eta, _) in zip(ss, solo):
        tb, etab = s.gaugeLog()
        assert_bitwise(tb, t, "batched log times")
        assert_bitwise(etab, eta, "batched log")
    b.close()
    for s in ss:
        s.close()
