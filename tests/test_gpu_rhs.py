"""GPU parity of the fused RHS kernel against the oracle (bitwise).

Every case builds the same still-depth arrays and state for libsgnx and for
the oracle and requires dU to agree bit for bit (finite inputs).
"""
import math

import numpy as np
import pytest

import oracle_lib as OL

pytestmark = pytest.mark.gpu


def assert_bitwise(got, want, what):
    g = np.ascontiguousarray(got, dtype=np.float64).view(np.uint64)
    w = np.ascontiguousarray(want, dtype=np.float64).view(np.uint64)
    assert g.shape == w.shape, what
    bad = np.flatnonzero(g != w)
    if bad.size:
        i = int(bad[0])
        a, b = got.ravel()[i], want.ravel()[i]
        raise AssertionError(f"{what}: {bad.size} of {g.size} entries differ; first at {i}: "
                             f"{a!r} ({float(a).hex()}) vs {b!r} ({float(b).hex()}), "
                             f"max abs diff {np.nanmax(np.abs(got - want)):.3e}")


# RHS kernel variants checked by every parity case: (variant, rows per chunk)
KERNELS = [(0, 16), (1, 16), (1, 3), (1, 1)]


def run_pair(case, t=0.0, U=None, kernels=KERNELS):
    """GPU RHS under every kernel variant; all must agree with the oracle bitwise.
    Returns the first variant's result and the oracle's."""
    import sgnx
    U = case.U0 if U is None else U
    s = sgnx.SGN2D(**case.solver_kwargs())
    o = OL.Oracle.from_case(case)
    try:
        want = o.rhs(U, t)
        outs = []
        for variant, rows in kernels:
            s.set_rhs_kernel(variant, rows)
            got = s.rhs(U, t)
            assert_bitwise(got, want, f"{case.name} kernel variant {variant} rows {rows}")
            outs.append(got)
        return outs[0], want
    finally:
        s.close()
        o.close()


def test_rhs_gate_case_bitwise():
    from sgnx import cases
    got, want = run_pair(cases.gate())
    assert_bitwise(got, want, "gate 8x6 RHS")


@pytest.mark.parametrize("Nx,Ny", [(2, 2), (3, 5), (10, 3), (11, 4), (23, 7), (40, 9), (21, 33)])
@pytest.mark.parametrize("sides", ["periodic", "walls_y", "walls"])
def test_rhs_random_bitwise(Nx, Ny, sides):
    from sgnx import cases
    bc = {"periodic": ("periodic",) * 4, "walls_y": ("periodic", "periodic", "wall", "wall"),
          "walls": ("wall",) * 4}[sides]
    Lx, Ly = 1.3 * Nx, 0.7 * Ny          # hx, hy not powers of two
    depth = cases.random_depth(4, Nx, Ny, Lx, Ly, seed=Nx * 31 + Ny, bc=bc)
    case = cases.Case("random", 4, Nx, Ny, Lx, Ly, cases.Params(9.81, 3.0 * math.sqrt(9.81), 1.5),
                      depth, bc, cases.random_state(Nx * Ny * 25, seed=Nx + 100 * Ny))
    got, want = run_pair(case)
    assert_bitwise(got, want, f"random {Nx}x{Ny} {sides}")


def test_rhs_manning_within_tolerance():
    """Manning friction on (n = 0.03, sgn_flux2d.h:135-141): the friction
    coefficient divides by cbrt(h), and the device's cbrt is the device
    library's, not the C library's, so a friction-on RHS is held to a
    tolerance: each entry within 1e-13 of the oracle's relative to its
    variable's largest magnitude (a one-ulp cbrt difference moves the
    friction term by about 1e-16 of itself). The trajectory bar holds too:
    20 adaptive steps with the oracle's step counts and relative L2 <= 1e-12."""
    import dataclasses
    import sgnx
    from sgnx import cases
    c = cases.gate(Nx=11, Ny=7, bc=("periodic", "periodic", "wall", "wall"))
    c = dataclasses.replace(c, par=cases.Params(c.par.g, c.par.c, c.par.gamma, 0.03))
    s = sgnx.SGN2D(**c.solver_kwargs())
    o = OL.Oracle.from_case(c)
    got, want = s.rhs(c.U0), o.rhs(c.U0)
    nd = c.nd
    for v in range(5):
        g_, w_ = got[v * nd:(v + 1) * nd], want[v * nd:(v + 1) * nd]
        scale = max(np.abs(w_).max(), 1e-300)
        assert np.abs(g_ - w_).max() <= 1e-13 * scale, v
    # the friction changes the result (the test is not vacuous)
    o0 = OL.Oracle.from_case(dataclasses.replace(c, par=cases.Params(c.par.g, c.par.c, c.par.gamma)))
    assert np.abs(o0.rhs(c.U0) - want).max() > 1e-6
    o0.close()
    dt0 = o.cfl_time_step(c.U0, 0.5)
    s.upload(c.U0)
    st = s.integrateAdaptive(0.0, 1e9, dt0, sgnx.RkOpts(atol=1e-10, rtol=1e-8, max_accepted=20))
    Ug = s.download()
    Uo, sto, _, _, _ = o.integrate(c.U0, 0.0, 1e9, dt0,
                                   OL.OrcOpts(atol=1e-10, rtol=1e-8, max_accepted=20))
    s.close()
    o.close()
    assert (st.accepted, st.rejected, st.rhsEvals) == (sto.accepted, sto.rejected, sto.rhsEvals)
    assert np.linalg.norm(Ug - Uo) <= 1e-12 * np.linalg.norm(Uo)


def test_manning_solver_keeps_to_the_tile_kernel():
    """The row sweep has no friction term: a solver with friction runs the
    element-row tiles (the test above), and refuses the row sweep, a batch
    and the autotuner with SGNX_EUNSUPPORTED, leaving itself usable."""
    import dataclasses
    import sgnx
    from sgnx import _lib as L
    from sgnx import cases
    c = cases.gate(Nx=11, Ny=7, bc=("periodic", "periodic", "wall", "wall"))
    c = dataclasses.replace(c, par=cases.Params(c.par.g, c.par.c, c.par.gamma, 0.03))
    s = sgnx.SGN2D(**c.solver_kwargs())
    with pytest.raises(L.SgnxError, match="Manning") as e:
        s.set_rhs_kernel(1, 8)
    assert e.value.code == L.SGNX_EUNSUPPORTED
    with pytest.raises(L.SgnxError, match="Manning"):
        sgnx.Batch([s])
    s.upload(c.U0)
    with pytest.raises(L.SgnxError, match="Manning"):
        s.autotune()
    o = OL.Oracle.from_case(c)
    got, want = s.rhs(c.U0), o.rhs(c.U0)
    o.close()
    s.close()
    nd = c.nd
    for v in range(5):
        scale = max(np.abs(want[v * nd:(v + 1) * nd]).max(), 1e-300)
        assert np.abs(got[v * nd:(v + 1) * nd] - want[v * nd:(v + 1) * nd]).max() <= 1e-13 * scale


def test_rhs_flat_no_depth_bitwise():
    """Flat bottom given as no still depth at all (the reference's H == 0)."""
    from sgnx import cases
    Nx, Ny = 12, 5
    case = cases.Case("flat", 4, Nx, Ny, 6.0, 2.5, cases.Params(9.81, 2.0, 2.0), None,
                      ("periodic",) * 4, cases.random_state(Nx * Ny * 25, seed=3))
    got, want = run_pair(case)
    assert_bitwise(got, want, "flat H == 0")


def test_rhs_inflow_bitwise():
    from sgnx import cases
    Nx, Ny = 21, 6
    Lx, Ly = 4.0, 1.5
    bc = ("inflow", "wall", "wall", "wall")
    depth = {k: 0.45 * v for k, v in cases.random_depth(4, Nx, Ny, Lx, Ly, seed=5, bc=bc).items()}
    HfX = depth["HfaceX"].reshape(Ny, Nx + 1, 5)[:, 0, :].reshape(-1)
    table, omega = cases.inflow_table(Ny, 4, HfX, period=1.0, amp=0.01)
    case = cases.Case("inflow", 4, Nx, Ny, Lx, Ly, cases.Params(9.81, 3.0 * math.sqrt(9.81 * 0.45), 1.5),
                      depth, bc, cases.random_state(Nx * Ny * 25, seed=9), inflow=(table, omega))
    for t in (0.0, 0.123, 0.37):
        got, want = run_pair(case, t=t)
        assert_bitwise(got, want, f"inflow t={t}")


def test_rhs_solitary_128x64_bitwise():
    from sgnx import cases
    got, want = run_pair(cases.solitary())
    assert_bitwise(got, want, "solitary 128x64 walls-in-y RHS")


def test_lake_at_rest_gpu():
    """Well-balancing (test/test_sgn2d.cpp:55-66): h = H, u = v = w = p = 0."""
    import sgnx
    from sgnx import cases
    case = cases.gate()
    Hn = case.depth["Hnod"]
    z = np.zeros_like(Hn)
    U = cases.prim_to_state(Hn.copy(), z, z, z, z)
    s = sgnx.SGN2D(**case.solver_kwargs())
    R = s.rhs(U)
    s.close()
    assert np.abs(R).max() < 1e-12


def test_bathymetry_gradients_bitwise():
    import sgnx
    from sgnx import cases
    case = cases.gate()
    s = sgnx.SGN2D(**case.solver_kwargs())
    Hx, Hy = s.stillDepthGrad()
    s.close()
    o = OL.Oracle.from_case(case)
    _, oHx, oHy = o.bathymetry()
    assert_bitwise(Hx, oHx, "Hx")
    assert_bitwise(Hy, oHy, "Hy")


def test_rhs_device_pointers_and_determinism():
    """sgnx_rhs on caller-owned device buffers; repeated evaluations give
    identical bits."""
    import sgnx
    from sgnx import cases, hip
    case = cases.gate(Nx=17, Ny=9)
    s = sgnx.SGN2D(**case.solver_kwargs())
    U = hip.DeviceArray.from_host(case.U0)
    d1 = hip.DeviceArray(U.n)
    d2 = hip.DeviceArray(U.n)
    s.rhs_device(U.address, d1.address)
    s.rhs_device(U.address, d2.address)
    hip.synchronize()
    a, b = d1.download(), d2.download()
    s.close()
    assert_bitwise(a, b, "run-to-run")
    want = OL.Oracle.from_case(case).rhs(case.U0)
    assert_bitwise(a, want, "device-pointer RHS")
    for d in (U, d1, d2):
        d.free()


def test_kernel_settings_read_back(monkeypatch):
    """sgnx_get_rhs_kernel reports the settings in use: the waves mask from
    SGNX_SWEEP_WAVES3 at creation (decimal, a leading zero included, or
    0x-hexadecimal; 0x2 without it), the tile height sgnx_set_rhs_kernel sets
    and what the autotuner chose."""
    import sgnx
    from sgnx import cases
    case = cases.gate(Nx=8, Ny=6)
    for text, want in (("0x3", 0x3), ("10", 0xA), ("010", 0xA), ("0xF", 0xF), ("2", 0x2),
                       ("0", 0x0)):
        monkeypatch.setenv("SGNX_SWEEP_WAVES3", text)
        with sgnx.SGN2D(**case.solver_kwargs()) as s:
            assert s.rhs_kernel()[2] == want, (text, s.rhs_kernel())
    monkeypatch.delenv("SGNX_SWEEP_WAVES3")
    with sgnx.SGN2D(**case.solver_kwargs()) as s:
        variant, rows, waves3 = s.rhs_kernel()
        assert variant == 1 and rows >= 1 and waves3 == 0x2
        s.set_rhs_kernel(1, 8)
        assert s.rhs_kernel() == (1, 8, 0x2)
        s.upload(case.U0)
        r, w = s.autotune(1)
        assert s.rhs_kernel() == (1, r, w)


@pytest.mark.parametrize("ci", range(3))
def test_wave_and_gradient_builds_give_the_same_bits(ci, monkeypatch):
    """The builds a launch kind can take (two or three waves per SIMD, Hx, Hy
    read or formed from Hnod in phase C, the face operands issued after
    phase A or at the start of phase C) are the same arithmetic: 12 adaptive
    steps under every wave mask and either gradient setting give the oracle's
    step counts and one final state bit for bit, and the plain RHS of that
    state is the oracle's."""
    import sgnx
    from sgnx import cases
    case = [cases.gate(Nx=11, Ny=7, bc=("periodic", "periodic", "wall", "wall")),
            cases.solitary(Nx=48, Ny=24, walls=True),
            cases.shoal(seed=3, Nx=40, Ny=20)][ci]
    o = OL.Oracle.from_case(case)
    dt0 = o.cfl_time_step(case.U0, 0.5)
    opts = dict(atol=1e-8, rtol=1e-6, max_accepted=12)
    Uo, sto, *_ = o.integrate(case.U0, 0.0, 1e9, dt0, OL.OrcOpts(**opts))
    finals = []
    for mask in (0x0, 0x2, 0xF, 0x1, 0xA):
        for hxr in ("1", "0"):
            # the mask as decimal or as 0x-hexadecimal text, both read alike
            monkeypatch.setenv("SGNX_SWEEP_WAVES3", f"{mask:#x}" if hxr == "1" else str(mask))
            monkeypatch.setenv("SGNX_SWEEP_HXR", hxr)
            s = sgnx.SGN2D(**case.solver_kwargs())
            assert s.rhs_kernel()[2] == mask, (mask, hxr, s.rhs_kernel())
            s.upload(case.U0)
            st = s.integrateAdaptive(0.0, 1e9, dt0, sgnx.RkOpts(**opts))
            assert (st.accepted, st.rejected) == (sto.accepted, sto.rejected), (mask, hxr)
            U = s.download()
            finals.append(((mask, hxr), U))
            assert_bitwise(s.rhs(U), o.rhs(U), f"{case.name} plain RHS, waves {mask:#x} hxr {hxr}")
            s.close()
    o.close()
    for key, U in finals[1:]:
        assert_bitwise(U, finals[0][1], f"{case.name} state after 12 steps, {key}")
    rel = np.linalg.norm(finals[0][1] - Uo) / np.linalg.norm(Uo)
    assert rel <= 1e-12, rel
