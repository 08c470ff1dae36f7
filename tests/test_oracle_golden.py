"""The oracle against the reference's own outputs (golden fixtures written by
tests/golden/gen_golden.py from the compiled reference) -- bitwise."""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as OL

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def unhex(lst):
    return np.array([float.fromhex(s) for s in lst])


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.fixture(scope="module")
def meta():
    with open(os.path.join(GOLD, "golden.json")) as f:
        return json.load(f)


def test_fr_tables_match_reference():
    with open(os.path.join(GOLD, "fr_tables.json")) as f:
        tab = json.load(f)
    lib = OL.oracle()
    for key, want in tab.items():
        gauss = key.startswith("gauss")
        P = int(key.split("_P")[1])
        n = P + 1
        arrs = [np.zeros(n), np.zeros(n * n)] + [np.zeros(n) for _ in range(5)]
        assert lib.orc_fr_operators(P, int(gauss), *[OL.dp(a) for a in arrs]) == 0
        for name, got in zip(["x", "D", "lL", "lR", "gLp", "gRp", "qw"], arrs):
            assert np.array_equal(bits(got), bits(unhex(want[name]))), (key, name)


def test_pointwise_physics_match_reference():
    z = np.load(os.path.join(GOLD, "pointwise.npz"))
    lib = OL.oracle()
    b = np.zeros(5)
    n = z["U"].shape[0]
    for i in range(n):
        u = np.ascontiguousarray(z["U"][i])
        ur = np.ascontiguousarray(z["UR"][i])
        p0 = np.ascontiguousarray(z["pars"][0])
        p1 = np.ascontiguousarray(z["pars"][1])
        H = z["H"][i]
        checks = []
        lib.orc_flux_x(OL.dp(u), H, OL.dp(p0), OL.dp(b)); checks.append(("fx", b.copy()))
        lib.orc_flux_y(OL.dp(u), H, OL.dp(p0), OL.dp(b)); checks.append(("fy", b.copy()))
        lib.orc_source(OL.dp(u), H, z["Hx"][i], z["Hy"][i], OL.dp(p0), OL.dp(b))
        checks.append(("src", b.copy()))
        lib.orc_source(OL.dp(u), H, z["Hx"][i], z["Hy"][i], OL.dp(p1), OL.dp(b))
        checks.append(("src_manning", b.copy()))
        lib.orc_ncp_x(OL.dp(u), z["ghx"][i], OL.dp(p0), OL.dp(b)); checks.append(("ncpx", b.copy()))
        lib.orc_ncp_y(OL.dp(u), z["ghy"][i], OL.dp(p0), OL.dp(b)); checks.append(("ncpy", b.copy()))
        lib.orc_rusanov(OL.dp(u), OL.dp(ur), 0, H, OL.dp(p0), OL.dp(b)); checks.append(("rus0", b.copy()))
        lib.orc_rusanov(OL.dp(u), OL.dp(ur), 1, H, OL.dp(p0), OL.dp(b)); checks.append(("rus1", b.copy()))
        for name, got in checks:
            assert np.array_equal(bits(got), bits(z[name][i])), (name, i)
        for name, got in (("ce", lib.orc_signal_speed(OL.dp(u), OL.dp(p0))),
                          ("ws0", lib.orc_max_wave_speed(OL.dp(u), 0, OL.dp(p0))),
                          ("ws1", lib.orc_max_wave_speed(OL.dp(u), 1, OL.dp(p0))),
                          ("energy", lib.orc_energy_density(OL.dp(u), H, OL.dp(p0)))):
            assert bits(np.array([got]))[0] == bits(np.array([z[name][i]]))[0], (name, i)


def gate_oracle(z, m, bc=("periodic",) * 4):
    """Oracle on the gate case with the reference's still-depth arrays, the
    left/bottom edge depths extended to the interface layout (last line = the
    periodic wrap of the first, as the reference samples it)."""
    Nx, Ny = m["Nx"], m["Ny"]
    HfX = np.zeros((Ny, Nx + 1, 5))
    HfX[:, :Nx] = z["HfaceX_left"].reshape(Ny, Nx, 5)
    HfX[:, Nx] = HfX[:, 0]
    HfY = np.zeros((Ny + 1, Nx, 5))
    HfY[:Ny] = z["HfaceY_bottom"].reshape(Ny, Nx, 5)
    HfY[Ny] = HfY[0]
    depth = {"Hnod": z["Hnod"], "HfaceX": HfX.ravel(), "HfaceY": HfY.ravel()}
    return OL.Oracle(4, Nx, Ny, m["Lx"], m["Ly"], tuple(unhex(m["par"])), depth, bc)


def test_gate_rhs_and_setup(meta):
    z = np.load(os.path.join(GOLD, "gate_8x6.npz"))
    m = meta["gate_8x6"]
    o = gate_oracle(z, m)
    Hn, Hx, Hy = o.bathymetry()
    assert np.array_equal(bits(Hx), bits(z["Hx"])) and np.array_equal(bits(Hy), bits(z["Hy"]))
    x = np.zeros(o.Nd)
    y = np.zeros(o.Nd)
    OL.oracle().orc_node_coords(4, m["Nx"], m["Ny"], m["Lx"], m["Ly"], OL.dp(x), OL.dp(y))
    assert np.array_equal(bits(x), bits(z["x"])) and np.array_equal(bits(y), bits(z["y"]))
    assert np.array_equal(bits(o.rhs(z["U"])), bits(z["R"]))
    assert np.array_equal(bits(o.rhs(np.concatenate([z["Hnod"]] + [np.zeros(o.Nd)] * 4))),
                          bits(z["Rrest"]))
    assert o.max_wave_speed(z["U"]).hex() == m["maxWaveSpeed"]
    assert o.cfl_time_step(z["U"], 0.5).hex() == m["cfl0.5"]
    assert o.mass(z["U"]).hex() == m["mass"]
    assert o.energy(z["U"]).hex() == m["energy"]


def test_gate_fixed_steps(meta):
    z = np.load(os.path.join(GOLD, "gate_8x6.npz"))
    m = meta["gate_8x6"]
    o = gate_oracle(z, m)
    dt = float.fromhex(m["dt_fixed"])
    U4, U3 = z["U"].copy(), z["U"].copy()
    for k in range(3):
        U4 = o.rk4_step(U4, k * dt, dt)
        U3 = o.bs3_step(U3, k * dt, dt)
    assert np.array_equal(bits(U4), bits(z["Urk4"]))
    assert np.array_equal(bits(U3), bits(z["Ubs3"]))


def test_gate_adaptive_trajectory_bitwise(meta):
    """Full adaptive run of the reference gate: the oracle's trajectory is the
    reference's bit for bit (every RHS call's time and output digest)."""
    z = np.load(os.path.join(GOLD, "gate_8x6.npz"))
    m = meta["gate_8x6"]
    a = m["adaptive"]
    o = gate_oracle(z, m)
    opts = OL.OrcOpts(*unhex(a["opts"]))
    dt0 = float.fromhex(m["cfl0.5"])
    ncall = a["rhsEvals"]
    U, st, _, ct, ch = o.integrate(z["U"], 0.0, a["tEnd"], dt0, opts, call_cap=ncall)
    assert (st.accepted, st.rejected, st.rhsEvals) == (a["accepted"], a["rejected"], a["rhsEvals"])
    assert st.tReached.hex() == a["tReached"] and st.dtLast.hex() == a["dtLast"]
    assert st.maxNormLast.hex() == a["maxNormLast"]
    assert np.array_equal(bits(ct), bits(z["call_t"]))
    assert np.array_equal(ch, z["call_h"])
    assert np.array_equal(bits(U), bits(z["Uend"]))


def test_solitary_periodic_digests(meta):
    """Config-1 geometry with periodic sides: RHS digest, the 31 RHS calls of
    the first 10 steps, and 720 further calls, against the reference."""
    from sgnx import cases
    m = meta["solitary_128x64_periodic"]
    case = cases.solitary(walls=False)
    if OL.fnv64_py(case.U0) != m["IC_digest"]:
        pytest.skip("initial state differs from the fixture's (host libm differs); "
                    "golden comparison not applicable on this host")
    o = OL.Oracle.from_case(case)
    assert OL.fnv64(o.rhs(case.U0)) == m["RHS_digest"]
    dt0 = o.cfl_time_step(case.U0, 0.5)
    assert dt0.hex() == m["dt0"]
    assert o.mass(case.U0).hex() == m["mass0"] and o.energy(case.U0).hex() == m["energy0"]
    want_t = unhex(m["calls240"]["call_t"])
    want_h = np.array([int(v) for v in m["calls240"]["call_h"]], dtype=np.uint64)
    U, st, _, ct, ch = o.integrate(case.U0, 0.0, 1e9, dt0, OL.OrcOpts(max_accepted=240),
                                   call_cap=want_t.size)
    assert np.array_equal(bits(ct[:want_t.size]), bits(want_t))
    assert np.array_equal(ch[:want_h.size], want_h)
    f10 = m["first10"]
    assert np.array_equal(bits(ct[:31]), bits(unhex(f10["call_t"])))
    assert np.array_equal(ch[:31], np.array([int(v) for v in f10["call_h"]], dtype=np.uint64))


def test_rk_ode_match_reference(meta):
    m = meta["rk_ode"]
    lib = OL.oracle()
    opts = OL.OrcOpts(*unhex(m["opts"]))
    for key in ("kind0", "kind1", "kind2"):
        r = m[key]
        kind = int(key[-1])
        u = np.array(r["u0"], dtype=np.float64)
        st = OL.OrcStats()
        lib.orc_integrate_ode(kind, r["param"], OL.dp(u), u.size, 0.0, r["tEnd"], 1e-3,
                              OL.C.byref(opts), OL.C.byref(st))
        assert np.array_equal(bits(u), bits(unhex(r["u"]))), key
        assert [st.accepted, st.rejected, st.rhsEvals] == r["stats"], key
        u4 = np.array(r["u0"], dtype=np.float64)
        u3 = np.array(r["u0"], dtype=np.float64)
        for k in range(32):
            lib.orc_rk4_step_ode(kind, r["param"], OL.dp(u4), u4.size, k / 32, 1.0 / 32)
            lib.orc_bs3_step_ode(kind, r["param"], OL.dp(u3), u3.size, k / 32, 1.0 / 32)
        assert np.array_equal(bits(u4), bits(unhex(r["rk4_32"]))), key
        assert np.array_equal(bits(u3), bits(unhex(r["bs3_32"]))), key


def test_fnv_digest_implementations_agree():
    a = np.random.default_rng(1).standard_normal(1000)
    assert OL.fnv64(a) == OL.fnv64_py(a)
    assert math.isfinite(float(a.sum()))
