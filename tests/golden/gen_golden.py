"""Write the golden fixtures of tests/golden/ from the reference CPU solver.

Runs in the build container only (needs oracle/_ref/libsgnref.so, which
oracle/Makefile compiles from the reference checkout's own sources). The
fixtures are data: inputs and the reference's outputs, as hex doubles, .npy
arrays or FNV-1a 64 digests of output bits.

    python tests/golden/gen_golden.py
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as OL  # noqa: E402

G, GAMMA, ALPHA = 9.81, 1.5, 3.0


def hx(a):
    return [float(v).hex() for v in np.ravel(a)]


def par4(g=G, c=ALPHA * math.sqrt(G * 1.0), gamma=GAMMA, nm=0.0):
    return np.array([g, c, gamma, nm])


def fr_tables(ref):
    out = {}
    for gauss in (0, 1):
        for P in range(1, 7):
            n = P + 1
            arrs = [np.zeros(n), np.zeros(n * n)] + [np.zeros(n) for _ in range(5)]
            assert ref.ref_fr_operators(P, gauss, *[OL.dp(a) for a in arrs]) == 0
            key = f"{'gauss' if gauss else 'gll'}_P{P}"
            out[key] = dict(zip(["x", "D", "lL", "lR", "gLp", "gRp", "qw"], [hx(a) for a in arrs]))
    return out


def pointwise(ref, n=1000, seed=20260101):
    rng = np.random.default_rng(seed)
    h = rng.uniform(0.05, 3.0, n)
    U = np.stack([h, h * rng.uniform(-2, 2, n), h * rng.uniform(-2, 2, n),
                  h * rng.uniform(-0.5, 0.5, n), h * rng.uniform(-1.0, 1.0, n)], axis=1)
    UR = U * (1.0 + 0.1 * rng.uniform(-1, 1, (n, 5)))
    Hl = rng.uniform(0.0, 3.0, n)
    Hx_ = rng.uniform(-1, 1, n)
    Hy_ = rng.uniform(-1, 1, n)
    ghx = rng.uniform(-1, 1, n)
    ghy = rng.uniform(-1, 1, n)
    pars = np.stack([par4(), par4(nm=0.03)], axis=0)
    outs = {k: [] for k in ("fx", "fy", "src", "src_manning", "ncpx", "ncpy", "ce", "ws0", "ws1",
                            "rus0", "rus1", "energy")}
    b = np.zeros(5)
    for i in range(n):
        u = np.ascontiguousarray(U[i])
        ur = np.ascontiguousarray(UR[i])
        p0 = np.ascontiguousarray(pars[0])
        p1 = np.ascontiguousarray(pars[1])
        for name, fn in (("fx", ref.ref_flux_x), ("fy", ref.ref_flux_y)):
            fn(OL.dp(u), Hl[i], OL.dp(p0), OL.dp(b))
            outs[name].append(b.copy())
        ref.ref_source(OL.dp(u), Hl[i], Hx_[i], Hy_[i], OL.dp(p0), OL.dp(b))
        outs["src"].append(b.copy())
        ref.ref_source(OL.dp(u), Hl[i], Hx_[i], Hy_[i], OL.dp(p1), OL.dp(b))
        outs["src_manning"].append(b.copy())
        ref.ref_ncp_x(OL.dp(u), ghx[i], OL.dp(p0), OL.dp(b))
        outs["ncpx"].append(b.copy())
        ref.ref_ncp_y(OL.dp(u), ghy[i], OL.dp(p0), OL.dp(b))
        outs["ncpy"].append(b.copy())
        outs["ce"].append(ref.ref_signal_speed(OL.dp(u), OL.dp(p0)))
        outs["ws0"].append(ref.ref_max_wave_speed(OL.dp(u), 0, OL.dp(p0)))
        outs["ws1"].append(ref.ref_max_wave_speed(OL.dp(u), 1, OL.dp(p0)))
        ref.ref_rusanov(OL.dp(u), OL.dp(ur), 0, Hl[i], OL.dp(p0), OL.dp(b))
        outs["rus0"].append(b.copy())
        ref.ref_rusanov(OL.dp(u), OL.dp(ur), 1, Hl[i], OL.dp(p0), OL.dp(b))
        outs["rus1"].append(b.copy())
        outs["energy"].append(ref.ref_energy_density(OL.dp(u), Hl[i], OL.dp(p0)))
    arrays = {k: np.array(v) for k, v in outs.items()}
    arrays.update(U=U, UR=UR, H=Hl, Hx=Hx_, Hy=Hy_, ghx=ghx, ghy=ghy, pars=pars)
    np.savez(os.path.join(HERE, "pointwise.npz"), **arrays)


class RefSolver:
    def __init__(self, ref, P, Nx, Ny, Lx, Ly, par, depth_fn):
        self.ref = ref
        self.cb = OL.DEPTH_CB(depth_fn) if depth_fn else OL.DEPTH_CB()
        self.par = np.ascontiguousarray(par)
        err = C.create_string_buffer(256)
        self.h = ref.ref_create(P, Nx, Ny, Lx, Ly, OL.dp(self.par), self.cb, 0, err, 256)
        assert self.h, err.value
        self.Nd = ref.ref_ndof(self.h)
        self.P, self.Nx, self.Ny = P, Nx, Ny

    def coords(self):
        x, y = np.zeros(self.Nd), np.zeros(self.Nd)
        self.ref.ref_node_coords(self.h, OL.dp(x), OL.dp(y))
        return x, y

    def bathy(self):
        ne5 = self.Nx * self.Ny * (self.P + 1)
        a = [np.zeros(self.Nd) for _ in range(3)] + [np.zeros(ne5), np.zeros(ne5)]
        self.ref.ref_bathymetry(self.h, *[OL.dp(v) for v in a])
        return a

    def rhs(self, U):
        R = np.zeros_like(U)
        assert self.ref.ref_rhs(self.h, OL.dp(U), OL.dp(R)) == 0
        return R

    def integrate(self, U, t0, tEnd, dt0, opt7, max_calls=0, call_cap=8192):
        U = np.array(U)
        ct = np.zeros(call_cap)
        ch = np.zeros(call_cap, dtype=np.uint64)
        sl = (C.c_long * 3)()
        sd = np.zeros(3)
        ab = C.c_int()
        rc = self.ref.ref_integrate(self.h, OL.dp(U), t0, tEnd, dt0, OL.dp(opt7), max_calls,
                                    OL.dp(ct), ch.ctypes.data_as(C.POINTER(C.c_ulonglong)),
                                    call_cap, sl, OL.dp(sd), C.byref(ab))
        return rc, U, list(sl), sd, ab.value, ct, ch


def gate_fixture(ref):
    """8x6 gate of test/test_cuda_sgn2d.cu:52-128 (bump bathymetry, smooth state)."""
    Lx, Ly, Nx, Ny, P = 10.0, 8.0, 8, 6, 4

    def bump(x, y):
        dx, dy = x - 0.5 * Lx, y - 0.5 * Ly
        return 1.0 - 0.4 * math.exp(-(dx * dx + dy * dy) / 6.0)

    rs = RefSolver(ref, P, Nx, Ny, Lx, Ly, par4(), bump)
    x, y = rs.coords()
    Hn, Hx_, Hy_, HfX, HfY = rs.bathy()
    kx, ky = 2.0 * math.pi / Lx, 2.0 * math.pi / Ly
    h = np.array([Hn[i] + 0.04 * math.sin(kx * x[i]) * math.cos(ky * y[i]) for i in range(rs.Nd)])
    u = np.array([0.08 * math.cos(kx * v) for v in x])
    v_ = np.array([-0.05 * math.sin(ky * v) for v in y])
    w = np.array([0.02 * math.sin(kx * v) for v in x])
    p = np.array([0.01 * math.cos(ky * v) for v in y])
    U = np.concatenate([h, h * u, h * v_, h * w, h * p])
    R = rs.rhs(U)
    lam = ref.ref_max_wave_speed_state(rs.h, OL.dp(U))
    dt0 = ref.ref_cfl_time_step(rs.h, OL.dp(U), 0.5)
    mass = ref.ref_mass(rs.h, OL.dp(U))
    energy = ref.ref_energy(rs.h, OL.dp(U))
    # lake at rest over the bump (test/test_sgn2d.cpp:55-66)
    z = np.zeros(rs.Nd)
    Rrest = rs.rhs(np.concatenate([Hn, z, z, z, z]))
    # fixed steps
    dt = 0.5 * dt0
    Urk4 = U.copy()
    Ubs3 = U.copy()
    for k in range(3):
        ref.ref_rk4_step(rs.h, k * dt, dt, OL.dp(Urk4))
        ref.ref_bs3_step(rs.h, k * dt, dt, OL.dp(Ubs3))
    # adaptive run of the reference gate (atol 1e-10, rtol 1e-8, tEnd 0.3)
    opt7 = np.array([1e-10, 1e-8, 1e-13, 1e30, 0.9, 0.2, 5.0])
    rc, Uend, sl, sd, ab, ct, ch = rs.integrate(U, 0.0, 0.3, dt0, opt7)
    ncall = sl[2]
    np.savez(os.path.join(HERE, "gate_8x6.npz"), x=x, y=y, Hnod=Hn, Hx=Hx_, Hy=Hy_,
             HfaceX_left=HfX, HfaceY_bottom=HfY, U=U, R=R, Rrest=Rrest, Urk4=Urk4, Ubs3=Ubs3,
             Uend=Uend, call_t=ct[:ncall], call_h=ch[:ncall])
    meta = {"P": P, "Nx": Nx, "Ny": Ny, "Lx": Lx, "Ly": Ly, "par": hx(par4()),
            "maxWaveSpeed": float(lam).hex(), "cfl0.5": float(dt0).hex(),
            "mass": float(mass).hex(), "energy": float(energy).hex(), "dt_fixed": float(dt).hex(),
            "adaptive": {"opts": hx(opt7), "tEnd": 0.3, "rc": rc, "accepted": sl[0],
                         "rejected": sl[1], "rhsEvals": sl[2], "tReached": float(sd[0]).hex(),
                         "dtLast": float(sd[1]).hex(), "maxNormLast": float(sd[2]).hex(),
                         "aborted": ab}}
    return meta


def solitary_fixture(ref):
    """Config-1 geometry, all sides periodic (the reference has no walls):
    128x64 on 64x32 m, H == 1 through the bathymetry path, solitary wave a = 0.1
    at x = 16 m, atol 1e-8 / rtol 1e-6, dt0 = cflTimeStep(U, 0.5)."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "sgn-hyperbolic-gpu_amd"))
    from sgnx import cases  # the same IC builder the tests use
    P, Nx, Ny, Lx, Ly = 4, 128, 64, 64.0, 32.0
    rs = RefSolver(ref, P, Nx, Ny, Lx, Ly, par4(), lambda x, y: 1.0)
    x, _ = rs.coords()
    Hn, *_ = rs.bathy()
    h, u, w, p = cases.solitary_state(x, Hn, 1.0, 0.1, G, GAMMA, x0=16.0)
    U = cases.prim_to_state(h, u, np.zeros_like(h), w, p)
    R = rs.rhs(U)
    dt0 = ref.ref_cfl_time_step(rs.h, OL.dp(U), 0.5)
    opt7 = np.array([1e-8, 1e-6, 1e-13, 1e30, 0.9, 0.2, 5.0])
    # first 10 attempted steps: 31 RHS calls, cut when call 32 begins
    rc10, U10, _, _, _, ct10, ch10 = rs.integrate(U, 0.0, 1e9, dt0, opt7, max_calls=31)
    # 200 accepted steps: run to a far time and cut at 1 + 3*(200 + rejections) calls.
    # The reference has no step-count stop, so find the call count from a long run's
    # t-log: accepted steps end where t advances.
    rc, Ulong, _, _, _, ct, ch = rs.integrate(U, 0.0, 1e9, dt0, opt7, max_calls=1 + 3 * 240,
                                              call_cap=1 + 3 * 240)
    meta = {"IC_digest": OL.fnv64_py(U), "RHS_digest": OL.fnv64_py(R), "dt0": float(dt0).hex(),
            "mass0": float(ref.ref_mass(rs.h, OL.dp(U))).hex(),
            "energy0": float(ref.ref_energy(rs.h, OL.dp(U))).hex(),
            "first10": {"rc": rc10, "state_digest": OL.fnv64_py(U10),
                        "call_t": hx(ct10[:31]), "call_h": [str(int(v)) for v in ch10[:31]]},
            "calls240": {"call_t": hx(ct), "call_h": [str(int(v)) for v in ch]}}
    return meta


def rk_ode_fixture(ref):
    opt7 = np.array([1e-10, 1e-8, 1e-13, 1e30, 0.9, 0.2, 5.0])
    out = {}
    for kind, u0, tEnd, param in ((0, [1.0], 1.0, 0.0), (1, [1.0, 0.0], 20.0, 0.0),
                                  (2, [1.0], 2.0, -3.0)):
        u = np.array(u0)
        sl = (C.c_long * 3)()
        sd = np.zeros(3)
        ab = C.c_int()
        ref.ref_integrate_ode(kind, param, OL.dp(u), len(u0), 0.0, tEnd, 1e-3, OL.dp(opt7), sl,
                              OL.dp(sd), C.byref(ab))
        u4 = np.array(u0)
        u3 = np.array(u0)
        for k in range(32):
            ref.ref_rk4_step_ode(kind, param, OL.dp(u4), len(u0), k / 32, 1.0 / 32)
            ref.ref_bs3_step_ode(kind, param, OL.dp(u3), len(u0), k / 32, 1.0 / 32)
        out[f"kind{kind}"] = {"u0": u0, "tEnd": tEnd, "param": param, "u": hx(u),
                              "stats": list(sl), "stats_d": hx(sd), "aborted": ab.value,
                              "rk4_32": hx(u4), "bs3_32": hx(u3)}
    out["opts"] = hx(opt7)
    return out


def main():
    ref = OL.ref()
    with open(os.path.join(HERE, "fr_tables.json"), "w") as f:
        json.dump(fr_tables(ref), f, indent=1)
    pointwise(ref)
    meta = {"gate_8x6": gate_fixture(ref), "solitary_128x64_periodic": solitary_fixture(ref),
            "rk_ode": rk_ode_fixture(ref)}
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
