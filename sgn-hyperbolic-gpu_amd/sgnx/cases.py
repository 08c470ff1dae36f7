"""Problem builders: still depths, initial states and boundary data of the
workloads the solver is measured and checked on.

Each builder returns a `Case` with the constructor arguments of SGN2D and the
initial state; the same arrays feed the GPU solver and (in the tests) the
oracle, so any libm difference between machines is shared by both.

Workloads
  gate       8x6 elements on 10x8 m over a Gaussian bump with a smooth
             (h,u,v,w,p) state (the reference's GPU-vs-CPU gate,
             test/test_cuda_sgn2d.cu:52-64)
  solitary   128x64 on 64x32 m, flat 1 m still depth, sech^2 solitary wave
             a/h0 = 0.1 with its crest at x = -16 m (x = 16 m in mesh
             coordinates), periodic in x and walls in y (formulas of
             apps/shoal_focus_mpi.cu:117-137)
  shoal      1024x512 on 96x48 m: 1:20 slope 0.45 -> 0.10 m plus a seeded
             elliptic Gaussian shoal, 1 s / 0.01 m incident wave on the left
             (slow-branch linear eigenmode), walls elsewhere, fluid at rest
  inversion  256x128 on 64x32 m, 0.5 m minus six seeded mounds, walls, eight
             seeded surface humps at rest, 20 seeded gauges (the adjoint's
             gauge-inversion workload)
  margin     2000x1000 on 200x100 km: a tanh continental margin 4000 -> 50 m
             across a seeded shelf break with 2 % seeded roughness, walls in x,
             periodic in y, a seeded 0.5-2 m elliptical Gaussian uplift at rest
             (the large single-grid case)
  shoal_focus  384x192 on 40x20 m, periodic: a 0.3 m solitary wave over 1 m
             crossing a Gaussian submerged shoal (the reference's production
             focusing run, apps/shoal_focus_mpi.cu:88-137)
  mms        the manufactured periodic solution of apps/convergence2d.cpp
             (flat bottom, unit square): its state and forcing
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import Params
from .solver import node_coords

G = 9.81
GAMMA = 1.5
ALPHA = 3.0


@dataclass
class Case:
    name: str
    P: int
    Nx: int
    Ny: int
    Lx: float
    Ly: float
    par: Params
    depth: object            # dict of arrays or None
    bc: tuple
    U0: np.ndarray
    inflow: object = None    # (table, omega) or None
    meta: dict = field(default_factory=dict)

    def solver_kwargs(self):
        return dict(P=self.P, Nx=self.Nx, Ny=self.Ny, Lx=self.Lx, Ly=self.Ly, par=self.par,
                    still_depth=self.depth, bc=self.bc, inflow=self.inflow)

    @property
    def nd(self):
        return self.Nx * self.Ny * (self.P + 1) ** 2


def _face_points(P, Nx, Ny, Lx, Ly, wrapx, wrapy):
    from .solver import face_coords
    return face_coords(P, Nx, Ny, Lx, Ly, wrapx, wrapy)


def depth_arrays(P, Nx, Ny, Lx, Ly, fn, bc):
    """Still depth at nodes and on every interface line for H = fn(x, y)."""
    x, y = node_coords(P, Nx, Ny, Lx, Ly)
    fxx, fxy, fyx, fyy = _face_points(P, Nx, Ny, Lx, Ly, bc[0] == "periodic", bc[2] == "periodic")
    return {"Hnod": np.ascontiguousarray(fn(x, y), dtype=np.float64),
            "HfaceX": np.ascontiguousarray(fn(fxx, fxy), dtype=np.float64),
            "HfaceY": np.ascontiguousarray(fn(fyx, fyy), dtype=np.float64)}


def prim_to_state(h, u, v, w, p):
    """fromPrimitive (src/physics/sgn_flux2d.h:65-68), var-major."""
    return np.concatenate([h, h * u, h * v, h * w, h * p]).astype(np.float64)


# ---------------------------------------------------------------------------
def gate(P=4, Nx=8, Ny=6, bc=("periodic",) * 4):
    Lx, Ly = 10.0, 8.0
    par = Params(g=G, c=ALPHA * math.sqrt(G * 1.0), gamma=GAMMA)

    def bump(x, y):
        dx = x - 0.5 * Lx
        dy = y - 0.5 * Ly
        return 1.0 - 0.4 * np.exp(-(dx * dx + dy * dy) / 6.0)

    depth = depth_arrays(P, Nx, Ny, Lx, Ly, bump, bc)
    x, y = node_coords(P, Nx, Ny, Lx, Ly)
    kx, ky = 2.0 * math.pi / Lx, 2.0 * math.pi / Ly
    h = depth["Hnod"] + 0.04 * np.sin(kx * x) * np.cos(ky * y)
    u = 0.08 * np.cos(kx * x)
    v = -0.05 * np.sin(ky * y)
    w = 0.02 * np.sin(kx * x)
    p = 0.01 * np.cos(ky * y)
    return Case("gate", P, Nx, Ny, Lx, Ly, par, depth, tuple(bc), prim_to_state(h, u, v, w, p))


# ---------------------------------------------------------------------------
def solitary_state(x, H, h0, a, g, gamma, x0):
    """Travelling SGN solitary wave (h, u, w, p) riding on still depth H."""
    cs = math.sqrt(g * (h0 + a))
    kap = math.sqrt(3.0 * a / (4.0 * h0 * h0 * (h0 + a)))
    th = np.tanh(kap * (x - x0))
    s = 1.0 - th * th
    hs = h0 + a * s
    hx = -2.0 * kap * a * s * th
    hxx = 4.0 * kap * kap * a * s * th * th - 2.0 * kap * kap * a * s * s
    h = H + a * s
    u = cs * (1.0 - h0 / hs)
    w = -cs * h0 * hx / (2.0 * hs)
    p = -(cs * h0 / gamma) * (-0.5 * cs * h0 * (hxx * hs - hx * hx) / (hs * hs))
    return h, u, w, p


def solitary(Nx=128, Ny=64, walls=True, P=4):
    Lx, Ly = 64.0, 32.0
    h0, a = 1.0, 0.1
    par = Params(g=G, c=ALPHA * math.sqrt(G * h0), gamma=GAMMA)
    bc = ("periodic", "periodic", "wall", "wall") if walls else ("periodic",) * 4
    depth = depth_arrays(P, Nx, Ny, Lx, Ly, lambda x, y: np.full_like(x, h0), bc)
    x, _ = node_coords(P, Nx, Ny, Lx, Ly)
    h, u, w, p = solitary_state(x, depth["Hnod"], h0, a, G, GAMMA, x0=16.0)
    U0 = prim_to_state(h, u, np.zeros_like(h), w, p)
    return Case("solitary", P, Nx, Ny, Lx, Ly, par, depth, bc, U0,
                meta={"atol": 1e-8, "rtol": 1e-6, "cfl": 0.5, "steps": 200})


# ---------------------------------------------------------------------------
def slow_cp2(kH, gamma, alpha):
    """Slow-branch phase speed^2/(gH) of the relaxed system, Escalante-Dumbser-
    Castro (2019) eq. (15) with the + root."""
    J = 1.0 + kH * kH / (2.0 * gamma)
    mu = (kH / alpha) * (kH / alpha) / (2.0 * gamma)
    half = 0.5 * (J + mu)
    disc = half * half - mu
    return 1.0 / (half + math.sqrt(disc if disc > 0.0 else 0.0))


def wavenumber(omega, H, g=G, gamma=GAMMA, alpha=ALPHA):
    """k with k * Cp(k) = omega on the slow branch. The slow-branch frequency
    k*Cp(k) increases with k, so bisection on [0, k_hi] converges to the
    double nearest the root without a derivative."""
    def f(k):
        return k * math.sqrt(slow_cp2(k * H, gamma, alpha) * g * H) - omega
    lo, hi = 0.0, omega / math.sqrt(g * H)
    while f(hi) < 0.0:
        hi *= 2.0
        if hi * H > 1e6:
            # k*Cp(k) saturates at sqrt(2 gamma g / (H (1 + 1/alpha^2))) for large k
            raise ValueError(f"no slow-branch wave of frequency {omega} at depth {H}")
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mid <= lo or mid >= hi:
            break
        if f(mid) < 0.0:
            lo = mid
        else:
            hi = mid
    return hi if abs(f(hi)) < abs(f(lo)) else lo


def inflow_table(Ny, P, H_left, period, amp, g=G, gamma=GAMMA, alpha=ALPHA):
    """Left-side ghost coefficients of the linear slow-branch eigenmode
        eta = A sin(w t),  u = (Cp/H) eta,  p = (Cp^2/H - g) eta,
        w = -(gamma (Cp^2/H - g) A / (w H)) cos(w t)
    (x = 0, phase k x - w t shifted by a quarter period), as linearised
    conserved variables h = H + eta, hu = H u, hv = 0, hw = H w, hp = H p.
    Table layout [(ey*(P+1) + t)*15 + 3*var + {base, sin, cos}]."""
    omega = 2.0 * math.pi / period
    np_ = P + 1
    tab = np.zeros((Ny * np_, 5, 3))
    Hl = np.broadcast_to(np.asarray(H_left, dtype=np.float64), (Ny * np_,))
    for i in range(Ny * np_):
        H = float(Hl[i])
        k = wavenumber(omega, H, g, gamma, alpha)
        cp = omega / k
        Pamp = cp * cp / H - g
        tab[i, 0] = (H, amp, 0.0)
        tab[i, 1] = (0.0, cp * amp, 0.0)
        tab[i, 3] = (0.0, 0.0, -gamma * Pamp * amp / omega)
        tab[i, 4] = (0.0, H * Pamp * amp, 0.0)
    return tab.reshape(-1), omega


def shoal_focus(Nx=384, Ny=192, P=4, d_shoal=0.65, r_shoal=3.0):
    """Dispersive focusing over a submerged shoal (apps/shoal_focus_mpi.cu:
    98-137): still depth h0 - d exp(-((x - 18)^2 + (y - 10)^2) / (2 r^2)) on
    a periodic 40x20 m domain, a solitary wave of amplitude a = 0.3 over
    h0 = 1 m with its crest at x = 6 m riding on that depth, c = 3 sqrt(g h0).
    meta: the shoal centre, the amplitude, the end time 8 s and the 100
    samples of the reference's run."""
    Lx, Ly = 40.0, 20.0
    h0, a = 1.0, 0.3
    xs, ys = 18.0, 10.0
    par = Params(g=G, c=ALPHA * math.sqrt(G * h0), gamma=GAMMA)

    def depth_fn(x, y):
        dx, dy = x - xs, y - ys
        return h0 - d_shoal * np.exp(-(dx * dx + dy * dy) / (2.0 * r_shoal * r_shoal))

    bc = ("periodic",) * 4
    depth = depth_arrays(P, Nx, Ny, Lx, Ly, depth_fn, bc)
    x, _ = node_coords(P, Nx, Ny, Lx, Ly)
    h, u, w, p = solitary_state(x, depth["Hnod"], h0, a, G, GAMMA, x0=6.0)
    U0 = prim_to_state(h, u, np.zeros_like(h), w, p)
    return Case("shoal_focus", P, Nx, Ny, Lx, Ly, par, depth, bc, U0,
                meta={"atol": 1e-8, "rtol": 1e-6, "cfl": 0.5, "a": a, "x_shoal": xs,
                      "y_shoal": ys, "t_end": 8.0, "samples": 100, "d_shoal": d_shoal,
                      "r_shoal": r_shoal})


def tsunami_basin(Nx=1000, Ny=500, P=4, w_slope=5.0e4):
    """The basin-scale tsunami of apps/tsunami_margin_mpi.cu:130-186: a
    4000 x 2000 km periodic basin, 4000 m deep between two continental slopes
    (tanh of half-width w_slope centred at x = 1300 km and at a curved eastern
    margin 2700 km + 200 km cos(2 pi (y - ySrc) / Ly)) rising to a 150 m shelf;
    a 2 m elliptical Gaussian uplift (semi-axes 50 and 200 km) at (2000, 1000)
    km with the fluid at rest; c = 3 sqrt(g 4000). meta: the shelf window's
    margin curve, the five gauges on the eastern shelf (:187-192) and the
    adaptive tolerances."""
    Lx, Ly = 4.0e6, 2.0e6
    h_deep, h_shelf = 4000.0, 150.0
    xw, xe, bend = 1.3e6, 2.7e6, 2.0e5
    xs, ys, a_src, sx, sy = 2.0e6, 1.0e6, 2.0, 5.0e4, 2.0e5
    par = Params(g=G, c=ALPHA * math.sqrt(G * h_deep), gamma=GAMMA)

    def margin_x(y):
        return xe + bend * np.cos(2.0 * math.pi * (y - ys) / Ly)

    def depth_fn(x, y):
        s_ = 0.5 * (np.tanh((x - xw) / w_slope) - np.tanh((x - margin_x(y)) / w_slope))
        return h_shelf + (h_deep - h_shelf) * s_

    bc = ("periodic",) * 4
    depth = depth_arrays(P, Nx, Ny, Lx, Ly, depth_fn, bc)
    x, y = node_coords(P, Nx, Ny, Lx, Ly)
    dx, dy = (x - xs) / sx, (y - ys) / sy
    h = depth["Hnod"] + a_src * np.exp(-(dx * dx + dy * dy))
    z = np.zeros_like(h)
    gy = np.array([ys + (k - 2) * 0.18 * Ly for k in range(5)])
    gx = margin_x(gy) + 1.2e5
    return Case("tsunami_basin", P, Nx, Ny, Lx, Ly, par, depth, bc, prim_to_state(h, z, z, z, z),
                meta={"atol": 1e-8, "rtol": 1e-6, "cfl": 0.5, "margin_x": margin_x,
                      "gauges_xy": (gx, gy), "source": a_src})


# manufactured solution of apps/convergence2d.cpp:38-84 (phi = kx x + ky y - w t)
MMS = {"g": 9.81, "gamma": 1.5, "alpha": 3.0, "L": 1.0, "k": 2.0 * math.pi, "omega": 1.0,
       "ah": 0.10, "au": 0.15, "av": 0.12, "aw": 0.05, "ap": 0.03}


def mms_prims(x, y, t):
    """(h, u, v, w, p) of the manufactured solution and their x, y and t
    derivatives, as convergence2d.cpp:43-52 forms them."""
    m = MMS
    kx = ky = m["k"]
    om = m["omega"]
    ah, au, av, aw, ap = m["ah"], m["au"], m["av"], m["aw"], m["ap"]
    ph = kx * x + ky * y - om * t
    C, S = np.cos(ph), np.sin(ph)
    val = (1.0 + ah * C, au * C, av * C, aw * S, ap * C)
    dx = (-ah * kx * S, -au * kx * S, -av * kx * S, aw * kx * C, -ap * kx * S)
    dy = (-ah * ky * S, -au * ky * S, -av * ky * S, aw * ky * C, -ap * ky * S)
    dt = (ah * om * S, au * om * S, av * om * S, -aw * om * C, ap * om * S)
    return val, dx, dy, dt


def mms_forcing(x, y, t, c2):
    """G = d_t U* - R_cont(U*) of the manufactured solution, the five rows of
    convergence2d.cpp:54-84 term by term (shape (5, n))."""
    g, gamma = MMS["g"], MMS["gamma"]
    (h, u, v, w, p), (hx, ux, vx, wx, px), (hy, uy, vy, wy, py), (ht, ut, vt, wt, pt) = \
        mms_prims(x, y, t)
    dtU = (ht, ht * u + h * ut, ht * v + h * vt, ht * w + h * wt, ht * p + h * pt)
    dxF1 = hx * u + h * ux
    dxF2 = hx * u * u + 2.0 * h * u * ux + g * h * hx + hx * p + h * px
    dxF3 = hx * u * v + h * ux * v + h * u * vx
    dxF4 = hx * u * w + h * ux * w + h * u * wx
    dxF5 = (hx * u + h * ux) * (p + c2) + h * u * px
    dyG1 = hy * v + h * vy
    dyG2 = hy * u * v + h * uy * v + h * u * vy
    dyG3 = hy * v * v + 2.0 * h * v * vy + g * h * hy + hy * p + h * py
    dyG4 = hy * v * w + h * vy * w + h * v * wy
    dyG5 = (hy * v + h * vy) * (p + c2) + h * v * py
    R = (-(dxF1 + dyG1), -(dxF2 + dyG2), -(dxF3 + dyG3), -(dxF4 + dyG4) + gamma * p,
         -(dxF5 + dyG5) + c2 * u * hx + c2 * v * hy - 2.0 * c2 * w)
    return np.stack([dtU[k] - R[k] for k in range(5)])


def mms(N, P=4):
    """The manufactured-solution case at N x N elements on the unit square,
    flat bottom (H = 0), periodic; meta: the RK4 step of convergence2d.cpp:
    116-122 (dt = 0.15 dx_eff / sqrt(1.2 g + c^2), rounded to land on t = 0.1)."""
    m = MMS
    L = m["L"]
    c = m["alpha"] * math.sqrt(m["g"] * 1.0)
    par = Params(g=m["g"], c=c, gamma=m["gamma"])
    x, y = node_coords(P, N, N, L, L)
    (h, u, v, w, p), _, _, _ = mms_prims(x, y, 0.0)
    U0 = prim_to_state(h, u, v, w, p)
    t_end = 0.1
    dxeff = (L / N) / (2.0 * P + 1.0)
    sig = math.sqrt(m["g"] * 1.2 + c * c)
    dt = 0.15 * dxeff / sig
    n_step = int(math.ceil(t_end / dt))
    dt = t_end / n_step
    return Case(f"mms{N}", P, N, N, L, L, par, None, ("periodic",) * 4, U0,
                meta={"t_end": t_end, "dt": dt, "steps": n_step, "c2": c * c, "x": x, "y": y})


def shoal_params(seed):
    rng = np.random.default_rng(seed)
    return {
        "xc": float(rng.uniform(18.0, 30.0)),
        "yc": float(rng.uniform(14.0, 34.0)),
        "ax": float(rng.uniform(1.5, 6.0)),
        "ay": float(rng.uniform(1.5, 6.0)),
        "crest": float(rng.uniform(0.3, 0.9)),
    }


def slope_depth(x, h_deep=0.45, h_shelf=0.10, x_toe=40.0, slope=1.0 / 20.0):
    """1:20 linear slope from h_deep (x <= x_toe) to h_shelf."""
    return np.clip(h_deep - slope * (x - x_toe), h_shelf, h_deep)


def shoal(seed=0, Nx=1024, Ny=512, P=4, period=1.0, amp=0.01):
    Lx, Ly = 96.0, 48.0
    h_deep = 0.45
    par = Params(g=G, c=ALPHA * math.sqrt(G * h_deep), gamma=GAMMA)
    sp = shoal_params(seed)

    def depth_fn(x, y):
        base = slope_depth(x)
        local = slope_depth(np.full_like(x, sp["xc"]))
        r2 = ((x - sp["xc"]) / sp["ax"]) ** 2 + ((y - sp["yc"]) / sp["ay"]) ** 2
        return base - sp["crest"] * local * np.exp(-r2)

    bc = ("inflow", "wall", "wall", "wall")
    depth = depth_arrays(P, Nx, Ny, Lx, Ly, depth_fn, bc)
    # the ghost sits on the x = 0 interface line; its depth is the face depth there
    HfX = depth["HfaceX"].reshape(Ny, Nx + 1, P + 1)[:, 0, :].reshape(-1)
    table, omega = inflow_table(Ny, P, HfX, period, amp)
    Hn = depth["Hnod"]
    z = np.zeros_like(Hn)
    U0 = prim_to_state(Hn.copy(), z, z, z, z)
    return Case("shoal", P, Nx, Ny, Lx, Ly, par, depth, bc, U0, inflow=(table, omega),
                meta={"atol": 1e-8, "rtol": 1e-6, "cfl": 0.5, "steps": 500, "seed": seed,
                      "shoal": sp, "period": period, "amp": amp})


# ---------------------------------------------------------------------------
def point_noise(x, y, seed):
    """A seeded value in [-1, 1) per physical point: every sample of one point
    (the nodes of neighbouring elements, the face lines) gets the same value."""
    ix = np.round(np.asarray(x, dtype=np.float64) * 1048576.0).astype(np.int64).view(np.uint64)
    iy = np.round(np.asarray(y, dtype=np.float64) * 1048576.0).astype(np.int64).view(np.uint64)
    z = (ix * np.uint64(0x9E3779B97F4A7C15)) ^ (iy * np.uint64(0xC2B2AE3D27D4EB4F)) \
        ^ np.uint64((seed * 0x165667B19E3779F9) % (1 << 64))
    # splitmix64 finaliser
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def sweep_member(seed, Nx, P=4, period=1.0, amp=0.01):
    """One member of the batched shoal sweep: the production workload's domain,
    slope, inflow and walls at Nx x Nx/2 elements, a seeded round shoal (crest
    ratio in [0.3, 0.9], radius in [1.5, 6] m) and seeded per-point relative
    bathymetry noise of amplitude in [0, 2 %]."""
    Ny = Nx // 2
    Lx, Ly = 96.0, 48.0
    h_deep = 0.45
    par = Params(g=G, c=ALPHA * math.sqrt(G * h_deep), gamma=GAMMA)
    rng = np.random.default_rng(seed)
    sp = {"crest": float(rng.uniform(0.3, 0.9)), "r": float(rng.uniform(1.5, 6.0)),
          "xc": float(rng.uniform(18.0, 30.0)), "yc": float(rng.uniform(14.0, 34.0)),
          "noise": float(rng.uniform(0.0, 0.02))}

    def depth_fn(x, y):
        base = slope_depth(x)
        local = slope_depth(np.full_like(x, sp["xc"]))
        r2 = ((x - sp["xc"]) ** 2 + (y - sp["yc"]) ** 2) / sp["r"] ** 2
        H = base - sp["crest"] * local * np.exp(-r2)
        return H * (1.0 + sp["noise"] * point_noise(x, y, seed))

    bc = ("inflow", "wall", "wall", "wall")
    depth = depth_arrays(P, Nx, Ny, Lx, Ly, depth_fn, bc)
    HfX = depth["HfaceX"].reshape(Ny, Nx + 1, P + 1)[:, 0, :].reshape(-1)
    table, omega = inflow_table(Ny, P, HfX, period, amp)
    Hn = depth["Hnod"]
    z = np.zeros_like(Hn)
    U0 = prim_to_state(Hn.copy(), z, z, z, z)
    return Case(f"sweep{seed}", P, Nx, Ny, Lx, Ly, par, depth, bc, U0, inflow=(table, omega),
                meta={"atol": 1e-8, "rtol": 1e-6, "cfl": 0.5, "seed": seed, "shoal": sp})


def inversion(seed=0, Nx=256, Ny=128, P=4, n_gauges=20):
    """The gauge-inversion workload: 256x128 elements on 64x32 m, still depth
    H = 0.5 - (six seeded Gaussian mounds of 0.05-0.2 m), walls on every side,
    the fluid at rest under a free surface of eight seeded Gaussian humps
    (0.005-0.03 m high, 1-4 m wide), and n_gauges seeded gauge nodes
    (meta["gauges"]). The depth is a nodal field: the face depths follow from
    it as SGN2D.setDepth forms them (the mean of the two edge values).
    Tolerances atol 1e-10, rtol 1e-8: at 1e-8 / 1e-6 the controller accepts
    steps whose embedded error peaks near 20 times the tolerance at single
    nodes, and with those steps held fixed the run amplifies rounding noise
    in under-resolved grid modes, which swamps finite-difference checks of
    its gradient."""
    Lx, Ly = 64.0, 32.0
    h0 = 0.5
    par = Params(g=G, c=ALPHA * math.sqrt(G * h0), gamma=GAMMA)
    rng = np.random.default_rng(seed)
    mounds = [(rng.uniform(0.1, 0.9) * Lx, rng.uniform(0.1, 0.9) * Ly, rng.uniform(0.05, 0.2),
               rng.uniform(2.0, 6.0)) for _ in range(6)]
    humps = [(rng.uniform(0.15, 0.85) * Lx, rng.uniform(0.15, 0.85) * Ly,
              rng.uniform(0.005, 0.03), rng.uniform(1.0, 4.0)) for _ in range(8)]

    def bumps(x, y, items):
        out = np.zeros_like(x)
        for xc, yc, a, w in items:
            out = out + a * np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / (w * w))
        return out

    x, y = node_coords(P, Nx, Ny, Lx, Ly)
    theta = h0 - bumps(x, y, mounds)
    eta0 = bumps(x, y, humps)
    T = theta.reshape(Ny, Nx, P + 1, P + 1)
    fx = np.empty((Ny, Nx + 1, P + 1))
    fx[:, 1:Nx] = 0.5 * (T[:, :-1, :, P] + T[:, 1:, :, 0])
    fx[:, 0], fx[:, Nx] = T[:, 0, :, 0], T[:, -1, :, P]
    fy = np.empty((Ny + 1, Nx, P + 1))
    fy[1:Ny] = 0.5 * (T[:-1, :, P, :] + T[1:, :, 0, :])
    fy[0], fy[Ny] = T[0, :, 0, :], T[-1, :, P, :]
    depth = {"Hnod": np.ascontiguousarray(theta), "HfaceX": fx.ravel(), "HfaceY": fy.ravel()}
    z = np.zeros_like(theta)
    U0 = prim_to_state(theta + eta0, z, z, z, z)
    nd = theta.size
    gauges = np.sort(rng.choice(nd, size=n_gauges, replace=False)).astype(np.int64)
    return Case("inversion", P, Nx, Ny, Lx, Ly, par, depth, ("wall",) * 4, U0,
                meta={"atol": 1e-10, "rtol": 1e-8, "cfl": 0.5, "steps": 2000, "seed": seed,
                      "theta": theta, "eta0": eta0, "gauges": gauges, "checkpoints": 40})


def margin_params(seed):
    rng = np.random.default_rng(seed)
    return {
        "x_break": float(rng.uniform(0.55, 0.75)),   # shelf break, fraction of Lx
        "width": float(rng.uniform(5.0e3, 15.0e3)),  # slope half-width [m]
        "amp": float(rng.uniform(0.5, 2.0)),         # uplift [m]
        "xs": float(rng.uniform(0.2, 0.4)),          # uplift centre, fractions of Lx, Ly
        "ys": float(rng.uniform(0.3, 0.7)),
        "sx": float(rng.uniform(5.0e3, 2.0e4)),      # uplift semi-axes [m]
        "sy": float(rng.uniform(1.0e4, 4.0e4)),
    }


def margin(seed=0, Nx=2000, Ny=1000, P=4, rough=0.02):
    """The large single-grid margin workload, a regional cut of the
    co-seismic uplift over a continental margin of
    apps/tsunami_margin_mpi.cu:130-172: 200x100 km, still depth
    50 + 3950 * (1 - tanh((x - x_break) / width)) / 2 (abyssal plain 4000 m to
    the west, shelf 50 m to the east) times (1 + rough * seeded per-point noise),
    walls on the two x sides, periodic along the margin (y), the fluid at rest
    under a seeded elliptical Gaussian uplift over the deep basin;
    c = 3 sqrt(g 4000) from the greatest depth, as there."""
    Lx, Ly = 2.0e5, 1.0e5
    h_deep, h_shelf = 4000.0, 50.0
    par = Params(g=G, c=ALPHA * math.sqrt(G * h_deep), gamma=GAMMA)
    mp = margin_params(seed)
    xb = mp["x_break"] * Lx

    def depth_fn(x, y):
        s = 0.5 * (1.0 - np.tanh((x - xb) / mp["width"]))
        H = h_shelf + (h_deep - h_shelf) * s
        # the roughness is periodic in y like the domain: y = Ly is y = 0
        yw = y - Ly * np.floor(y / Ly + 1e-9)
        return H * (1.0 + rough * point_noise(x, yw, seed))

    bc = ("wall", "wall", "periodic", "periodic")
    depth = depth_arrays(P, Nx, Ny, Lx, Ly, depth_fn, bc)
    x, y = node_coords(P, Nx, Ny, Lx, Ly)
    dx = (x - mp["xs"] * Lx) / mp["sx"]
    dy = (y - mp["ys"] * Ly) / mp["sy"]
    Hn = depth["Hnod"]
    z = np.zeros_like(Hn)
    U0 = prim_to_state(Hn + mp["amp"] * np.exp(-(dx * dx + dy * dy)), z, z, z, z)
    return Case("margin", P, Nx, Ny, Lx, Ly, par, depth, bc, U0,
                meta={"atol": 1e-8, "rtol": 1e-6, "cfl": 0.5, "steps": 1000, "seed": seed,
                      "margin": mp, "rough": rough})


def margin_gauges(case, n=16, seed=0):
    """n seeded gauge nodes (ascending) inside the margin case's uplift
    ellipse, where a short run's surface moves: the stations of a source
    inversion over the first steps."""
    mp = case.meta["margin"]
    x, y = node_coords(case.P, case.Nx, case.Ny, case.Lx, case.Ly)
    dx = (x - mp["xs"] * case.Lx) / mp["sx"]
    dy = (y - mp["ys"] * case.Ly) / mp["sy"]
    inside = np.flatnonzero(dx * dx + dy * dy < 1.0)
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(inside, size=min(n, inside.size), replace=False)).astype(np.int64)


def smooth_direction(case, rng, n=4, amp=0.01, near=None, scale=1.0):
    """A seeded smooth nodal field (n Gaussians of width 2-6 m times scale and
    amplitude up to amp, continuous across elements): a perturbation direction
    for finite-difference checks of gradients of runs whose recorded step
    sizes are held fixed, where only resolved (smooth) perturbations stay
    bounded. near: node indices (e.g. gauges) to centre the Gaussians on
    (within 1 m times scale), so that the direction reaches them within a
    short run. scale = 1 suits grids of metres (the shoal, the inversion);
    the margin's 100 m elements take kilometres (scale 1000)."""
    x, y = node_coords(case.P, case.Nx, case.Ny, case.Lx, case.Ly)
    out = np.zeros_like(x)
    for _ in range(n):
        if near is not None:
            k = int(near[rng.integers(len(near))])
            xc = x[k] + scale * rng.uniform(-1.0, 1.0)
            yc = y[k] + scale * rng.uniform(-1.0, 1.0)
        else:
            xc, yc = rng.uniform(0.0, case.Lx), rng.uniform(0.0, case.Ly)
        a, w = rng.uniform(-amp, amp), scale * rng.uniform(2.0, 6.0)
        out = out + a * np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / (w * w))
    return out


def sweep_sizes(n=64, seed=0, nx_min=128, nx_max=768):
    """Seeded element widths of the sweep's members (even, Ny = Nx / 2)."""
    rng = np.random.default_rng(seed)
    return [int(v) * 2 for v in rng.integers(nx_min // 2, nx_max // 2 + 1, size=n)]


# ---------------------------------------------------------------------------
def random_state(nd, seed, h0=1.0, amp=1.0):
    """Node-wise random state with h in h0*[0.6, 1.4] (exercises every branch
    of the arithmetic; not a smooth field)."""
    rng = np.random.default_rng(seed)
    h = h0 * (1.0 + 0.4 * amp * rng.uniform(-1.0, 1.0, nd))
    u = 0.5 * amp * rng.uniform(-1.0, 1.0, nd)
    v = 0.5 * amp * rng.uniform(-1.0, 1.0, nd)
    w = 0.1 * amp * rng.uniform(-1.0, 1.0, nd)
    p = 0.2 * amp * rng.uniform(-1.0, 1.0, nd)
    return prim_to_state(h, u, v, w, p)


def random_depth(P, Nx, Ny, Lx, Ly, seed, bc):
    """Smooth random still depth in [0.5, 1.5] (a few Gaussian bumps)."""
    rng = np.random.default_rng(seed)
    cx = rng.uniform(0.0, Lx, 4)
    cy = rng.uniform(0.0, Ly, 4)
    amp = rng.uniform(-0.2, 0.2, 4)
    wid = rng.uniform(0.1, 0.3, 4) * min(Lx, Ly)

    def fn(x, y):
        out = np.ones_like(x)
        for i in range(4):
            out = out + amp[i] * np.exp(-((x - cx[i]) ** 2 + (y - cy[i]) ** 2) / (wid[i] ** 2))
        return out

    return depth_arrays(P, Nx, Ny, Lx, Ly, fn, bc)
