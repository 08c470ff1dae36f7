"""Python mirror of the reference's solver API over libsgnx.

`SGN2D` takes the constructor arguments of sgn::SGN2D (src/solver/sgn2d.h:57-59)
and exposes its methods (nDof, nTotal, nodeCoords, makeState, rhs, rhsFn,
maxWaveSpeed, cflTimeStep, mass, energy) together with the device-class methods
of SGN2DCuda (upload, download, rhsOnce, integrateAdaptive;
src/cuda/sgn2d_cuda.h:41-47). Every evaluation runs on the GPU through the
C-ABI; setup (operator tables, coordinates, still-depth samples) is host work,
as in the reference constructor (src/solver/sgn2d.cpp:30-84).

Sides: the reference 2D mesh is periodic everywhere. `bc` adds reflective walls
and a left-side inflow ghost; with the default all-periodic sides the solver is
a drop-in for the reference path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

NVAR = 5
TIMING_MODES = 4          # launch kinds of sgnx_rhs_timing (SGNX_TIMING_MODES)
H, HU, HV, HW, HP = range(5)

_BC_CODES = {"periodic": L.BC_PERIODIC, "wall": L.BC_WALL, "inflow": L.BC_INFLOW}


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


class FROperators1D:
    """sgn::FROperators1D(P, GaussLobatto) tables (src/core/fr_operators.h:57-115)."""

    def __init__(self, P):
        lib = L.load()
        n = P + 1
        self.P = P
        self._x = np.zeros(n)
        self._D = np.zeros(n * n)
        self._lL = np.zeros(n)
        self._lR = np.zeros(n)
        self._gLp = np.zeros(n)
        self._gRp = np.zeros(n)
        self._qw = np.zeros(n)
        L.check(lib.sgnx_fr_operators(P, _dp(self._x), _dp(self._D), _dp(self._lL),
                                      _dp(self._lR), _dp(self._gLp), _dp(self._gRp),
                                      _dp(self._qw)))

    def nodes(self):
        return self._x

    def D(self):
        return self._D

    def lL(self):
        return self._lL

    def lR(self):
        return self._lR

    def gLprime(self):
        return self._gLp

    def gRprime(self):
        return self._gRp

    def weights(self):
        return self._qw


def node_coords(P, Nx, Ny, Lx, Ly):
    """(x, y) of every node, gnode = e*np^2 + a + np*b (sgn2d.cpp:89-102)."""
    nd = Nx * Ny * (P + 1) ** 2
    x = np.zeros(nd)
    y = np.zeros(nd)
    L.check(L.load().sgnx_node_coords(P, Nx, Ny, Lx, Ly, _dp(x), _dp(y)))
    return x, y


def face_coords(P, Nx, Ny, Lx, Ly, wrapx, wrapy):
    """Sample points of the x- and y-interface still depths (sgn2d.cpp:75-83)."""
    np_ = P + 1
    nfx = (Nx + 1) * Ny * np_
    nfy = Nx * (Ny + 1) * np_
    fxx, fxy, fyx, fyy = (np.zeros(nfx), np.zeros(nfx), np.zeros(nfy), np.zeros(nfy))
    L.check(L.load().sgnx_face_coords(P, Nx, Ny, Lx, Ly, int(wrapx), int(wrapy), _dp(fxx),
                                      _dp(fxy), _dp(fyx), _dp(fyy)))
    return fxx, fxy, fyx, fyy


def still_depth_arrays(P, Nx, Ny, Lx, Ly, depth, bc):
    """Hnod, HfaceX, HfaceY of a still-depth callable depth(x, y) (vectorised)."""
    wrapx = bc[0] == "periodic"
    wrapy = bc[2] == "periodic"
    x, y = node_coords(P, Nx, Ny, Lx, Ly)
    fxx, fxy, fyx, fyy = face_coords(P, Nx, Ny, Lx, Ly, wrapx, wrapy)
    Hn = np.ascontiguousarray(np.asarray(depth(x, y), dtype=np.float64))
    Hfx = np.ascontiguousarray(np.asarray(depth(fxx, fxy), dtype=np.float64))
    Hfy = np.ascontiguousarray(np.asarray(depth(fyx, fyy), dtype=np.float64))
    return Hn, Hfx, Hfy


class SGN2D:
    """Device-backed counterpart of sgn::SGN2D + sgn::SGN2DCuda.

    still_depth: None (flat H == 0), a vectorised callable H(x, y), or a dict
    {"Hnod", "HfaceX", "HfaceY"} of precomputed arrays (interface layout of
    include/sgnx.h).
    bc: four side kinds (left, right, bottom, top) out of "periodic", "wall",
    "inflow"; inflow (left side) needs `inflow=(table, omega)`.
    """

    def __init__(self, P, Nx, Ny, Lx, Ly, par, still_depth=None,
                 bc=("periodic", "periodic", "periodic", "periodic"), inflow=None):
        if Nx < 2 or Ny < 2:
            raise ValueError("SGN2D: Nx,Ny >= 2")
        self.lib = L.load()
        self.P, self.Nx, self.Ny, self.Lx, self.Ly = P, Nx, Ny, float(Lx), float(Ly)
        self.par = par if isinstance(par, L.Params) else L.Params(*par)
        self.bc = tuple(bc)
        self.ops = FROperators1D(P)
        np_ = P + 1
        self._Nd = Nx * Ny * np_ * np_
        if still_depth is None:
            self.Hnod = self.HfaceX = self.HfaceY = None
        elif isinstance(still_depth, dict):
            self.Hnod = np.ascontiguousarray(still_depth["Hnod"], dtype=np.float64)
            self.HfaceX = np.ascontiguousarray(still_depth["HfaceX"], dtype=np.float64)
            self.HfaceY = np.ascontiguousarray(still_depth["HfaceY"], dtype=np.float64)
        else:
            self.Hnod, self.HfaceX, self.HfaceY = still_depth_arrays(P, Nx, Ny, Lx, Ly,
                                                                     still_depth, self.bc)
        d = L.Desc()
        d.P, d.Nx, d.Ny, d.Lx, d.Ly = P, Nx, Ny, self.Lx, self.Ly
        d.par = self.par
        o = self.ops
        d.D, d.lL, d.lR = _dp(o.D()), _dp(o.lL()), _dp(o.lR())
        d.gLp, d.gRp, d.qw = _dp(o.gLprime()), _dp(o.gRprime()), _dp(o.weights())
        d.Hnod, d.HfaceX, d.HfaceY = _dp(self.Hnod), _dp(self.HfaceX), _dp(self.HfaceY)
        d.Hx = d.Hy = None
        for k in range(4):
            d.bc[k] = _BC_CODES[self.bc[k]]
        self._inflow = None
        if inflow is not None:
            table, omega = inflow
            self._inflow = np.ascontiguousarray(table, dtype=np.float64)
            d.inflow = _dp(self._inflow)
            d.inflow_omega = float(omega)
        h = C.c_void_p()
        L.check(self.lib.sgnx_create(C.byref(d), C.byref(h)))
        self.h = h

    # ---- sizes and setup ----------------------------------------------------
    def order(self):
        return self.P

    def nElem(self):
        return self.Nx * self.Ny

    def nDof(self):
        return self._Nd

    def nTotal(self):
        return NVAR * self._Nd

    def nodeCoords(self):
        return node_coords(self.P, self.Nx, self.Ny, self.Lx, self.Ly)

    def makeState(self, h, u, v, w, p):
        """(h, u, v, w, p) -> var-major conserved state (sgn2d.cpp:107-120)."""
        h = np.asarray(h, dtype=np.float64)
        U = np.empty(self.nTotal())
        Nd = self._Nd
        U[0:Nd] = h
        U[Nd:2 * Nd] = h * np.asarray(u, dtype=np.float64)
        U[2 * Nd:3 * Nd] = h * np.asarray(v, dtype=np.float64)
        U[3 * Nd:4 * Nd] = h * np.asarray(w, dtype=np.float64)
        U[4 * Nd:5 * Nd] = h * np.asarray(p, dtype=np.float64)
        return U

    def stillDepthGrad(self):
        Hx = np.zeros(self._Nd)
        Hy = np.zeros(self._Nd)
        L.check(self.lib.sgnx_bathymetry(self.h, _dp(Hx), _dp(Hy)), self.h)
        return Hx, Hy

    # ---- state transfer ------------------------------------------------------
    def _vec(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.size != self.nTotal():
            raise ValueError("SGN2D: size mismatch")
        return U

    def upload(self, U):
        U = self._vec(U)
        L.check(self.lib.sgnx_upload(self.h, _dp(U), U.size), self.h)

    def download(self):
        U = np.empty(self.nTotal())
        L.check(self.lib.sgnx_download(self.h, _dp(U), U.size), self.h)
        return U

    def state_ptr(self):
        return self.lib.sgnx_state_ptr(self.h)

    def set_stream(self, stream_ptr):
        L.check(self.lib.sgnx_set_stream(self.h, C.c_void_p(stream_ptr)), self.h)

    # ---- evaluations (GPU) -----------------------------------------------------
    def rhs(self, U, t=0.0):
        """dU = R(U, t) (sgn2d.cpp:125-290); t only enters an inflow ghost."""
        U = self._vec(U)
        dU = np.empty_like(U)
        L.check(self.lib.sgnx_rhs(self.h, U.ctypes.data, dU.ctypes.data, U.size, float(t),
                                  L.MEM_HOST), self.h)
        return dU

    def rhs_device(self, U_ptr, dU_ptr, t=0.0):
        """dU = R(U, t) on device pointers (e.g. torch tensors' data_ptr())."""
        L.check(self.lib.sgnx_rhs(self.h, C.c_void_p(U_ptr), C.c_void_p(dU_ptr), self.nTotal(),
                                  float(t), L.MEM_DEVICE), self.h)

    def rhsFn(self):
        return lambda t, U: self.rhs(U, t)

    def rhsOnce(self, t=0.0):
        dU = np.empty(self.nTotal())
        L.check(self.lib.sgnx_rhs_once(self.h, _dp(dU), dU.size, float(t)), self.h)
        return dU

    def integrateAdaptive(self, t0, tEnd, dt0, opt=None, step_log_cap=0):
        """Device-resident BS3(2) on the uploaded state (rk.cpp:67-164)."""
        opt = opt if opt is not None else L.RkOpts()
        st = L.RkStats()
        log = np.zeros(2 * step_log_cap) if step_log_cap else None
        L.check(self.lib.sgnx_integrate(self.h, float(t0), float(tEnd), float(dt0), C.byref(opt),
                                        C.byref(st), _dp(log), int(step_log_cap)), self.h)
        if step_log_cap:
            return st, log.reshape(-1, 2)
        return st

    def rk4Step(self, t, dt):
        L.check(self.lib.sgnx_rk4_step(self.h, float(t), float(dt)), self.h)

    def bs3Step(self, t, dt):
        L.check(self.lib.sgnx_bs3_step(self.h, float(t), float(dt)), self.h)

    def _diag(self, fn, U, *extra):
        if U is not None:
            self.upload(U)
        out = C.c_double()
        L.check(fn(self.h, *extra, C.byref(out)), self.h)
        return out.value

    def maxWaveSpeed(self, U=None):
        return self._diag(self.lib.sgnx_max_wave_speed, U)

    def cflTimeStep(self, U=None, cfl=0.5):
        return self._diag(self.lib.sgnx_cfl_time_step, U, C.c_double(cfl))

    def mass(self, U=None):
        return self._diag(self.lib.sgnx_mass, U)

    def energy(self, U=None):
        return self._diag(self.lib.sgnx_energy, U)

    def set_rhs_kernel(self, variant, rows=16):
        """0: one workgroup per 5-element row tile; 1: row sweep of `rows` rows."""
        L.check(self.lib.sgnx_set_rhs_kernel(self.h, int(variant), int(rows)), self.h)

    def rhs_kernel(self):
        """(variant, rows, waves3) in use: the kernel variant, the row sweep's
        tile height and its waves mask (bit k: launch kind k at three waves)."""
        v, r, w = C.c_int(), C.c_int(), C.c_int()
        L.check(self.lib.sgnx_get_rhs_kernel(self.h, C.byref(v), C.byref(r), C.byref(w)), self.h)
        return v.value, r.value, w.value

    # ---- diagnostics (eta = h - H) ---------------------------------------------
    def etaMax(self, window=None):
        """(max eta, node index) over the nodes strictly inside window =
        (x0, x1, y0, y1) (None: all nodes); the lowest index among ties."""
        x0, x1, y0, y1 = window if window is not None else (-np.inf, np.inf, -np.inf, np.inf)
        v, i = C.c_double(), C.c_longlong()
        L.check(self.lib.sgnx_eta_max(self.h, float(x0), float(x1), float(y0), float(y1),
                                      C.byref(v), C.byref(i)), self.h)
        return v.value, i.value

    def gauges(self, nodes, noalloc=False):
        """eta at the given node indices of the resident state (noalloc: through
        sgnx_gauges_noalloc, at most SGNX_GAUGES_NOALLOC_MAX nodes)."""
        nodes = np.ascontiguousarray(nodes, dtype=np.int64)
        out = np.zeros(nodes.size)
        fn = self.lib.sgnx_gauges_noalloc if noalloc else self.lib.sgnx_gauges
        L.check(fn(self.h, nodes.ctypes.data_as(C.POINTER(C.c_longlong)), int(nodes.size),
                   _dp(out)), self.h)
        return out

    def elementMeans(self):
        """Element means of eta, u, v, p, H: array (5, Ny, Nx)."""
        out = np.zeros(5 * self.Nx * self.Ny)
        L.check(self.lib.sgnx_element_means(self.h, _dp(out)), self.h)
        return out.reshape(5, self.Ny, self.Nx)

    def setGauges(self, nodes, cap):
        """Record eta at `nodes` after every accepted step of integrateAdaptive
        (up to cap rows; clears the log)."""
        nodes = np.ascontiguousarray(nodes, dtype=np.int64)
        L.check(self.lib.sgnx_set_gauges(self.h, nodes.ctypes.data_as(C.POINTER(C.c_longlong)),
                                         int(nodes.size), int(cap)), self.h)
        self._nG = int(nodes.size)
        self._gcap = int(cap)

    def gaugeValues(self):
        """Number of values in the gauge log (rows recorded x gauges): the size
        a gauge-log cotangent must have."""
        return self.gaugeLog()[1].size

    def gaugeLog(self):
        """(times, eta[rows, n]) recorded since setGauges."""
        n, cap = getattr(self, "_nG", 0), getattr(self, "_gcap", 0)
        eta = np.zeros(max(n * cap, 1))
        t = np.zeros(max(cap, 1))
        rows = C.c_long()
        L.check(self.lib.sgnx_gauge_log(self.h, _dp(eta), _dp(t), int(cap), C.byref(rows)),
                self.h)
        r = rows.value
        return t[:r].copy(), eta[:r * n].reshape(r, n).copy()

    def autotune(self, reps=3):
        """Tile height and waves per SIMD of the row sweep chosen by timing the
        launches of an adaptive step on the resident state; returns them."""
        r, w = C.c_int(), C.c_int()
        L.check(self.lib.sgnx_autotune(self.h, int(reps), C.byref(r), C.byref(w)), self.h)
        return r.value, w.value

    # ---- discrete adjoint ------------------------------------------------------
    def setDepth(self, theta, device=False):
        """Still depth from the nodal field theta (Nd values): Hnod = theta,
        its element-line gradients, and face depths = the mean of the two edge
        values meeting at each interface line node (one interior value on a
        non-periodic side). device=True: theta is a device address."""
        if device:
            ptr = int(theta)
        else:
            theta = np.ascontiguousarray(theta, dtype=np.float64)
            if theta.size != self._Nd:
                raise ValueError("setDepth: size mismatch")
            ptr = theta.ctypes.data
        L.check(self.lib.sgnx_set_depth(self.h, C.c_void_p(ptr), self._Nd,
                                        L.MEM_DEVICE if device else L.MEM_HOST), self.h)

    def rhs_vjp(self, U, W, t=0.0, theta=True):
        """(Ubar, thetaBar) = ((dR/dU)^T W, (dR/dtheta)^T W) at (U, t)."""
        U = self._vec(U)
        W = self._vec(W)
        Ub = np.empty_like(U)
        tb = np.empty(self._Nd) if theta else None
        L.check(self.lib.sgnx_rhs_vjp(self.h, U.ctypes.data, W.ctypes.data, Ub.ctypes.data,
                                      tb.ctypes.data if theta else None, U.size, float(t),
                                      L.MEM_HOST), self.h)
        return (Ub, tb) if theta else Ub

    def rhs_vjp_device(self, U_ptr, W_ptr, Ubar_ptr, theta_bar_ptr=None, t=0.0):
        L.check(self.lib.sgnx_rhs_vjp(self.h, C.c_void_p(U_ptr), C.c_void_p(W_ptr),
                                      C.c_void_p(Ubar_ptr),
                                      C.c_void_p(theta_bar_ptr) if theta_bar_ptr else None,
                                      self.nTotal(), float(t), L.MEM_DEVICE), self.h)

    def tapeBegin(self, max_snapshots=1 << 20):
        """Record the accepted steps and state snapshots of the following
        integrateAdaptive calls (at most max_snapshots states held)."""
        L.check(self.lib.sgnx_tape_begin(self.h, int(max_snapshots)), self.h)

    def tapeBeginPlan(self, steps, states):
        """Record like tapeBegin, for a run of `steps` accepted steps whose
        reverse sweep holds at most `states` states (snapshots included):
        snapshots at the first sweep of the binomial schedule, so the sweep
        recomputes the fewest steps that budget allows."""
        L.check(self.lib.sgnx_tape_begin_plan(self.h, int(steps), int(states)), self.h)

    def tapeSlots(self, slots):
        """States the reverse sweep may hold besides the snapshots (-1: all a
        segment needs; fewer: binomial recomputation, same result)."""
        L.check(self.lib.sgnx_tape_set_slots(self.h, int(slots)), self.h)

    def tapeKeepStages(self, on=True):
        """Keep each taped step's stage-2 and stage-3 inputs (two more states
        per step, written by the forward's own stage launches), so the reverse
        sweep recomputes no stage; dropped once the snapshot budget is full.
        Set before the run; the adjoint is the same bit for bit."""
        L.check(self.lib.sgnx_tape_keep_stages(self.h, 1 if on else 0), self.h)

    def tapeStages(self):
        """Steps whose stage inputs the tape keeps."""
        a = C.c_long()
        L.check(self.lib.sgnx_tape_stages(self.h, C.byref(a)), self.h)
        return a.value

    def adjointStats(self):
        """(steps recomputed, slots held at once) of the last adjoint()."""
        a, b = C.c_long(), C.c_long()
        L.check(self.lib.sgnx_adjoint_stats(self.h, C.byref(a), C.byref(b)), self.h)
        return a.value, b.value

    def tapeEnd(self):
        L.check(self.lib.sgnx_tape_end(self.h), self.h)

    def tapeInfo(self):
        """(steps recorded, snapshot stride, snapshots held)."""
        a, b, c = C.c_long(), C.c_long(), C.c_long()
        L.check(self.lib.sgnx_tape_info(self.h, C.byref(a), C.byref(b), C.byref(c)), self.h)
        return a.value, b.value, c.value

    def tapeSteps(self):
        """(t, dt) of every recorded step."""
        n = self.tapeInfo()[0]
        t, dt = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        got = C.c_long()
        L.check(self.lib.sgnx_tape_steps(self.h, _dp(t), _dp(dt), n, C.byref(got)), self.h)
        return t[:n].copy(), dt[:n].copy()

    def tapeReplay(self):
        """The taped steps (same t, dt) from the resident state, with gauge recording."""
        L.check(self.lib.sgnx_tape_replay(self.h), self.h)

    def adjoint(self, Ubar_final=None, gauge_bar=None, theta=True):
        """Cotangents (U0bar, thetaBar) of the taped run's initial state and still
        depth for the final-state cotangent Ubar_final and the gauge-log
        cotangent gauge_bar (rows x gauges, as gaugeLog)."""
        Uf = self._vec(Ubar_final) if Ubar_final is not None else None
        gb = None
        if gauge_bar is not None:
            gb = np.ascontiguousarray(gauge_bar, dtype=np.float64)
            want = self.gaugeValues()
            if gb.size != want:
                # the library reads rows x gauges values from this buffer
                raise ValueError(f"gauge_bar has {gb.size} values, the gauge log {want}")
        U0b = np.empty(self.nTotal())
        tb = np.empty(self._Nd) if theta else None
        L.check(self.lib.sgnx_adjoint(self.h, Uf.ctypes.data if Uf is not None else None,
                                      gb.ctypes.data if gb is not None else None,
                                      U0b.ctypes.data, tb.ctypes.data if theta else None,
                                      L.MEM_HOST), self.h)
        return (U0b, tb) if theta else U0b

    def adjoint_device(self, Ubar_final_ptr, gauge_bar_ptr, U0bar_ptr, theta_bar_ptr=None):
        """adjoint() on device addresses (0 / None for an absent input)."""
        L.check(self.lib.sgnx_adjoint(self.h, C.c_void_p(Ubar_final_ptr or None),
                                      C.c_void_p(gauge_bar_ptr or None), C.c_void_p(U0bar_ptr),
                                      C.c_void_p(theta_bar_ptr or None), L.MEM_DEVICE), self.h)

    def enable_timing(self, on=True):
        L.check(self.lib.sgnx_enable_timing(self.h, int(on)), self.h)

    def rhs_timing(self):
        """[(launches, total device ms)] of RHS launches since timing was enabled,
        by launch kind: plain state, stage 2 (u + a*k), stage 3 (u + a*k writing
        the candidate), k4 with the embedded error."""
        n = (C.c_long * TIMING_MODES)()
        ms = (C.c_double * TIMING_MODES)()
        L.check(self.lib.sgnx_rhs_timing(self.h, n, ms), self.h)
        return [(n[m], ms[m]) for m in range(TIMING_MODES)]

    def close(self):
        if getattr(self, "h", None):
            self.lib.sgnx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
