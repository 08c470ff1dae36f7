"""ctypes access to the CPU oracle libraries (test infrastructure).

  oracle   oracle/liboracle.so       the C++ restatement of the reference path
  ref      oracle/_ref/libsgnref.so  the reference CPU solver built from its own
                                     sources (present where the checkout was at
                                     build time; tests that need it skip without)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libsgnref.so")

_dp = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_ulonglong)


def dp(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class OrcDesc(C.Structure):
    _fields_ = [("P", C.c_int), ("Nx", C.c_int), ("Ny", C.c_int), ("gauss_points", C.c_int),
                ("Lx", C.c_double), ("Ly", C.c_double), ("g", C.c_double), ("c", C.c_double),
                ("gamma", C.c_double), ("nManning", C.c_double), ("bc", C.c_int * 4),
                ("Hnod", _dp), ("HfaceX", _dp), ("HfaceY", _dp), ("inflow", _dp),
                ("inflow_omega", C.c_double)]


class OrcOpts(C.Structure):
    _fields_ = [("atol", C.c_double), ("rtol", C.c_double), ("dtMin", C.c_double),
                ("dtMax", C.c_double), ("safety", C.c_double), ("facMin", C.c_double),
                ("facMax", C.c_double), ("max_accepted", C.c_long)]

    def __init__(self, atol=1e-8, rtol=1e-6, dtMin=1e-13, dtMax=1e30, safety=0.9, facMin=0.2,
                 facMax=5.0, max_accepted=0):
        super().__init__(atol, rtol, dtMin, dtMax, safety, facMin, facMax, max_accepted)

    def as7(self):
        return np.array([self.atol, self.rtol, self.dtMin, self.dtMax, self.safety, self.facMin,
                         self.facMax])


class OrcStats(C.Structure):
    _fields_ = [("accepted", C.c_long), ("rejected", C.c_long), ("rhsEvals", C.c_long),
                ("tReached", C.c_double), ("dtLast", C.c_double), ("maxNormLast", C.c_double),
                ("aborted", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_orc = None
_ref = None


def oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"{ORACLE_PATH} missing: run `make -C oracle`")
        lib = C.CDLL(ORACLE_PATH)
        lib.orc_create.restype = C.c_void_p
        lib.orc_create.argtypes = [C.POINTER(OrcDesc), C.c_char_p, C.c_size_t]
        lib.orc_destroy.argtypes = [C.c_void_p]
        lib.orc_ndof.restype = C.c_long
        lib.orc_ndof.argtypes = [C.c_void_p]
        lib.orc_bathymetry.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.orc_rhs.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
        for f in ("orc_max_wave_speed_state", "orc_mass", "orc_energy"):
            getattr(lib, f).restype = C.c_double
            getattr(lib, f).argtypes = [C.c_void_p, _dp]
        lib.orc_cfl_time_step.restype = C.c_double
        lib.orc_cfl_time_step.argtypes = [C.c_void_p, _dp, C.c_double]
        lib.orc_rk4_step.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp]
        lib.orc_bs3_step.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp]
        lib.orc_integrate.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_double,
                                      C.POINTER(OrcOpts), C.POINTER(OrcStats), _dp, C.c_long,
                                      _dp, _u64p, C.c_long]
        lib.orc_fnv64.restype = C.c_ulonglong
        lib.orc_fnv64.argtypes = [_dp, C.c_size_t]
        lib.orc_fr_operators.argtypes = [C.c_int, C.c_int] + [_dp] * 7
        lib.orc_legendre.restype = C.c_double
        lib.orc_legendre.argtypes = [C.c_int, C.c_double]
        lib.orc_legendre_deriv.restype = C.c_double
        lib.orc_legendre_deriv.argtypes = [C.c_int, C.c_double]
        for f in ("orc_flux_x", "orc_flux_y"):
            getattr(lib, f).argtypes = [_dp, C.c_double, _dp, _dp]
        lib.orc_source.argtypes = [_dp, C.c_double, C.c_double, C.c_double, _dp, _dp]
        lib.orc_ncp_x.argtypes = [_dp, C.c_double, _dp, _dp]
        lib.orc_ncp_y.argtypes = [_dp, C.c_double, _dp, _dp]
        lib.orc_signal_speed.restype = C.c_double
        lib.orc_signal_speed.argtypes = [_dp, _dp]
        lib.orc_max_wave_speed.restype = C.c_double
        lib.orc_max_wave_speed.argtypes = [_dp, C.c_int, _dp]
        lib.orc_rusanov.argtypes = [_dp, _dp, C.c_int, C.c_double, _dp, _dp]
        lib.orc_energy_density.restype = C.c_double
        lib.orc_energy_density.argtypes = [_dp, C.c_double, _dp]
        lib.orc_node_coords.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _dp,
                                        _dp]
        lib.orc_face_coords.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        lib.orc_integrate_ode.argtypes = [C.c_int, C.c_double, _dp, C.c_int, C.c_double,
                                          C.c_double, C.c_double, C.POINTER(OrcOpts),
                                          C.POINTER(OrcStats)]
        lib.orc_rk4_step_ode.argtypes = [C.c_int, C.c_double, _dp, C.c_int, C.c_double,
                                         C.c_double]
        lib.orc_bs3_step_ode.argtypes = [C.c_int, C.c_double, _dp, C.c_int, C.c_double,
                                         C.c_double]
        _orc = lib
    return _orc


def ref_available():
    return os.path.exists(REF_PATH)


DEPTH_CB = C.CFUNCTYPE(C.c_double, C.c_double, C.c_double)


def ref():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_PATH)
        lib.ref_fr_operators.argtypes = [C.c_int, C.c_int] + [_dp] * 7
        for f in ("ref_flux_x", "ref_flux_y"):
            getattr(lib, f).argtypes = [_dp, C.c_double, _dp, _dp]
        lib.ref_source.argtypes = [_dp, C.c_double, C.c_double, C.c_double, _dp, _dp]
        lib.ref_ncp_x.argtypes = [_dp, C.c_double, _dp, _dp]
        lib.ref_ncp_y.argtypes = [_dp, C.c_double, _dp, _dp]
        lib.ref_signal_speed.restype = C.c_double
        lib.ref_signal_speed.argtypes = [_dp, _dp]
        lib.ref_max_wave_speed.restype = C.c_double
        lib.ref_max_wave_speed.argtypes = [_dp, C.c_int, _dp]
        lib.ref_rusanov.argtypes = [_dp, _dp, C.c_int, C.c_double, _dp, _dp]
        lib.ref_energy_density.restype = C.c_double
        lib.ref_energy_density.argtypes = [_dp, C.c_double, _dp]
        lib.ref_create.restype = C.c_void_p
        lib.ref_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _dp,
                                   DEPTH_CB, C.c_int, C.c_char_p, C.c_size_t]
        lib.ref_destroy.argtypes = [C.c_void_p]
        lib.ref_ndof.restype = C.c_long
        lib.ref_ndof.argtypes = [C.c_void_p]
        lib.ref_node_coords.argtypes = [C.c_void_p, _dp, _dp]
        lib.ref_bathymetry.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        lib.ref_rhs.argtypes = [C.c_void_p, _dp, _dp]
        for f in ("ref_max_wave_speed_state", "ref_mass", "ref_energy"):
            getattr(lib, f).restype = C.c_double
            getattr(lib, f).argtypes = [C.c_void_p, _dp]
        lib.ref_cfl_time_step.restype = C.c_double
        lib.ref_cfl_time_step.argtypes = [C.c_void_p, _dp, C.c_double]
        lib.ref_rk4_step.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp]
        lib.ref_bs3_step.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp]
        lib.ref_integrate.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_double, _dp,
                                      C.c_long, _dp, _u64p, C.c_long, C.POINTER(C.c_long), _dp,
                                      C.POINTER(C.c_int)]
        lib.ref_integrate_ode.argtypes = [C.c_int, C.c_double, _dp, C.c_int, C.c_double,
                                          C.c_double, C.c_double, _dp, C.POINTER(C.c_long), _dp,
                                          C.POINTER(C.c_int)]
        lib.ref_rk4_step_ode.argtypes = [C.c_int, C.c_double, _dp, C.c_int, C.c_double,
                                         C.c_double]
        lib.ref_bs3_step_ode.argtypes = [C.c_int, C.c_double, _dp, C.c_int, C.c_double,
                                         C.c_double]
        _ref = lib
    return _ref


def fnv64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return int(oracle().orc_fnv64(dp(a), a.size))


def fnv64_py(a):
    """FNV-1a 64 in Python (independent of the oracle's C implementation)."""
    h = 1469598103934665603
    for byte in np.ascontiguousarray(a, dtype="<f8").tobytes():
        h ^= byte
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


_BC = {"periodic": 0, "wall": 1, "inflow": 2}


class Oracle:
    """The oracle solver on the same arguments as sgnx.SGN2D (and a Case)."""

    def __init__(self, P, Nx, Ny, Lx, Ly, par, still_depth=None, bc=("periodic",) * 4,
                 inflow=None, gauss_points=False):
        lib = oracle()
        self.lib = lib
        self.P, self.Nx, self.Ny = P, Nx, Ny
        d = OrcDesc()
        d.P, d.Nx, d.Ny, d.gauss_points = P, Nx, Ny, int(gauss_points)
        d.Lx, d.Ly = float(Lx), float(Ly)
        g, c, gamma, nm = par.as4() if hasattr(par, "as4") else par
        d.g, d.c, d.gamma, d.nManning = g, c, gamma, nm
        for k in range(4):
            d.bc[k] = _BC[bc[k]]
        self._keep = []
        if still_depth is not None:
            for name in ("Hnod", "HfaceX", "HfaceY"):
                arr = np.ascontiguousarray(still_depth[name], dtype=np.float64)
                self._keep.append(arr)
                setattr(d, name, dp(arr))
        if inflow is not None:
            tab = np.ascontiguousarray(inflow[0], dtype=np.float64)
            self._keep.append(tab)
            d.inflow = dp(tab)
            d.inflow_omega = float(inflow[1])
        err = C.create_string_buffer(256)
        h = lib.orc_create(C.byref(d), err, 256)
        if not h:
            raise ValueError(err.value.decode())
        self.h = h
        self.Nd = int(lib.orc_ndof(h))

    @classmethod
    def from_case(cls, case, **kw):
        return cls(case.P, case.Nx, case.Ny, case.Lx, case.Ly, case.par, case.depth, case.bc,
                   case.inflow, **kw)

    def nTotal(self):
        return 5 * self.Nd

    def rhs(self, U, t=0.0):
        U = np.ascontiguousarray(U, dtype=np.float64)
        dU = np.empty_like(U)
        self.lib.orc_rhs(self.h, dp(U), dp(dU), float(t))
        return dU

    def bathymetry(self):
        a = [np.empty(self.Nd) for _ in range(3)]
        self.lib.orc_bathymetry(self.h, *[dp(x) for x in a])
        return a

    def max_wave_speed(self, U):
        return self.lib.orc_max_wave_speed_state(self.h, dp(np.ascontiguousarray(U)))

    def cfl_time_step(self, U, cfl):
        return self.lib.orc_cfl_time_step(self.h, dp(np.ascontiguousarray(U)), float(cfl))

    def mass(self, U):
        return self.lib.orc_mass(self.h, dp(np.ascontiguousarray(U)))

    def energy(self, U):
        return self.lib.orc_energy(self.h, dp(np.ascontiguousarray(U)))

    def rk4_step(self, U, t, dt):
        U = np.array(U, dtype=np.float64)
        self.lib.orc_rk4_step(self.h, float(t), float(dt), dp(U))
        return U

    def bs3_step(self, U, t, dt):
        U = np.array(U, dtype=np.float64)
        self.lib.orc_bs3_step(self.h, float(t), float(dt), dp(U))
        return U

    def integrate(self, U, t0, tEnd, dt0, opts=None, step_cap=0, call_cap=0):
        opts = opts or OrcOpts()
        U = np.array(U, dtype=np.float64)
        st = OrcStats()
        steps = np.zeros(2 * step_cap) if step_cap else None
        ct = np.zeros(call_cap) if call_cap else None
        ch = np.zeros(call_cap, dtype=np.uint64) if call_cap else None
        self.lib.orc_integrate(self.h, dp(U), float(t0), float(tEnd), float(dt0), C.byref(opts),
                               C.byref(st), dp(steps), int(step_cap), dp(ct),
                               ch.ctypes.data_as(_u64p) if ch is not None else None,
                               int(call_cap))
        return U, st, (steps.reshape(-1, 2) if step_cap else None), ct, ch

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pointwise(lib_prefix, name):
    """Handle to a pointwise physics function of the oracle or the reference."""
    lib = oracle() if lib_prefix == "orc" else ref()
    return getattr(lib, f"{lib_prefix}_{name}")
