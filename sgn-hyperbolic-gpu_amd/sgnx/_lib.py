"""ctypes binding of libsgnx (include/sgnx.h).

The shared library is the product: every numerical entry point below runs the
HIP kernels. If the library is missing the import fails loudly; there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SGNX_LIB") or os.path.normpath(os.path.join(_HERE, "..", "libsgnx.so"))

SGNX_OK = 0
SGNX_EINVAL = 1
SGNX_EDEVICE = 2
SGNX_ENOMEM = 3
SGNX_EUNSUPPORTED = 4

BC_PERIODIC = 0
BC_WALL = 1
BC_INFLOW = 2

MEM_HOST = 0
MEM_DEVICE = 1
BATCH_GRAPH = 1   # SGNX_BATCH_GRAPH

_dp = C.POINTER(C.c_double)


class Params(C.Structure):
    """sgn2d::Params (src/physics/sgn_flux2d.h:57-62)."""

    _fields_ = [("g", C.c_double), ("c", C.c_double), ("gamma", C.c_double),
                ("nManning", C.c_double)]

    def __init__(self, g=9.81, c=0.0, gamma=1.5, nManning=0.0):
        super().__init__(g, c, gamma, nManning)

    def as4(self):
        return (self.g, self.c, self.gamma, self.nManning)


class Desc(C.Structure):
    _fields_ = [("P", C.c_int), ("Nx", C.c_int), ("Ny", C.c_int),
                ("Lx", C.c_double), ("Ly", C.c_double), ("par", Params),
                ("D", _dp), ("lL", _dp), ("lR", _dp), ("gLp", _dp), ("gRp", _dp), ("qw", _dp),
                ("Hnod", _dp), ("Hx", _dp), ("Hy", _dp), ("HfaceX", _dp), ("HfaceY", _dp),
                ("bc", C.c_int * 4), ("inflow", _dp), ("inflow_omega", C.c_double)]


class RkOpts(C.Structure):
    """AdaptiveRKOptions (src/solver/rk.h:53-61) plus max_accepted."""

    _fields_ = [("atol", C.c_double), ("rtol", C.c_double), ("dtMin", C.c_double),
                ("dtMax", C.c_double), ("safety", C.c_double), ("facMin", C.c_double),
                ("facMax", C.c_double), ("max_accepted", C.c_long)]

    def __init__(self, atol=1e-8, rtol=1e-6, dtMin=1e-13, dtMax=1e30, safety=0.9, facMin=0.2,
                 facMax=5.0, max_accepted=0):
        super().__init__(atol, rtol, dtMin, dtMax, safety, facMin, facMax, max_accepted)


class RkStats(C.Structure):
    """AdaptiveRKStats (src/solver/rk.h:64-72)."""

    _fields_ = [("accepted", C.c_long), ("rejected", C.c_long), ("rhsEvals", C.c_long),
                ("tReached", C.c_double), ("dtLast", C.c_double), ("maxNormLast", C.c_double),
                ("aborted", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# (name, restype, argtypes) for every symbol include/sgnx.h declares
SIGNATURES = [
    ("sgnx_last_error", C.c_char_p, [C.c_void_p]),
    ("sgnx_create", C.c_int, [C.POINTER(Desc), C.POINTER(C.c_void_p)]),
    ("sgnx_destroy", C.c_int, [C.c_void_p]),
    ("sgnx_set_stream", C.c_int, [C.c_void_p, C.c_void_p]),
    ("sgnx_ndof", C.c_longlong, [C.c_void_p]),
    ("sgnx_state_ptr", C.c_void_p, [C.c_void_p]),
    ("sgnx_upload", C.c_int, [C.c_void_p, _dp, C.c_size_t]),
    ("sgnx_download", C.c_int, [C.c_void_p, _dp, C.c_size_t]),
    ("sgnx_rhs", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_double, C.c_int]),
    ("sgnx_rhs_once", C.c_int, [C.c_void_p, _dp, C.c_size_t, C.c_double]),
    ("sgnx_integrate", C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                 C.POINTER(RkOpts), C.POINTER(RkStats), _dp, C.c_long]),
    ("sgnx_rk4_step", C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    ("sgnx_bs3_step", C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    ("sgnx_max_wave_speed", C.c_int, [C.c_void_p, _dp]),
    ("sgnx_cfl_time_step", C.c_int, [C.c_void_p, C.c_double, _dp]),
    ("sgnx_mass", C.c_int, [C.c_void_p, _dp]),
    ("sgnx_energy", C.c_int, [C.c_void_p, _dp]),
    ("sgnx_bathymetry", C.c_int, [C.c_void_p, _dp, _dp]),
    ("sgnx_set_rhs_kernel", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("sgnx_get_rhs_kernel", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("sgnx_enable_timing", C.c_int, [C.c_void_p, C.c_int]),
    ("sgnx_rhs_timing", C.c_int, [C.c_void_p, C.POINTER(C.c_long), _dp]),
    ("sgnx_eta_max", C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, _dp,
                               C.POINTER(C.c_longlong)]),
    ("sgnx_gauges", C.c_int, [C.c_void_p, C.POINTER(C.c_longlong), C.c_int, _dp]),
    ("sgnx_gauges_noalloc", C.c_int, [C.c_void_p, C.POINTER(C.c_longlong), C.c_int, _dp]),
    ("sgnx_element_means", C.c_int, [C.c_void_p, _dp]),
    ("sgnx_set_gauges", C.c_int, [C.c_void_p, C.POINTER(C.c_longlong), C.c_int, C.c_long]),
    ("sgnx_gauge_log", C.c_int, [C.c_void_p, _dp, _dp, C.c_long, C.POINTER(C.c_long)]),
    ("sgnx_batch_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]),
    ("sgnx_batch_destroy", C.c_int, [C.c_void_p]),
    ("sgnx_batch_last_error", C.c_char_p, [C.c_void_p]),
    ("sgnx_batch_set_waves", C.c_int, [C.c_void_p, C.c_int]),
    ("sgnx_batch_rhs", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), _dp]),
    ("sgnx_batch_integrate", C.c_int, [C.c_void_p, _dp, _dp, _dp, C.POINTER(RkOpts),
                                       C.POINTER(RkStats), C.c_int]),
    ("sgnx_batch_adjoint", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int]),
    ("sgnx_autotune", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("sgnx_batch_autotune", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    ("sgnx_set_depth", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    ("sgnx_rhs_vjp", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_size_t, C.c_double, C.c_int]),
    ("sgnx_tape_begin", C.c_int, [C.c_void_p, C.c_long]),
    ("sgnx_tape_begin_plan", C.c_int, [C.c_void_p, C.c_long, C.c_long]),
    ("sgnx_tape_plan_positions", C.c_int, [C.c_long, C.c_long, C.POINTER(C.c_long), C.c_long,
                                           C.POINTER(C.c_long)]),
    ("sgnx_tape_end", C.c_int, [C.c_void_p]),
    ("sgnx_tape_set_slots", C.c_int, [C.c_void_p, C.c_long]),
    ("sgnx_tape_keep_stages", C.c_int, [C.c_void_p, C.c_int]),
    ("sgnx_tape_stages", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
    ("sgnx_adjoint_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    ("sgnx_tape_info", C.c_int, [C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_long),
                                 C.POINTER(C.c_long)]),
    ("sgnx_tape_replay", C.c_int, [C.c_void_p]),
    ("sgnx_tape_steps", C.c_int, [C.c_void_p, _dp, _dp, C.c_long, C.POINTER(C.c_long)]),
    ("sgnx_adjoint", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_int]),
    ("sgnx_fr_operators", C.c_int, [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    ("sgnx_node_coords", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp]),
    ("sgnx_face_coords", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                   C.c_int, _dp, _dp, _dp, _dp]),
]

_lib = None


def load():
    """Load libsgnx.so (built in-tree by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libsgnx.so not found at {LIB_PATH}: build it first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class SgnxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libsgnx error {code}: {msg}")
        self.code = code


def check(rc, handle=None):
    if rc != SGNX_OK:
        msg = load().sgnx_last_error(handle)
        raise SgnxError(rc, msg.decode() if msg else "")
