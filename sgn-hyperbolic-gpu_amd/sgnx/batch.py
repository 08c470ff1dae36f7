"""Batches of independent runs (sgnx_batch_* in include/sgnx.h).

Solvers of uneven sizes advanced together: each RHS launch of an adaptive
step covers every running member, each member keeps its own time, step size,
controller and tile decomposition, and so ends bit for bit where
``SGN2D.integrateAdaptive`` on it alone would. The reference runs such sweeps
as separate jobs (scripts/shoal_sweep.sbatch).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def _per_member(x, n, name):
    if np.isscalar(x):
        return [float(x)] * n
    x = [float(v) for v in x]
    if len(x) != n:
        raise ValueError(f"{name}: {len(x)} values for {n} members")
    return x


class Batch:
    """``Batch([solver, ...])``: the members must use the row-sweep kernel and
    outlive the batch; their states stay in their own buffers (``download``)."""

    def __init__(self, members):
        self.members = list(members)
        self.lib = L.load()
        n = len(self.members)
        arr = (C.c_void_p * n)(*[m.h for m in self.members])
        h = C.c_void_p()
        rc = self.lib.sgnx_batch_create(arr, n, C.byref(h))
        if rc != L.SGNX_OK:
            msg = self.lib.sgnx_batch_last_error(None)
            raise L.SgnxError(rc, msg.decode() if msg else "")
        self.h = h.value

    def _check(self, rc):
        if rc != L.SGNX_OK:
            msg = self.lib.sgnx_batch_last_error(self.h)
            raise L.SgnxError(rc, msg.decode() if msg else "")

    def __len__(self):
        return len(self.members)

    def set_waves(self, mask):
        """Waves per SIMD of the batched launches by kind (bit k set = three)."""
        self._check(self.lib.sgnx_batch_set_waves(self.h, int(mask)))

    def autotune(self, reps=3):
        """One tile height for every member and the batch's waves per SIMD, by
        timing the batched launches; returns (rows, waves mask)."""
        r, w = C.c_int(), C.c_int()
        self._check(self.lib.sgnx_batch_autotune(self.h, int(reps), C.byref(r), C.byref(w)))
        return r.value, w.value

    def rhs_device(self, U_ptrs, dU_ptrs, t=0.0):
        """dU[m] = R(U[m]) for every member in one launch (device addresses)."""
        n = len(self.members)
        u = (C.c_void_p * n)(*[int(p) for p in U_ptrs])
        du = (C.c_void_p * n)(*[int(p) for p in dU_ptrs])
        tt = np.asarray(_per_member(t, n, "t"), dtype=np.float64)
        self._check(self.lib.sgnx_batch_rhs(self.h, u, du, tt.ctypes.data_as(L._dp)))

    def integrateAdaptive(self, t0, tEnd, dt0, opts=None, graph=True):
        """Every member's integrateAdaptive, in lockstep; returns their stats.
        t0, tEnd, dt0: one value or one per member; opts: one RkOpts or a list."""
        n = len(self.members)
        a0 = np.asarray(_per_member(t0, n, "t0"), dtype=np.float64)
        a1 = np.asarray(_per_member(tEnd, n, "tEnd"), dtype=np.float64)
        a2 = np.asarray(_per_member(dt0, n, "dt0"), dtype=np.float64)
        if opts is None:
            opts = L.RkOpts()
        if isinstance(opts, L.RkOpts):
            opts = [opts] * n
        if len(opts) != n:
            raise ValueError(f"opts: {len(opts)} entries for {n} members")
        oa = (L.RkOpts * n)(*opts)
        st = (L.RkStats * n)()
        flags = L.BATCH_GRAPH if graph else 0
        self._check(self.lib.sgnx_batch_integrate(
            self.h, a0.ctypes.data_as(L._dp), a1.ctypes.data_as(L._dp),
            a2.ctypes.data_as(L._dp), oa, st, flags))
        return [st[i] for i in range(n)]

    def adjoint(self, Ubar_final=None, gauge_bar=None, theta=True):
        """Every member's SGN2D.adjoint (its tape recorded with a snapshot of
        every step), the reverse sweeps in lockstep with one launch per kind
        (or, when every member alone fills the GPU, the members' own sweeps
        back to back); each member's result is bit for bit its own adjoint's. Ubar_final,
        gauge_bar: None or one array (or None) per member. Returns a list of
        (U0bar, thetaBar) (or U0bar when theta is False)."""
        n = len(self.members)
        keep = []

        def ptrs(arrs, sizes):
            out = (C.c_void_p * n)()
            for i in range(n):
                a = None if arrs is None else arrs[i]
                if a is not None:
                    a = np.ascontiguousarray(a, dtype=np.float64)
                    if sizes is not None and a.size != sizes[i]:
                        raise ValueError(f"member {i}: {a.size} values, expected {sizes[i]}")
                    keep.append(a)
                    out[i] = a.ctypes.data
            return out
        U0b = [np.empty(m.nTotal()) for m in self.members]
        thb = [np.empty(m.nDof()) for m in self.members] if theta else None
        uf = ptrs(Ubar_final, [m.nTotal() for m in self.members])
        gb = ptrs(gauge_bar, [m.gaugeValues() for m in self.members]
                  if gauge_bar is not None else None)
        ub = ptrs(U0b, None)
        tb = ptrs(thb, None) if theta else None
        self._check(self.lib.sgnx_batch_adjoint(
            self.h, uf if Ubar_final is not None else None,
            gb if gauge_bar is not None else None, ub, tb, L.MEM_HOST))
        return list(zip(U0b, thb)) if theta else U0b

    def close(self):
        if getattr(self, "h", None):
            self.lib.sgnx_batch_destroy(self.h)
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
