"""torch.autograd over adaptive runs: (U0, theta) -> (final state, gauge log).

theta is the nodal still depth (set with SGN2D.setDepth) and the gauge log
holds eta = h - theta at the gauge nodes after every accepted step. The
forward runs the device-resident BS3(2) loop on a tape (sgnx_tape_begin); the
backward is the device reverse sweep (sgnx_adjoint) over that tape, with the
recorded step sizes held fixed, so the gradient is that of the taped step
sequence (replay() evaluates exactly that map for other inputs).

Tensors cross as host memory: libsgnx runs on the system ROCm runtime, and a
torch wheel that bundles its own HIP runtime must not share device pointers
with it; inputs may live on any torch device and the outputs and gradients
are returned on the device of U0.
"""
from __future__ import annotations

import numpy as np
import torch
from torch.autograd.function import once_differentiable


def _np(x):
    return np.ascontiguousarray(x.detach().to("cpu", torch.float64).numpy())


class _Run(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, theta, solver, spec):
        solver.setDepth(_np(theta))
        solver.upload(_np(U0))
        solver.setGauges(spec["nodes"], spec["max_rows"])
        if spec["replay"]:
            if getattr(solver, "_tape_token", None) is None:
                raise RuntimeError("replay: the solver holds no tape from run_adaptive")
            solver.tapeReplay()
            ctx.stats = None
        else:
            solver.tapeBegin(spec["max_snapshots"])
            try:
                ctx.stats = solver.integrateAdaptive(spec["t0"], spec["tEnd"], spec["dt0"],
                                                     spec["opts"])
            finally:
                solver.tapeEnd()
            solver._tape_token = object()
        ctx.solver = solver
        ctx.token = solver._tape_token
        ctx.device = U0.device
        UN = solver.download()
        _, eta = solver.gaugeLog()
        return (torch.from_numpy(UN).to(U0.device), torch.from_numpy(eta).to(U0.device))

    @staticmethod
    @once_differentiable
    def backward(ctx, gU, gEta):
        s = ctx.solver
        if getattr(s, "_tape_token", None) is not ctx.token:
            raise RuntimeError("the solver recorded another run after this one; "
                               "its tape is gone")
        ub = _np(gU) if gU is not None else None
        gb = _np(gEta) if gEta is not None and gEta.numel() else None
        U0b, thb = s.adjoint(ub, gb)
        dev = ctx.device
        return torch.from_numpy(U0b).to(dev), torch.from_numpy(thb).to(dev), None, None


def run_adaptive(solver, U0, theta, t0, tEnd, dt0, opts=None, gauges=(), max_rows=None,
                 max_snapshots=1 << 20):
    """Differentiable adaptive run on `solver` from U0 (5*Nd) over still depth
    theta (Nd): returns (U_final, eta[rows, gauges]). Records the solver's tape
    (a later run_adaptive on the same solver replaces it)."""
    import sgnx
    opts = opts if opts is not None else sgnx.RkOpts()
    nodes = np.ascontiguousarray(gauges, dtype=np.int64)
    if max_rows is None:
        max_rows = opts.max_accepted if opts.max_accepted > 0 else 4096
    spec = dict(replay=False, t0=float(t0), tEnd=float(tEnd), dt0=float(dt0), opts=opts,
                nodes=nodes, max_rows=int(max_rows), max_snapshots=int(max_snapshots))
    return _Run.apply(U0, theta, solver, spec)


def replay(solver, U0, theta, gauges=(), max_rows=4096):
    """The taped step sequence of the solver's last run_adaptive (same t, dt)
    from other inputs, differentiable with the same tape."""
    nodes = np.ascontiguousarray(gauges, dtype=np.int64)
    spec = dict(replay=True, nodes=nodes, max_rows=int(max_rows))
    return _Run.apply(U0, theta, solver, spec)
