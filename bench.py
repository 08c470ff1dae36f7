#!/usr/bin/env python3
"""Throughput of the FR/SGN right-hand side inside the adaptive BS3(2) loop on
one MI355X, in the reference's own unit (apps/sgn_bench.cu:138-141,
README.md:372-376): DOF-RHS-evals/s = 5*Nx*Ny*(P+1)^2 * N_RHS / wall time,
N_RHS counting the RHS evaluations of rejected steps.

Workload (the single-GPU production share): 1024x512 elements, P = 4, FP64,
96x48 m, 1:20 slope 0.45 -> 0.10 m plus a seeded elliptic shoal, a 1 s /
0.01 m incident wave on the left side, walls elsewhere, fluid at rest;
atol 1e-8, rtol 1e-6, dt0 = cflTimeStep(U0, 0.5).

A "step" is one accepted adaptive step. Method as apps/sgn_bench.cu:111-128:
a warm-up integration, re-upload of the initial state, then ONE device-resident
integration of exactly --steps accepted steps, timed between two device
synchronisations (state resident in HBM before the clock starts).

Also reported:
  roofline      the RHS kernel (k_rhs_sweep) as the timed run launches it, with
                the RK stage fused: the RHS's algorithmic 104 B per node (U
                read, dU written, H, Hx, Hy read) per launch over the launches'
                mean duration, measured with HIP events around every launch,
                against the 8 TB/s HBM3E peak (`achieved`, `frac`). Beside it:
                the same launches priced with the stage operands each launch
                kind also reads and writes (ALG_BYTES_BY_KIND,
                `achieved_stage_bytes`), and the plain RHS on its own (20
                launches after the timed run); `traffic` is the
                measured HBM bytes per launch from the rocprofv3 counter runs
                summarised in profiles/rhs_traffic.json (per launch kind,
                weighted by the timed run's launch mix; null if absent)
  cpu_baseline  the CPU oracle (the reference path restated op for op) on one
                host core over a bounded sample of the same workload: the FSAL
                head plus two accepted steps; the same sample also checks the
                full-size RHS bitwise and the 2-step trajectory. Beside it, the
                aggregate of N concurrent single-threaded instances.
  batched       the 64-member uneven sweep (Nx 128..768): one at a time against
                one batch (one launch per stage kind, captured step graph),
                members bit-identical
  batched_small the same with 64 small members (Nx 32..128, 50 steps each), where
                per-step launch and synchronisation costs dominate a solo run
  adjoint       a gradient at 1024x512 (taped run + reverse sweep) against the
                plain run, with and without kept stage inputs
  inversion     the gauge-inversion gradient (256x128, 2000 steps, at most 40 states
                held: planned binomial snapshots and recomputation slots) checked
                by differences in 32 directions
  batched_adjoint  16 uneven members' gradients, one by one against one batch,
                each checked by differences
  margin        the 2000x1000 margin workload (50 M nodes), 1000 adaptive steps
  margin_adjoint  a source-inversion gradient on the margin workload (100 steps,
                at most 20 states of 2 GB held, planned binomial schedule) checked
                by differences in 3 directions

usage: python bench.py [--gpus 1] [--steps 500] [--warmup 20] [--seed 0]
       python bench.py --profile-rhs N    (N RHS launches + a copy of known size,
                                           for rocprofv3 counter runs)
       python bench.py --profile-vjp N    (N RHS vector-Jacobian products)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sgn-hyperbolic-gpu_amd"))

METRIC = "DOF-RHS-evals/s at P=4 1024x512, % HBM roofline; batched and adjoint vs plain"
HBM_PEAK_GBS = 8000.0
ALG_BYTES_PER_NODE = 104          # 13 doubles: U (5) + dU (5) + H, Hx, Hy
# Algorithmic bytes per node of each RHS launch kind of an adaptive step (the
# RHS's 104 B plus the stage operands read and the results written in the same
# launch): R(U); R(u + a k1) [u, k1 in]; R(u + a k2) with the candidate
# [u, k2, k1 in; uNew, z out; k3 not stored]; k4 = R(uNew) with the error
# [uNew, u, z in; k4 out].
ALG_BYTES_BY_KIND = (104, 144, 224, 184)
KIND_NAMES = ("plain", "stage2", "stage3_candidate", "k4_error")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_case(args):
    from sgnx import cases
    t = time.perf_counter()
    case = cases.shoal(seed=args.seed, Nx=args.nx, Ny=args.ny)
    log(f"[bench] case {case.name} {case.Nx}x{case.Ny} P={case.P} built in "
        f"{time.perf_counter() - t:.1f}s")
    return case


def profile_rhs(args):
    """RHS launches on device-resident buffers, plus a plain device-to-device
    copy of known size to calibrate the HBM byte counters."""
    import numpy as np
    import sgnx
    from sgnx import hip
    case = build_case(args)
    s = sgnx.SGN2D(**case.solver_kwargs())
    U = hip.DeviceArray.from_host(case.U0)
    dU = hip.DeviceArray(U.n)
    hip.synchronize()
    for _ in range(args.profile_rhs):
        s.rhs_device(U.address, dU.address)
    hip.synchronize()
    a = hip.DeviceArray.from_host(np.ones(args.copy_mib * 131072))
    b = hip.DeviceArray(a.n)
    for _ in range(3):
        b.copy_from(a)
    hip.synchronize()
    if args.profile_steps:
        s.upload(case.U0)
        dt0 = s.cflTimeStep(None, 0.5)
        st = s.integrateAdaptive(0.0, 1e9, dt0, sgnx.RkOpts(max_accepted=args.profile_steps))
        log(f"[bench] profile: {st.accepted} adaptive steps ({st.rhsEvals} RHS)")
    log(f"[bench] profile: {args.profile_rhs} RHS launches of {case.nd} nodes; "
        f"3 device copies of {args.copy_mib} MiB")
    s.close()


def profile_vjp(args):
    """RHS vector-Jacobian products (k_vjp_coef + k_vjp_elem) on device-resident
    buffers of the shoal state, with the theta cotangent, then one reverse sweep
    of --profile-steps taped steps."""
    import numpy as np
    import sgnx
    from sgnx import hip
    case = build_case(args)
    s = sgnx.SGN2D(**case.solver_kwargs())
    s.setDepth(case.depth["Hnod"])
    rng = np.random.default_rng(args.seed)
    U = hip.DeviceArray.from_host(case.U0)
    W = hip.DeviceArray.from_host(rng.normal(size=case.U0.size))
    Ub = hip.DeviceArray(U.n)
    tb = hip.DeviceArray(case.nd)
    for _ in range(args.profile_vjp):
        s.rhs_vjp_device(U.address, W.address, Ub.address, tb.address)
    if args.profile_steps:
        s.upload(case.U0)
        dt0 = s.cflTimeStep(None, 0.5)
        s.tapeBegin(args.profile_steps + 1)
        st = s.integrateAdaptive(0.0, 1e9, dt0, sgnx.RkOpts(max_accepted=args.profile_steps))
        s.tapeEnd()
        U0b = hip.DeviceArray(U.n)
        s.adjoint_device(W.address, 0, U0b.address, tb.address)
        log(f"[bench] profile: {st.accepted} taped steps and their reverse sweep")
    hip.synchronize()
    log(f"[bench] profile: {args.profile_vjp} RHS products of {case.nd} nodes")
    s.close()


def plain_rhs_rate(s, nd, n=20):
    """The RHS kernel on its own (no stage operands): n launches R(U) on the
    final state of the timed run, device-resident, timed with HIP events."""
    from sgnx import hip
    U = hip.DeviceArray.from_host(s.download())
    dU = hip.DeviceArray(U.n)
    for _ in range(3):
        s.rhs_device(U.address, dU.address)
    s.enable_timing(True)
    s.rhs_timing()
    for _ in range(n):
        s.rhs_device(U.address, dU.address)
    cnt, ms = s.rhs_timing()[0]
    s.enable_timing(False)
    mean_s = ms / cnt / 1e3
    gbs = ALG_BYTES_PER_NODE * nd / mean_s / 1e9
    log(f"[bench] plain RHS: {mean_s * 1e3:.4f} ms = {gbs:.0f} GB/s algorithmic "
        f"({100 * gbs / HBM_PEAK_GBS:.1f} % of peak)")
    return {"launches": cnt, "mean_ms": mean_s * 1e3, "achieved": gbs,
            "frac": gbs / HBM_PEAK_GBS}


def batched_sweep(args, nx_min=128, nx_max=768, n_members=None, steps=None, single_rate=None):
    """The batched shoal sweep: --batch-members grids of seeded uneven sizes
    (Nx in [128, 768], Ny = Nx/2; cases.sweep_member), each advanced
    --batch-steps accepted adaptive steps, first one at a time with the plain
    solver, then all together in one Batch (one launch per stage kind, a
    captured step graph); every member must end bit for bit where its solo run
    did. The tile height is chosen once by the batch autotuner and used by
    both runs. Before that, the members also run one at a time with their
    own default tile heights: the plain solver's rate without the batch,
    against which `speedup` is taken (each member's best solo setting).
    `speedup_bound` is what batching could gain at most if every member ran
    at `single_rate` (the rate of one production-size grid): the members'
    solo time over their work at that rate."""
    import hashlib
    import sgnx
    from sgnx import cases
    n_members = args.batch_members if n_members is None else n_members
    steps = args.batch_steps if steps is None else steps
    t = time.perf_counter()
    sizes = cases.sweep_sizes(n_members, seed=args.seed, nx_min=nx_min, nx_max=nx_max)
    members, U0s, dt0s = [], [], []
    for i, nx in enumerate(sizes):
        c = cases.sweep_member(args.seed * 1000 + i + 1, nx)
        s = sgnx.SGN2D(**c.solver_kwargs())
        s.upload(c.U0)
        dt0s.append(s.cflTimeStep(None, 0.5))
        members.append(s)
        U0s.append(c.U0)
    nodes = sum(s.nDof() for s in members)
    log(f"[bench] batch: {len(members)} members, Nx {min(sizes)}..{max(sizes)}, "
        f"{nodes / 1e6:.1f} M nodes, built in {time.perf_counter() - t:.1f}s")
    opts = sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=steps)
    one = sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=1)
    # the members one at a time with their own default tile heights (the
    # tallest that fills the chip, per grid): the sequential rate a user of
    # the plain solver gets
    members[0].integrateAdaptive(0.0, 1e9, dt0s[0], one)
    own_wall, own_dof = 0.0, 0.0
    for s, U0, dt0 in zip(members, U0s, dt0s):
        s.upload(U0)
        t0 = time.perf_counter()
        st = s.integrateAdaptive(0.0, 1e9, dt0, opts)
        own_wall += time.perf_counter() - t0
        own_dof += 5.0 * s.nDof() * st.rhsEvals
    b = sgnx.Batch(members)
    rows, waves = b.autotune()
    members[0].integrateAdaptive(0.0, 1e9, dt0s[0], one)       # warm both code paths
    b.integrateAdaptive(0.0, 1e9, dt0s, one, graph=True)

    def digest(s):
        return hashlib.sha1(s.download().tobytes()).hexdigest()

    seq_wall, seq_dof, solo = 0.0, 0.0, []
    for s, U0, dt0 in zip(members, U0s, dt0s):
        s.upload(U0)
        t0 = time.perf_counter()
        st = s.integrateAdaptive(0.0, 1e9, dt0, opts)
        seq_wall += time.perf_counter() - t0
        seq_dof += 5.0 * s.nDof() * st.rhsEvals
        solo.append((digest(s), st.as_dict()))
    for s, U0 in zip(members, U0s):
        s.upload(U0)
    t0 = time.perf_counter()
    sts = b.integrateAdaptive(0.0, 1e9, dt0s, opts, graph=True)
    bat_wall = time.perf_counter() - t0
    bat_dof = sum(5.0 * s.nDof() * st.rhsEvals for s, st in zip(members, sts))
    same = all(digest(s) == h and st.as_dict() == sd
               for s, st, (h, sd) in zip(members, sts, solo))
    b.close()
    for s in members:
        s.close()
    out = {"members": len(members), "nx_min": min(sizes), "nx_max": max(sizes),
           "nodes": nodes, "steps_per_member": steps,
           "autotune": {"rows": rows, "waves3": waves},
           "sequential": {"value": seq_dof / seq_wall, "wall_s": seq_wall},
           "sequential_own_tiles": {"value": own_dof / own_wall, "wall_s": own_wall},
           "batched": {"value": bat_dof / bat_wall, "wall_s": bat_wall, "graph": True},
           "speedup": (bat_dof / bat_wall) / (own_dof / own_wall),
           "speedup_basis": "each member's own default tile height (its best solo setting)",
           "speedup_vs_batch_tiles": (bat_dof / bat_wall) / (seq_dof / seq_wall),
           "members_bit_identical": bool(same), "unit": "DOF-RHS-evals/s"}
    if single_rate:
        out["single_grid_rate"] = single_rate
        out["speedup_bound"] = own_wall / (own_dof / single_rate)
    log(f"[bench] batch: sequential {seq_dof / seq_wall / 1e9:.2f} GDOF-RHS/s ({seq_wall:.2f}s), "
        f"batched {bat_dof / bat_wall / 1e9:.2f} GDOF-RHS/s ({bat_wall:.2f}s), "
        f"x{out['speedup_vs_batch_tiles']:.2f}, identical={same}, rows {rows} waves {waves:#x}; "
        f"one at a time with each grid's own tile height {own_dof / own_wall / 1e9:.2f} "
        f"GDOF-RHS/s: speed-up x{out['speedup']:.2f}"
        + (f" (bound x{out['speedup_bound']:.2f} at the single-grid rate)" if single_rate else ""))
    return out


def adjoint_leg(args):
    """Cost of a gradient on the timed workload: an adaptive run of
    --adjoint-steps accepted steps from the shoal state, its still depth set
    from the nodal field theta (SGN2D.setDepth), gauges at 16 nodes; then the
    same run taped (every step's input state kept, and its stage inputs, which
    the forward's stage launches write) and the reverse sweep for random
    cotangents of the final state and of the gauge log, giving the cotangents
    of the initial state and of theta. Reports the sweep and the whole
    gradient (taped run + sweep) against the plain run, the same for a tape
    without stages (the sweep recomputes them), and one RHS vector-Jacobian
    product against one RHS."""
    import numpy as np
    import sgnx
    from sgnx import cases, hip
    case = cases.shoal(seed=args.seed, Nx=args.nx, Ny=args.ny)
    s = sgnx.SGN2D(**case.solver_kwargs())
    s.setDepth(case.depth["Hnod"])
    n, nd = case.U0.size, case.nd
    rng = np.random.default_rng(args.seed)
    nodes = np.linspace(0, nd - 1, 16).astype(np.int64)
    steps = args.adjoint_steps
    s.upload(case.U0)
    dt0 = s.cflTimeStep(None, 0.5)
    opts = sgnx.RkOpts(atol=1e-8, rtol=1e-6, max_accepted=steps)
    Ubar = rng.normal(size=n)
    Gbar = rng.normal(size=(steps, nodes.size))

    def run(taped, keep=False):
        s.upload(case.U0)
        s.setGauges(nodes, steps)
        if taped:
            s.tapeKeepStages(keep)
            s.tapeBegin(steps + 1)
        t0 = time.perf_counter()
        st = s.integrateAdaptive(0.0, 1e9, dt0, opts)
        wall = time.perf_counter() - t0
        if taped:
            s.tapeEnd()
        return st, wall

    run(True, False)                               # warm-up: allocates the tape
    s.adjoint(Ubar, Gbar)
    st0, t_fwd = run(False)
    Ud = hip.DeviceArray.from_host(Ubar)
    Gd = hip.DeviceArray.from_host(Gbar)
    U0b = hip.DeviceArray(n)
    thb = hip.DeviceArray(nd)

    def sweep():
        t0 = time.perf_counter()
        s.adjoint_device(Ud.address, Gd.address, U0b.address, thb.address)
        return time.perf_counter() - t0, U0b.download(), thb.download()

    # the sweep recomputing each step's stages, then with the stages kept
    # (after a warm-up run that allocates their buffers)
    st1, t_tape_r = run(True, False)
    t_adj_r, U0b_r, thb_r = sweep()
    run(True, True)
    st2, t_tape = run(True, True)
    kept = s.tapeStages()
    t_adj, U0b_k, thb_k = sweep()
    same = bool(np.array_equal(U0b_r.view(np.int64), U0b_k.view(np.int64))
                and np.array_equal(thb_r.view(np.int64), thb_k.view(np.int64)))
    # one product against one RHS on the final state (device operands)
    U = hip.DeviceArray.from_host(s.download())
    W = hip.DeviceArray.from_host(rng.normal(size=n))
    dU = hip.DeviceArray(n)
    reps = 20
    s.rhs_vjp_device(U.address, W.address, dU.address, thb.address)
    t0 = time.perf_counter()
    for _ in range(reps):
        s.rhs_vjp_device(U.address, W.address, dU.address, thb.address)
    t_vjp = (time.perf_counter() - t0) / reps
    s.rhs_device(U.address, dU.address)
    hip.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        s.rhs_device(U.address, dU.address)
    hip.synchronize()
    t_rhs = (time.perf_counter() - t0) / reps
    ok = bool(np.all(np.isfinite(U0b_k)) and np.all(np.isfinite(thb_k)))
    info = s.tapeInfo()
    s.close()
    out = {"workload": f"shoal {case.Nx}x{case.Ny}, {steps} accepted steps, 16 gauges",
           "same_step_counts": st0.as_dict() == st1.as_dict() == st2.as_dict(),
           "accepted": st0.accepted, "rejected": st0.rejected,
           "forward_ms": 1e3 * t_fwd, "taped_forward_ms": 1e3 * t_tape, "adjoint_ms": 1e3 * t_adj,
           "adjoint_over_forward": t_adj / t_fwd,
           "gradient_over_forward": (t_tape + t_adj) / t_fwd,
           "stages_kept": kept,
           "recomputing": {"taped_forward_ms": 1e3 * t_tape_r, "adjoint_ms": 1e3 * t_adj_r,
                           "gradient_over_forward": (t_tape_r + t_adj_r) / t_fwd},
           "kept_equals_recomputed": same,
           "snapshots": info[2], "stride": info[1],
           "vjp_ms": 1e3 * t_vjp, "rhs_ms": 1e3 * t_rhs, "vjp_over_rhs": t_vjp / t_rhs,
           "cotangents_finite": ok}
    log(f"[bench] adjoint: forward {1e3 * t_fwd:.1f} ms, taped {1e3 * t_tape:.1f} ms, "
        f"reverse sweep {1e3 * t_adj:.1f} ms -> sweep/forward {t_adj / t_fwd:.2f}, "
        f"gradient/forward {(t_tape + t_adj) / t_fwd:.2f} (stages kept; recomputing "
        f"{(t_tape_r + t_adj_r) / t_fwd:.2f}, same bits {same}); VJP {1e3 * t_vjp:.3f} ms vs RHS "
        f"{1e3 * t_rhs:.3f} ms")
    return out


def inversion_leg(args):
    """The gauge-inversion workload (cases.inversion: 256x128, walls, surface
    humps at rest, 20 gauges): the gradient of the gauge misfit
    0.5 |G - G_obs|^2 with respect to the initial surface eta0 and the still
    depth theta over --inversion-steps adaptive steps (atol 1e-10, rtol 1e-8),
    under a budget of 40
    states held at once (snapshots and recomputation slots, the snapshots at the
    first sweep of the binomial schedule for the known step count). G_obs is the same
    run from 1.1 eta0. Reports its cost against the plain run, and the largest
    relative difference to central differences of the plain solver's replay
    of the taped steps over --inversion-dirs seeded smooth directions (held
    step sizes leave rough perturbations' grid modes unstable)."""
    import numpy as np
    import sgnx
    from sgnx import cases
    c = cases.inversion(seed=args.seed)
    theta, eta0, nodes = c.meta["theta"], c.meta["eta0"], c.meta["gauges"]
    nd, steps = c.nd, args.inversion_steps
    rest = c.U0[nd:]
    s = sgnx.SGN2D(**c.solver_kwargs())
    s.setDepth(theta)
    s.upload(c.U0)
    dt0 = s.cflTimeStep(None, 0.5)
    opts = sgnx.RkOpts(atol=c.meta["atol"], rtol=c.meta["rtol"], max_accepted=steps)
    s.upload(np.concatenate([theta + 1.1 * eta0, rest]))
    s.setGauges(nodes, steps)
    s.integrateAdaptive(0.0, 1e9, dt0, opts)
    G_obs = s.gaugeLog()[1]
    s.upload(c.U0)
    s.setGauges(nodes, steps)
    t0 = time.perf_counter()
    st = s.integrateAdaptive(0.0, 1e9, dt0, opts)
    t_fwd = time.perf_counter() - t0
    s.upload(c.U0)
    s.setGauges(nodes, steps)
    s.tapeBeginPlan(steps, 40)
    t0 = time.perf_counter()
    st1 = s.integrateAdaptive(0.0, 1e9, dt0, opts)
    t_tape = time.perf_counter() - t0
    s.tapeEnd()
    G = s.gaugeLog()[1]
    t0 = time.perf_counter()
    U0b, thb = s.adjoint(None, G - G_obs)
    t_adj = time.perf_counter() - t0
    recomputed, peak = s.adjointStats()
    _, stride, held = s.tapeInfo()
    g_eta, g_theta = U0b[:nd], thb + U0b[:nd]

    def misfit(e, th):
        s.setDepth(th)
        s.upload(np.concatenate([th + e, rest]))
        s.setGauges(nodes, steps)
        s.tapeReplay()
        return 0.5 * np.sum((s.gaugeLog()[1] - G_obs) ** 2)

    rng = np.random.default_rng(args.seed + 32)
    eps, errs, rels = 3e-3, [], []
    t0 = time.perf_counter()
    for _ in range(args.inversion_dirs):
        de = cases.smooth_direction(c, rng, near=nodes)
        dth = cases.smooth_direction(c, rng, near=nodes)
        # fourth-order central difference: truncation O(eps^4)
        f = {k: misfit(eta0 + k * eps * de, theta + k * eps * dth) for k in (-2, -1, 1, 2)}
        fd = (8.0 * (f[1] - f[-1]) - (f[2] - f[-2])) / (12.0 * eps)
        pe, pt = g_eta @ de, g_theta @ dth
        errs.append(abs(pe + pt - fd) / (abs(pe) + abs(pt)))
        rels.append(abs(pe + pt - fd) / max(abs(pe + pt), abs(fd), 1e-300))
    t_fd = time.perf_counter() - t0
    s.close()
    out = {"workload": (f"inversion {c.Nx}x{c.Ny} on 64x32 m, walls, 8 surface humps at rest, "
                        f"{nodes.size} gauges, {steps} accepted steps"),
           "accepted": st.accepted, "rejected": st.rejected,
           "same_step_counts": st.as_dict() == st1.as_dict(),
           "stored_states": {"snapshots": held, "slots_peak": peak, "budget": 40,
                             "schedule": "binomial, planned for the step count (snapshots "
                                         "at its first sweep, the taped run)"},
           "recomputed_steps": recomputed,
           "forward_ms": 1e3 * t_fwd, "taped_forward_ms": 1e3 * t_tape, "adjoint_ms": 1e3 * t_adj,
           "adjoint_over_forward": t_adj / t_fwd,
           "gradient_over_forward": (t_tape + t_adj) / t_fwd,
           "fd_directions": len(errs), "fd_max_rel_err": max(errs) if errs else None,
           "fd_median_rel_err": float(np.median(errs)) if errs else None,
           "fd_error_norm": "|adjoint - fd| / (|eta0 part| + |theta part|) of the directional derivative",
           "fd_max_plain_rel_err": max(rels) if rels else None,
           "fd_median_plain_rel_err": float(np.median(rels)) if rels else None,
           "fd_reference": ("fourth-order central differences (eps 3e-3) of the plain solver's "
                            "replay of the taped steps, along seeded smooth directions (4 "
                            "Gaussians each for eta0 and theta, centred near gauges)")}
    log(f"[bench] inversion: {st.accepted} steps, forward {1e3 * t_fwd:.0f} ms, taped "
        f"{1e3 * t_tape:.0f} ms, reverse {1e3 * t_adj:.0f} ms ({held} planned snapshots, "
        f"{peak} slots, {recomputed} steps recomputed) -> gradient/forward "
        f"{(t_tape + t_adj) / t_fwd:.2f}; FD over {len(errs)} directions: max rel err "
        f"{max(errs) if errs else float('nan'):.2e} ({t_fd:.1f}s)")
    return out


def batched_adjoint_leg(args):
    """Batched adjoints: --adjoint-members members of the shoal sweep (seeded
    widths 128-512, Ny = Nx/2: cases.sweep_member, inflow on the left, walls,
    a noisy seeded shoal), each run --adjoint-member-steps accepted steps at
    atol 1e-10, rtol 1e-8 with 8 gauges next to the inflow side; the gradient
    of the gauge energy 0.5 |G|^2 with respect to the still depth (the initial
    state being the fluid at rest, h = theta). Timed one member after another
    (plain run, taped run, reverse sweep) and as one batch (a batched taped
    run, one launch per kind over all members, then the batched reverse
    sweep); the tapes keep the stage inputs; the cotangents must agree bit
    for bit. The batch autotuner picks one tile height, used by both. Each member's gradient is
    checked in one smooth direction near its gauges against fourth-order
    differences of its plain solver's replay."""
    import numpy as np
    import sgnx
    from sgnx import cases
    sizes = cases.sweep_sizes(args.adjoint_members, seed=args.seed + 5, nx_min=128, nx_max=512)
    steps = args.adjoint_member_steps
    members, data = [], []
    for i, nx in enumerate(sizes):
        c = cases.sweep_member(args.seed * 1000 + 500 + i, nx)
        theta = c.depth["Hnod"]
        s = sgnx.SGN2D(**c.solver_kwargs())
        s.setDepth(theta)
        # gauges in the first two element columns, where the incident wave
        # arrives within the run: node (a=2, b=2) of elements (ex, ey)
        eys = np.linspace(0, c.Ny - 1, 4).astype(np.int64)
        nodes = np.array([(ey * c.Nx + ex) * 25 + 12 for ey in eys for ex in (0, 1)],
                         dtype=np.int64)
        s.upload(c.U0)
        dt0 = s.cflTimeStep(None, 0.5)
        members.append(s)
        data.append((c, theta, nodes, dt0))
    opts = sgnx.RkOpts(atol=1e-10, rtol=1e-8, max_accepted=steps)
    # one tile height for every member, chosen by the batch autotuner and used
    # by both the one-by-one and the batched runs (as in the batched sweep)
    b = sgnx.Batch(members)
    rows, waves = b.autotune()

    def start(s, c, nodes, taped):
        s.upload(c.U0)
        s.setGauges(nodes, steps)
        if taped:
            s.tapeKeepStages(True)
            s.tapeBegin(steps + 1)

    # one member after another
    seq = {"forward": 0.0, "taped": 0.0, "reverse": 0.0}
    for taped in (False, True):
        for s, (c, theta, nodes, dt0) in zip(members, data):
            start(s, c, nodes, taped)
            t0 = time.perf_counter()
            s.integrateAdaptive(0.0, 1e9, dt0, opts)
            seq["taped" if taped else "forward"] += time.perf_counter() - t0
            if taped:
                s.tapeEnd()
    Gs = [s.gaugeLog()[1] for s in members]
    for s, G in zip(members, Gs):
        s.adjoint(None, G)                         # warm-up
    t0 = time.perf_counter()
    solo = [s.adjoint(None, G) for s, G in zip(members, Gs)]
    seq["reverse"] = time.perf_counter() - t0
    # one batch: taped run and reverse sweep
    bat = {}
    for s, (c, theta, nodes, dt0) in zip(members, data):
        start(s, c, nodes, True)
    t0 = time.perf_counter()
    b.integrateAdaptive(0.0, 1e9, [d[3] for d in data], opts, graph=True)
    bat["taped"] = time.perf_counter() - t0
    for s in members:
        s.tapeEnd()
    Gb = [s.gaugeLog()[1] for s in members]
    b.adjoint(None, Gb)                            # warm-up
    t0 = time.perf_counter()
    res = b.adjoint(None, Gb)
    bat["reverse"] = time.perf_counter() - t0
    b.close()
    same = all(np.array_equal(r[0].view(np.uint64), q[0].view(np.uint64)) and
               np.array_equal(r[1].view(np.uint64), q[1].view(np.uint64))
               for r, q in zip(res, solo))
    # one smooth direction per member against its plain solver's replay
    errs = []
    rng = np.random.default_rng(args.seed + 55)
    eps = 3e-3
    for s, (c, theta, nodes, dt0), (U0b, thb) in zip(members, data, res):
        g_theta = thb + U0b[:c.nd]                  # h0 = theta at rest
        dth = cases.smooth_direction(c, rng, near=nodes)
        rest = c.U0[c.nd:]

        def energy(th):
            s.setDepth(th)
            s.upload(np.concatenate([th, rest]))
            s.setGauges(nodes, steps)
            s.tapeReplay()
            return 0.5 * np.sum(s.gaugeLog()[1] ** 2)
        f = {k: energy(theta + k * eps * dth) for k in (-2, -1, 1, 2)}
        fd = (8.0 * (f[1] - f[-1]) - (f[2] - f[-2])) / (12.0 * eps)
        ad = g_theta @ dth
        errs.append(abs(ad - fd) / max(abs(ad), abs(fd), 1e-300))
    log("[bench] batched adjoint FD errors: " + " ".join(f"{e:.1e}" for e in errs))
    nodes_total = sum(s.nDof() for s in members)
    for s in members:
        s.close()
    grad_seq = seq["taped"] + seq["reverse"]
    grad_bat = bat["taped"] + bat["reverse"]
    out = {"members": len(members), "nx_min": min(sizes), "nx_max": max(sizes),
           "nodes": nodes_total, "steps_per_member": steps, "rows": rows, "waves3_mask": waves,
           "sequential_s": seq, "batched_s": bat,
           "gradient_sequential_s": grad_seq, "gradient_batched_s": grad_bat,
           "gradient_speedup": grad_seq / grad_bat,
           "reverse_speedup": seq["reverse"] / bat["reverse"],
           "gradient_over_forward_sequential": grad_seq / seq["forward"],
           "gradient_over_forward_batched": grad_bat / seq["forward"],
           "members_bit_identical": bool(same),
           "fd_max_rel_err": max(errs), "fd_median_rel_err": float(np.median(errs)),
           "fd_reference": ("fourth-order central differences (eps 3e-3) of each member's plain "
                            "replay along one smooth direction near its gauges")}
    log(f"[bench] batched adjoint: {len(members)} members Nx {min(sizes)}..{max(sizes)}, "
        f"{nodes_total / 1e6:.1f} M nodes, {steps} steps each; one by one: run "
        f"{seq['forward']:.2f}s, taped {seq['taped']:.2f}s, reverse {seq['reverse']:.2f}s; "
        f"batched: taped {bat['taped']:.2f}s, reverse {bat['reverse']:.2f}s; gradient "
        f"x{grad_seq / grad_bat:.2f}, identical={same}; FD max rel err {max(errs):.2e}")
    return out


def margin_case(args):
    from sgnx import cases
    t = time.perf_counter()
    case = cases.margin(seed=args.seed)
    log(f"[bench] margin case {case.Nx}x{case.Ny} built in {time.perf_counter() - t:.1f}s")
    return case


def margin_leg(args, case):
    """The large single-grid margin workload (cases.margin: 2000x1000
    elements, 50 M nodes, 2 GB per state, on 200x100 km; a tanh margin
    4000 -> 50 m with 2 % roughness, walls in x, periodic in y, an elliptical
    uplift at rest): --margin-steps accepted adaptive steps in one
    device-resident call after a short warm-up, the same rate as the main
    line, plus the plain RHS's rate on the final state."""
    import sgnx
    s = sgnx.SGN2D(**case.solver_kwargs())
    s.upload(case.U0)
    dt0 = s.cflTimeStep(None, 0.5)
    opts = dict(atol=case.meta["atol"], rtol=case.meta["rtol"])
    s.integrateAdaptive(0.0, 1e9, dt0, sgnx.RkOpts(max_accepted=5, **opts))
    s.upload(case.U0)
    m0 = s.mass()
    t0 = time.perf_counter()
    st = s.integrateAdaptive(0.0, 1e9, dt0, sgnx.RkOpts(max_accepted=args.margin_steps, **opts))
    wall = time.perf_counter() - t0
    m1 = s.mass()
    plain = plain_rhs_rate(s, case.nd)
    s.close()
    dof = 5.0 * case.nd
    rate = dof * st.rhsEvals / wall
    log(f"[bench] margin: {st.accepted} acc / {st.rejected} rej / {st.rhsEvals} RHS in {wall:.2f}s "
        f"-> {rate / 1e9:.2f} GDOF-RHS/s, {1e3 * wall / max(st.accepted, 1):.2f} ms per step; "
        f"relative mass change {abs(m1 - m0) / abs(m0):.1e}")
    return {"workload": (f"margin {case.Nx}x{case.Ny} P=4 FP64 on 200x100 km: tanh margin "
                         f"4000 -> 50 m, 2 % roughness, walls in x, periodic in y, uplift at rest "
                         f"(seed {args.seed}); atol 1e-8 rtol 1e-6"),
            "nodes": case.nd, "accepted": st.accepted, "rejected": st.rejected,
            "rhs_evals": st.rhsEvals, "aborted": bool(st.aborted), "wall_s": wall,
            "dof_rhs_per_s": rate, "ms_per_step": 1e3 * wall / max(st.accepted, 1),
            "t_reached_s": st.tReached, "mass_rel_change": abs(m1 - m0) / abs(m0),
            "plain_rhs": plain}


def margin_adjoint_leg(args, case):
    """A source-inversion gradient on the margin workload: the gradient of
    the gauge misfit 0.5 |G(eta0) - G_obs|^2 with respect to the initial
    surface over --margin-adjoint-steps adaptive steps, 16 gauges inside the
    uplift, G_obs the same run from 1.1 times the uplift. At 2 GB per state
    the sweep holds at most 20 states (planned binomial schedule). Reports its cost against the plain run and the largest
    relative difference to fourth-order central differences of the plain
    solver's replay of the taped steps along --margin-adjoint-dirs seeded
    smooth directions (Gaussians 2-6 km wide near the gauges)."""
    import numpy as np
    import sgnx
    from sgnx import cases
    nd, steps = case.nd, args.margin_adjoint_steps
    Hn = case.depth["Hnod"]
    eta0 = case.U0[:nd] - Hn
    rest = case.U0[nd:]
    nodes = cases.margin_gauges(case, 16, seed=args.seed)
    s = sgnx.SGN2D(**case.solver_kwargs())
    s.upload(case.U0)
    dt0 = s.cflTimeStep(None, 0.5)
    opts = sgnx.RkOpts(atol=case.meta["atol"], rtol=case.meta["rtol"], max_accepted=steps)
    s.upload(np.concatenate([Hn + 1.1 * eta0, rest]))
    s.setGauges(nodes, steps)
    s.integrateAdaptive(0.0, 1e9, dt0, opts)
    G_obs = s.gaugeLog()[1]
    s.upload(case.U0)
    s.setGauges(nodes, steps)
    t0 = time.perf_counter()
    st = s.integrateAdaptive(0.0, 1e9, dt0, opts)
    t_fwd = time.perf_counter() - t0
    s.upload(case.U0)
    s.setGauges(nodes, steps)
    s.tapeBeginPlan(steps, 20)
    t0 = time.perf_counter()
    st1 = s.integrateAdaptive(0.0, 1e9, dt0, opts)
    t_tape = time.perf_counter() - t0
    s.tapeEnd()
    G = s.gaugeLog()[1]
    t0 = time.perf_counter()
    g_eta = s.adjoint(None, G - G_obs, theta=False)[:nd]
    t_adj = time.perf_counter() - t0
    recomputed, peak = s.adjointStats()
    _, stride, held = s.tapeInfo()

    def misfit(e):
        s.upload(np.concatenate([Hn + e, rest]))
        s.setGauges(nodes, steps)
        s.tapeReplay()
        return 0.5 * np.sum((s.gaugeLog()[1] - G_obs) ** 2)

    rng = np.random.default_rng(args.seed + 5)
    eps, errs = 3e-3, []
    t0 = time.perf_counter()
    for _ in range(args.margin_adjoint_dirs):
        de = cases.smooth_direction(case, rng, n=2, near=nodes, scale=1000.0)
        f = {k: misfit(eta0 + k * eps * de) for k in (-2, -1, 1, 2)}
        fd = (8.0 * (f[1] - f[-1]) - (f[2] - f[-2])) / (12.0 * eps)
        pe = g_eta @ de
        errs.append(abs(pe - fd) / max(abs(pe), abs(fd), 1e-300))
    t_fd = time.perf_counter() - t0
    s.close()
    log(f"[bench] margin adjoint: {st.accepted} steps, forward {1e3 * t_fwd:.0f} ms, taped "
        f"{1e3 * t_tape:.0f} ms, reverse {1e3 * t_adj:.0f} ms ({held} planned snapshots, "
        f"{peak} slots, {recomputed} steps recomputed) -> gradient/forward "
        f"{(t_tape + t_adj) / t_fwd:.2f}; FD over {len(errs)} directions: max rel err "
        f"{max(errs) if errs else float('nan'):.2e} ({t_fd:.1f}s)")
    return {"workload": (f"margin {case.Nx}x{case.Ny} (seed {args.seed}): gradient of the gauge "
                         f"misfit w.r.t. the initial surface, {nodes.size} gauges in the uplift, "
                         f"{steps} accepted steps, atol 1e-8 rtol 1e-6"),
            "accepted": st.accepted, "rejected": st.rejected,
            "same_step_counts": st.as_dict() == st1.as_dict(),
            "stored_states": {"snapshots": held, "slots_peak": peak, "budget": 20,
                              "bytes_per_state": 5 * 8 * nd,
                              "schedule": "binomial, planned for the step count"},
            "recomputed_steps": recomputed,
            "forward_ms": 1e3 * t_fwd, "taped_forward_ms": 1e3 * t_tape, "adjoint_ms": 1e3 * t_adj,
            "adjoint_over_forward": t_adj / t_fwd,
            "gradient_over_forward": (t_tape + t_adj) / t_fwd,
            "fd_directions": len(errs), "fd_max_rel_err": max(errs) if errs else None,
            "fd_reference": ("fourth-order central differences (eps 3e-3) of the plain solver's "
                             "replay of the taped steps, along seeded smooth directions (2 "
                             "Gaussians 2-6 km wide near the gauges)")}


def cpu_baseline(case, gpu_solver, dt0, args):
    """The oracle on one core: FSAL head + 2 accepted steps of the same workload."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    import sgnx
    o = OL.Oracle.from_case(case)
    t = time.perf_counter()
    R_cpu = o.rhs(case.U0)
    t_rhs = time.perf_counter() - t
    R_gpu = gpu_solver.rhs(case.U0)
    rhs_bitwise = bool(np.array_equal(R_cpu.view(np.uint64), R_gpu.view(np.uint64)))
    t = time.perf_counter()
    Uo, sto, _, _, _ = o.integrate(case.U0, 0.0, 1e9, dt0,
                                   OL.OrcOpts(atol=1e-8, rtol=1e-6, max_accepted=2))
    t_int = time.perf_counter() - t
    gpu_solver.upload(case.U0)
    stg = gpu_solver.integrateAdaptive(0.0, 1e9, dt0, sgnx.RkOpts(max_accepted=2))
    Ug = gpu_solver.download()
    rel = float(np.linalg.norm(Ug - Uo) / np.linalg.norm(Uo))
    dof = 5.0 * case.nd
    value = dof * sto.rhsEvals / t_int
    o.close()
    log(f"[bench] cpu: {sto.rhsEvals} RHS in {t_int:.2f}s ({value / 1e6:.1f} MDOF-RHS/s), "
        f"single RHS {t_rhs:.2f}s; full-size RHS bitwise={rhs_bitwise}, "
        f"2-step rel L2={rel:.2e}")
    agg = cpu_aggregate(case, OL)
    return {"value": value, "unit": "DOF-RHS-evals/s", "cores": 1, "kind": "port",
            "sample": (f"CPU oracle (reference path restated op for op, g++ -O2 "
                       f"-ffp-contract=off, 1 thread) on the same {case.Nx}x{case.Ny} workload: "
                       f"FSAL head + 2 accepted BS3(2) steps = {sto.rhsEvals} RHS evaluations "
                       f"in {t_int:.2f} s"),
            "check": {"full_size_rhs_bitwise": rhs_bitwise, "steps2_rel_l2": rel,
                      "steps2_counts_equal": (stg.accepted, stg.rejected, stg.rhsEvals)
                      == (sto.accepted, sto.rejected, sto.rhsEvals)},
            "aggregate": agg}


def host_cpus():
    """(nproc, the cgroup's CPU quota in whole CPUs or None): the cores this
    process may run on, and the share a CPU controller grants them."""
    nproc = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        if q != "max":
            quota = max(1, int(-(-int(q) // int(per))))
    except (OSError, ValueError):
        pass
    return nproc, quota


def oracle_instance_bytes(nd):
    """Host bytes one oracle instance holds at nd nodes: the state, its RHS and
    the solver's arrays (nodal depths, face values and fluxes, scratch), about
    40 doubles per node."""
    return 40 * 8 * nd


def cpu_aggregate(case, OL, reps=2):
    """N independent single-threaded oracle instances at once, each evaluating
    the full-size RHS `reps` times: the host's aggregate rate. N = nproc (the
    cores this process may use), lowered only where the cgroup's CPU quota or
    the host memory available cannot hold that many instances; both figures
    are reported. ctypes releases the interpreter lock during the oracle's
    calls, so the threads run in parallel."""
    import threading
    nproc, quota = host_cpus()
    n = nproc
    if quota:
        n = min(n, quota)
    avail = None
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemAvailable:"):
                    avail = int(line.split()[1]) * 1024
    except OSError:
        pass
    mem_cap = None
    if avail:
        mem_cap = max(1, int(0.5 * avail // oracle_instance_bytes(case.nd)))
        n = min(n, mem_cap)
    orcs = [OL.Oracle.from_case(case) for _ in range(n)]
    start = threading.Barrier(n + 1)

    def work(o):
        start.wait()
        for _ in range(reps):
            o.rhs(case.U0)

    th = [threading.Thread(target=work, args=(o,)) for o in orcs]
    for x in th:
        x.start()
    start.wait()
    t = time.perf_counter()
    for x in th:
        x.join()
    wall = time.perf_counter() - t
    for o in orcs:
        o.close()
    value = 5.0 * case.nd * n * reps / wall
    log(f"[bench] cpu aggregate: {n} instances x {reps} RHS in {wall:.2f}s "
        f"({value / 1e6:.1f} MDOF-RHS/s)")
    return {"value": value, "unit": "DOF-RHS-evals/s", "cores": n, "instances": n,
            "nproc": nproc, "cgroup_cpu_quota": quota, "memory_cap_instances": mem_cap,
            "sample": (f"{n} concurrent single-threaded oracle instances (nproc {nproc}, "
                       f"cgroup CPU quota {quota}, memory room for {mem_cap}), "
                       f"{reps} full-size RHS each")}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500, help="accepted adaptive steps timed")
    ap.add_argument("--warmup", type=int, default=20, help="accepted steps of the warm-up run")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--batch-members", type=int, default=64,
                    help="members of the batched sweep leg (0 = skip it)")
    ap.add_argument("--batch-steps", type=int, default=10,
                    help="accepted steps per member in the batched sweep leg")
    ap.add_argument("--batch-small-members", type=int, default=64,
                    help="members of the small-grid sweep leg, Nx in [32, 128] (0 = skip)")
    ap.add_argument("--batch-small-steps", type=int, default=50)
    ap.add_argument("--adjoint-steps", type=int, default=20,
                    help="accepted steps of the gradient leg (0 = skip)")
    ap.add_argument("--inversion-steps", type=int, default=2000,
                    help="accepted steps of the gauge-inversion gradient leg (0 = skip)")
    ap.add_argument("--inversion-dirs", type=int, default=32,
                    help="finite-difference directions checked on the inversion gradient")
    ap.add_argument("--adjoint-members", type=int, default=16,
                    help="members of the batched adjoint leg (0 = skip)")
    ap.add_argument("--adjoint-member-steps", type=int, default=60)
    ap.add_argument("--margin-steps", type=int, default=1000,
                    help="accepted steps of the 2000x1000 margin leg (0 = skip)")
    ap.add_argument("--margin-adjoint-steps", type=int, default=100,
                    help="accepted steps of the margin source-inversion gradient (0 = skip)")
    ap.add_argument("--margin-adjoint-dirs", type=int, default=3,
                    help="finite-difference directions checked on the margin gradient")
    ap.add_argument("--profile-rhs", type=int, default=0)
    ap.add_argument("--profile-vjp", type=int, default=0,
                    help="profile mode: this many RHS vector-Jacobian products")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="with --profile-rhs: also run this many adaptive steps")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: sgnx runs on one GPU only (asked for {max(args.gpus, world)}); "
              "a decomposition over several GPUs is designed but not implemented",
              file=sys.stderr)
        sys.exit(2)
    if args.profile_rhs:
        profile_rhs(args)
        return
    if args.profile_vjp:
        profile_vjp(args)
        return

    import sgnx
    case = build_case(args)
    s = sgnx.SGN2D(**case.solver_kwargs())
    s.upload(case.U0)
    dt0 = s.cflTimeStep(None, 0.5)
    opts = dict(atol=case.meta["atol"], rtol=case.meta["rtol"])

    # warm-up (untimed), then reset the state as apps/sgn_bench.cu:120 does
    t = time.perf_counter()
    stw = s.integrateAdaptive(0.0, 1e9, dt0, sgnx.RkOpts(max_accepted=args.warmup, **opts))
    log(f"[bench] warm-up: {stw.accepted} acc / {stw.rejected} rej in "
        f"{time.perf_counter() - t:.2f}s")
    s.upload(case.U0)
    s.enable_timing(True)
    s.rhs_timing()                                  # clear, and wait for the device
    t0 = time.perf_counter()
    st = s.integrateAdaptive(0.0, 1e9, dt0, sgnx.RkOpts(max_accepted=args.steps, **opts))
    wall = time.perf_counter() - t0                  # integrateAdaptive ends with a sync
    timing = s.rhs_timing()
    s.enable_timing(False)
    if st.aborted or st.accepted != args.steps:
        log(f"[bench] run did not complete: {st.as_dict()}")
        sys.exit(1)

    nd = case.nd
    dof = 5.0 * nd
    value = dof * st.rhsEvals / wall
    launches = sum(n for n, _ in timing)
    rhs_ms = sum(ms for _, ms in timing)
    mean_rhs_s = rhs_ms / launches / 1e3
    # the RHS's algorithmic bytes (104 B/node) per launch over the mean launch
    # duration; and, beside it, the bytes each fused launch kind also moves for
    # its RK stage operands
    achieved = ALG_BYTES_PER_NODE * nd / mean_rhs_s / 1e9
    alg_stage_per_launch = sum(ALG_BYTES_BY_KIND[m] * nd * timing[m][0]
                               for m in range(len(timing))) / launches
    achieved_stage = alg_stage_per_launch / mean_rhs_s / 1e9
    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE runs,
    # copy-calibrated, per launch kind), weighted by this run's launch mix
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "rhs_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        bk = tr.get("by_kind", {})
        if (tr.get("Nx"), tr.get("Ny")) == (case.Nx, case.Ny) and all(
                k in bk for m, k in enumerate(KIND_NAMES) if timing[m][0]):
            traffic = sum(bk[k]["bytes"] * timing[m][0]
                          for m, k in enumerate(KIND_NAMES) if timing[m][0]) / launches
    log(f"[bench] timed: {st.accepted} acc / {st.rejected} rej / {st.rhsEvals} RHS in "
        f"{wall:.3f}s -> {value / 1e9:.2f} GDOF-RHS/s; RHS kernel mean {mean_rhs_s * 1e3:.4f} ms "
        f"over {launches} launches = {achieved:.0f} GB/s algorithmic (104 B/node), "
        f"{achieved_stage:.0f} GB/s with the stage operands")
    plain = plain_rhs_rate(s, nd)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(case, s, dt0, args)
    s.close()
    batched = None
    if args.batch_members > 0:
        try:
            batched = batched_sweep(args, single_rate=value)
        except Exception as e:                    # reported, not fatal to the main line
            batched = {"error": f"{type(e).__name__}: {e}"}
            log(f"[bench] batch leg failed: {batched['error']}")

    batched_small = None
    if args.batch_small_members > 0:
        try:
            batched_small = batched_sweep(args, nx_min=32, nx_max=128,
                                          n_members=args.batch_small_members,
                                          steps=args.batch_small_steps, single_rate=value)
        except Exception as e:                    # reported, not fatal to the main line
            batched_small = {"error": f"{type(e).__name__}: {e}"}
            log(f"[bench] small-grid batch leg failed: {batched_small['error']}")

    adjoint = None
    if args.adjoint_steps > 0:
        try:
            adjoint = adjoint_leg(args)
        except Exception as e:                    # reported, not fatal to the main line
            adjoint = {"error": f"{type(e).__name__}: {e}"}
            log(f"[bench] adjoint leg failed: {adjoint['error']}")

    inversion = None
    if args.inversion_steps > 0:
        try:
            inversion = inversion_leg(args)
        except Exception as e:                    # reported, not fatal to the main line
            inversion = {"error": f"{type(e).__name__}: {e}"}
            log(f"[bench] inversion leg failed: {inversion['error']}")

    batched_adjoint = None
    if args.adjoint_members > 0:
        try:
            batched_adjoint = batched_adjoint_leg(args)
        except Exception as e:                    # reported, not fatal to the main line
            batched_adjoint = {"error": f"{type(e).__name__}: {e}"}
            log(f"[bench] batched adjoint leg failed: {batched_adjoint['error']}")

    margin = margin_adjoint = None
    if args.margin_steps > 0 or args.margin_adjoint_steps > 0:
        mcase = margin_case(args)
        if args.margin_steps > 0:
            try:
                margin = margin_leg(args, mcase)
            except Exception as e:                # reported, not fatal to the main line
                margin = {"error": f"{type(e).__name__}: {e}"}
                log(f"[bench] margin leg failed: {margin['error']}")
        if args.margin_adjoint_steps > 0:
            try:
                margin_adjoint = margin_adjoint_leg(args, mcase)
            except Exception as e:                # reported, not fatal to the main line
                margin_adjoint = {"error": f"{type(e).__name__}: {e}"}
                log(f"[bench] margin adjoint leg failed: {margin_adjoint['error']}")
        del mcase

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-RHS-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / st.accepted,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"shoal {case.Nx}x{case.Ny} P=4 FP64 on 96x48 m: 1:20 slope + seeded "
                         f"elliptic shoal (seed {args.seed}), 1 s / 0.01 m inflow on the left, "
                         f"walls elsewhere; {args.steps} accepted adaptive BS3(2) steps, "
                         f"atol 1e-8 rtol 1e-6"),
            "Nx": case.Nx, "Ny": case.Ny, "P": case.P, "dof": int(dof),
            "accepted": st.accepted, "rejected": st.rejected, "rhs_evals": st.rhsEvals,
            "wall_s": wall, "t_reached": st.tReached,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_rhs_sweep (FR/SGN right-hand side with the RK stage fused)",
            "basis": ("104 algorithmic bytes per node per RHS launch (U read, dU written, "
                      "H, Hx, Hy read) over the launches' mean duration"),
            "alg_bytes_per_launch": ALG_BYTES_PER_NODE * nd,
            "mean_launch_ms": mean_rhs_s * 1e3,
            "launches": launches,
            "achieved_rhs_bytes_only": achieved,
            "frac_rhs_bytes_only": achieved / HBM_PEAK_GBS,
            "achieved_stage_bytes": achieved_stage,
            "frac_stage_bytes": achieved_stage / HBM_PEAK_GBS,
            "alg_stage_bytes_per_launch": alg_stage_per_launch,
            "alg_stage_bytes_per_node_mean": alg_stage_per_launch / nd,
            "plain_rhs": plain,
            "launches_by_kind": {k: timing[m][0] for m, k in enumerate(KIND_NAMES)},
            "mean_ms_by_kind": {k: (timing[m][1] / timing[m][0] if timing[m][0] else None)
                                for m, k in enumerate(KIND_NAMES)},
            "alg_bytes_per_node_by_kind": dict(zip(KIND_NAMES, ALG_BYTES_BY_KIND)),
        },
        "cpu_baseline": cpu,
        "batched": batched,
        "batched_small": batched_small,
        "adjoint": adjoint,
        "inversion": inversion,
        "batched_adjoint": batched_adjoint,
        "margin": margin,
        "margin_adjoint": margin_adjoint,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
