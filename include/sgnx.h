/*
 * sgnx.h -- C-ABI of libsgnx: MI355X (gfx950) HIP implementation of the 2D
 * flux-reconstruction (P = 4, Gauss-Lobatto) right-hand side of the hyperbolic
 * Serre-Green-Naghdi system and of its adaptive Bogacki-Shampine 3(2) step.
 *
 * Plain C types only (pointers, sizes, doubles): every entry point returns an
 * int status (SGNX_OK = 0) and never aborts; sgnx_last_error() gives the text.
 * A solver handle is single-threaded and not re-entrant, like the reference's
 * device class. Each entry point below names the reference interface it
 * replaces (paths relative to the reference checkout).
 *
 * State layout (unchanged from the reference, src/solver/sgn2d.h:22-27):
 *   U[v*Nd + e*25 + a + 5*b],  v in {h, hu, hv, hw, hp},  e = ex + Nx*ey,
 *   node (a, b) of element e,  Nd = 25*Nx*Ny nodes per variable.
 */
#ifndef SGNX_H
#define SGNX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SGNX_OK 0
#define SGNX_EINVAL 1       /* bad argument or descriptor */
#define SGNX_EDEVICE 2      /* HIP runtime error */
#define SGNX_ENOMEM 3       /* device allocation failed */
#define SGNX_EUNSUPPORTED 4 /* valid for the reference but not implemented here */

/* side codes, side order (left, right, bottom, top) */
#define SGNX_BC_PERIODIC 0  /* the reference's only kind (src/core/mesh2d.h:36-45) */
#define SGNX_BC_WALL 1      /* reflective: ghost = face state, normal momentum negated */
#define SGNX_BC_INFLOW 2    /* prescribed ghost (left side only) */

/* where the pointers passed to sgnx_rhs live */
#define SGNX_MEM_HOST 0
#define SGNX_MEM_DEVICE 1

/* Physical constants; the same four fields and meaning as sgn2d::Params
 * (src/physics/sgn_flux2d.h:57-62). */
typedef struct sgnx_params {
  double g, c, gamma, nManning;
} sgnx_params;

/* Everything the device solver needs, taken from the host setup; the reference
 * device class reads the same items from its host SGN2D object
 * (src/cuda/sgn2d_cuda.cu:318-400). All pointers are host pointers and are
 * copied during sgnx_create. */
typedef struct sgnx_desc {
  int P;                       /* polynomial degree; the kernels implement P = 4 */
  int Nx, Ny;                  /* elements in x and y (>= 2) */
  double Lx, Ly;               /* domain size; hx = Lx/Nx, Jx = hx/2 (mesh2d.h:29-50) */
  sgnx_params par;
  const double *D;             /* (P+1)^2 row-major differentiation matrix */
  const double *lL, *lR;       /* face interpolation rows (must be GLL unit rows) */
  const double *gLp, *gRp;     /* FR correction derivatives at the nodes */
  const double *qw;            /* quadrature weights (mass, energy) */
  const double *Hnod;          /* Nd nodal still depth, or NULL = flat H == 0 */
  const double *Hx, *Hy;       /* Nd gradients, or NULL = computed on the device */
  /* still depth on every interface line (NULL with Hnod == NULL):
   *   HfaceX[(ey*(Nx+1) + i)*(P+1) + t]  x-interface i = 0..Nx, line node t
   *   HfaceY[(j*Nx + ex)*(P+1) + t]      y-interface j = 0..Ny, line node t
   * For a periodic axis the last line repeats the first (the reference samples
   * its right/top edge depth at the wrapped coordinate, sgn2d.cpp:75-83). */
  const double *HfaceX, *HfaceY;
  int bc[4];                   /* SGNX_BC_* for left, right, bottom, top */
  /* SGNX_BC_INFLOW on the left side: per face node (ey, t) and variable v the
   * ghost is (base + s*sin(w t)) + c*cos(w t); table [(ey*(P+1)+t)*15 + 3v + {0,1,2}]. */
  const double *inflow;
  double inflow_omega;
} sgnx_desc;

/* AdaptiveRKOptions (src/solver/rk.h:53-61) plus max_accepted (0 = off): stop
 * after that many accepted steps, so a fixed step count runs in ONE call and
 * keeps the PI controller memory. */
typedef struct sgnx_rk_opts {
  double atol, rtol, dtMin, dtMax, safety, facMin, facMax;
  long max_accepted;
} sgnx_rk_opts;

/* AdaptiveRKStats (src/solver/rk.h:64-72). */
typedef struct sgnx_rk_stats {
  long accepted, rejected, rhsEvals;
  double tReached, dtLast, maxNormLast;
  int aborted;
} sgnx_rk_stats;

typedef struct sgnx_solver sgnx_solver;

/* Text of the last failure on this handle; with s == NULL, of the last
 * sgnx_create on the calling thread. */
const char* sgnx_last_error(const sgnx_solver* s);

/* Replaces SGN2DCuda::SGN2DCuda(const SGN2D&) (src/cuda/sgn2d_cuda.h:33,
 * sgn2d_cuda.cu:318-400): allocates the device state, stages and tables. */
int sgnx_create(const sgnx_desc* d, sgnx_solver** out);
/* Replaces SGN2DCuda::~SGN2DCuda (sgn2d_cuda.cu:403-412). */
int sgnx_destroy(sgnx_solver* s);
/* Stream for all work of this handle (a hipStream_t; NULL = default stream).
 * Not in the reference, which uses the legacy default stream only. */
int sgnx_set_stream(sgnx_solver* s, void* stream);
/* Nodes per variable Nd (SGN2DCuda::nDof, sgn2d_cuda.h:38). */
long long sgnx_ndof(const sgnx_solver* s);
/* Device address of the resident state (5*Nd doubles), for zero-copy callers. */
void* sgnx_state_ptr(sgnx_solver* s);

/* Replace SGN2DCuda::upload / download (sgn2d_cuda.h:41-42). n = 5*Nd. */
int sgnx_upload(sgnx_solver* s, const double* U, size_t n);
int sgnx_download(const sgnx_solver* s, double* U, size_t n);

/* Replaces SGN2D::rhs(U, dUdt) (src/solver/sgn2d.cpp:125-290) and
 * SGN2DCuda::rhsDevice (sgn2d_cuda.cu:429-452): dU = R(U, t), bitwise equal to
 * the CPU rhs() for finite U. t only enters inflow ghosts. mem = SGNX_MEM_HOST
 * (copied through the device) or SGNX_MEM_DEVICE (both pointers on this
 * device, U and dU distinct). */
int sgnx_rhs(sgnx_solver* s, const double* U, double* dU, size_t n, double t, int mem);
/* Replaces SGN2DCuda::rhsOnce (sgn2d_cuda.cu:456-462): RHS of the resident
 * state into host dU. */
int sgnx_rhs_once(sgnx_solver* s, double* dU, size_t n, double t);

/* Replaces SGN2DCuda::integrateAdaptive (sgn2d_cuda.cu:467-540) with the
 * arithmetic of sgn::integrateAdaptive (src/solver/rk.cpp:67-164) on the
 * resident state. step_log (optional, capacity cap steps) receives (dt,
 * errNorm) of every attempted step. */
int sgnx_integrate(sgnx_solver* s, double t0, double tEnd, double dt0, const sgnx_rk_opts* o,
                   sgnx_rk_stats* st, double* step_log, long cap);
/* Replace sgn::rk4Step / sgn::bs3Step (src/solver/rk.cpp:23-54) on the
 * resident state. */
int sgnx_rk4_step(sgnx_solver* s, double t, double dt);
int sgnx_bs3_step(sgnx_solver* s, double t, double dt);

/* Replace SGN2D::maxWaveSpeed / cflTimeStep / mass / energy
 * (sgn2d.cpp:295-350) on the resident state. maxWaveSpeed and cflTimeStep are
 * bitwise; mass and energy are fixed-order tree sums. */
int sgnx_max_wave_speed(sgnx_solver* s, double* out);
int sgnx_cfl_time_step(sgnx_solver* s, double cfl, double* out);
int sgnx_mass(sgnx_solver* s, double* out);
int sgnx_energy(sgnx_solver* s, double* out);
/* The still-depth gradients the solver holds (Nd each, host pointers). */
int sgnx_bathymetry(const sgnx_solver* s, double* Hx, double* Hy);

/* RHS kernel choice (tuning, no counterpart in the reference): variant 0 =
 * one workgroup per 5-element row tile, 1 = row sweep over a strip of 5
 * element columns and `rows` element rows per workgroup (default 1, rows 16,
 * or fewer on grids too small to give every compute unit three workgroups;
 * the environment variables SGNX_RHS_VARIANT / SGNX_SWEEP_ROWS set the
 * defaults at creation). Both are bitwise equal to the reference rhs(). The
 * row sweep has no Manning friction term: a solver with nManning != 0 uses
 * variant 0 (variant 1 then gives SGNX_EUNSUPPORTED, and such a solver
 * cannot join a batch or be autotuned). */
int sgnx_set_rhs_kernel(sgnx_solver* s, int variant, int rows);

/* The RHS kernel configuration in use (no counterpart in the reference; for
 * callers that record or reproduce a run's settings): the variant and tile
 * height of sgnx_set_rhs_kernel, and the row sweep's waves mask (bit k set:
 * launch kind k = plain, stage 2, stage 3, k4 built for three waves per SIMD),
 * as set at creation (SGNX_SWEEP_WAVES3, decimal or 0x-hexadecimal) or by
 * sgnx_autotune. Any output pointer may be null. */
int sgnx_get_rhs_kernel(const sgnx_solver* s, int* variant, int* rows, int* waves3);

/* Launch timing (instrumentation, no counterpart in the reference): while
 * enabled, every RHS launch of this handle is bracketed by HIP events on the
 * handle's stream. sgnx_rhs_timing waits for the stream, fills
 * launches[SGNX_TIMING_MODES] and total_ms[SGNX_TIMING_MODES] (summed device
 * milliseconds) by launch kind: [0] R of a plain state, [1] R(u + a k) (stage 2
 * of a step), [2] R(u + a k) writing the candidate uNew and the partial
 * embedded sum instead of R (stage 3), [3] k4 = R(uNew) with the scaled error
 * partials, and clears the record. */
#define SGNX_TIMING_MODES 4
int sgnx_enable_timing(sgnx_solver* s, int on);
int sgnx_rhs_timing(sgnx_solver* s, long* launches, double* total_ms);

/* ---- diagnostics of the resident state (the host loops of the reference's
 * applications, on the device) ------------------------------------------------
 * Max of eta = h - H over the nodes strictly inside the window (x0, x1) x
 * (y0, y1) (pass +-INFINITY for no bound) and the lowest node index attaining
 * it, which is what a sequential scan with a strict ">" finds
 * (apps/shoal_focus_mpi.cu:196-231, apps/tsunami_margin_mpi.cu:255-290); no
 * node in the window gives (-1e300, -1). */
int sgnx_eta_max(sgnx_solver* s, double x0, double x1, double y0, double y1, double* eta,
                 long long* node);
/* eta = h - H at n nodes (the gauges of apps/tsunami_margin_mpi.cu:274-278). */
int sgnx_gauges(sgnx_solver* s, const long long* nodes, int n, double* eta);
/* sgnx_gauges without allocating: no counterpart in the reference. For callers
 * that cannot allocate device memory (serving loops, harnesses that forbid
 * hipMalloc). The node list and the values pass through a static device array
 * of the library (64 KiB per GPU, held until the process ends), so at most
 * SGNX_GAUGES_NOALLOC_MAX gauges per call (more: SGNX_EINVAL). Every call on
 * one GPU shares that array: calls that overlap on one GPU (several host
 * threads, or solvers on different streams) race. Same values as
 * sgnx_gauges, bit for bit. */
#define SGNX_GAUGES_NOALLOC_MAX 4096
int sgnx_gauges_noalloc(sgnx_solver* s, const long long* nodes, int n, double* eta);
/* Element means of eta, u, v, p, H as src/io/vtk2d.h:39-81 forms them (0.25 *
 * sum over b, a of qw[a] qw[b] f): out[f * Nx * Ny + ey * Nx + ex], 5 fields. */
int sgnx_element_means(sgnx_solver* s, double* out);
/* Gauges recorded by sgnx_integrate (and batches) after every accepted step:
 * eta at n nodes, up to cap rows (set_gauges clears the log; n = 0 stops it).
 * gauge_log copies min(cap, rows recorded) rows of n values and, if t is
 * given, the time of each row. */
int sgnx_set_gauges(sgnx_solver* s, const long long* nodes, int n, long cap);
int sgnx_gauge_log(sgnx_solver* s, double* eta, double* t, long cap, long* rows);

/* ---- batches of independent runs (no counterpart in the reference, whose
 * sweeps run as separate jobs: scripts/shoal_sweep.sbatch) -------------------
 * Solvers of uneven sizes advanced together: each RHS launch of an adaptive
 * step covers every running member (their row-sweep workgroups packed behind
 * a per-member tile offset table), one launch per stage kind. Each member keeps
 * its own buffers, time, step size, controller and tile decomposition, so its
 * state and statistics are bit for bit those of sgnx_integrate on it alone.
 * Members must use the row-sweep kernel on one device and outlive the batch,
 * which runs on a stream of its own and waits for the members' streams first. */
typedef struct sgnx_batch sgnx_batch;
#define SGNX_BATCH_GRAPH 1   /* capture a step into a hipGraph, replay it while
                                the set of running members is unchanged */
int sgnx_batch_create(sgnx_solver* const* members, int n, sgnx_batch** out);
int sgnx_batch_destroy(sgnx_batch* b);
const char* sgnx_batch_last_error(const sgnx_batch* b);   /* b may be NULL after create */
/* waves per SIMD the batched launches are built for, by launch kind (bit k set
 * = three; kinds as sgnx_rhs_timing); defaults to the first member's setting */
int sgnx_batch_set_waves(sgnx_batch* b, int waves3);
/* dU[m] = R(U[m]) at t[m] for every member in one launch (device pointers) */
int sgnx_batch_rhs(sgnx_batch* b, const double* const* U, double* const* dU, const double* t);
/* sgnx_integrate(member m, t0[m], tEnd[m], dt0[m], &o[m], &st[m], NULL, 0) for
 * every member, advanced in lockstep (members that finish drop out) */
int sgnx_batch_integrate(sgnx_batch* b, const double* t0, const double* tEnd, const double* dt0,
                         const sgnx_rk_opts* o, sgnx_rk_stats* st, int flags);

/* sgnx_adjoint for every member, their reverse sweeps in lockstep with one
 * launch per kind over all members (aligned at each member's last step), or,
 * when every member alone fills the GPU, the members' own sweeps back to
 * back; each member's cotangents are bit for bit its own sgnx_adjoint's.
 * Arrays of n pointers (Ubar_final, gauge_bar, theta_bar may be NULL, or hold
 * NULLs); every member's tape must hold every step's snapshot. */
int sgnx_batch_adjoint(sgnx_batch* b, const double* const* Ubar_final,
                       const double* const* gauge_bar, double* const* U0bar,
                       double* const* theta_bar, int mem);

/* Autotuning (no counterpart in the reference): times every launch kind of an
 * adaptive step on the resident state for each candidate tile height (element
 * rows per workgroup: 2, 4, 8, 12, 16, 24, 32, up to Ny) and each build
 * (two or three waves per SIMD), keeps the fastest, and sets it (row-sweep
 * kernel; reps = launches timed per candidate). Only scratch stage buffers are
 * written; U and k1 are untouched. The tile height fixes the summation order
 * of the error norm, so a run is reproducible bit for bit under the same
 * setting (a batch member and its solo run share it). rows_out / waves3_out
 * (optional) receive the choice. The batch form picks one tile height for
 * every member and the batch's waves. */
int sgnx_autotune(sgnx_solver* s, int reps, int* rows_out, int* waves3_out);
int sgnx_batch_autotune(sgnx_batch* b, int reps, int* rows_out, int* waves3_out);

/* ---- discrete adjoint (no counterpart in the reference, which has none) ----
 * Exact reverse-mode derivatives of the discrete map, nManning = 0 only
 * (SGNX_EUNSUPPORTED otherwise). theta is the nodal still depth: Hnod = theta,
 * Hx, Hy = (D theta)/J per element line (sgn2d.cpp:54-71), and each interface
 * line node's face depth the mean of the two edge values meeting there (the
 * wrapped pair on a periodic axis; the one interior value on a non-periodic
 * side). For a still depth sampled from a continuous function at the nodes
 * this is the reference's sampling (sgn2d.cpp:75-83). */
/* Sets the still depth from theta (Nd values; mem as sgnx_rhs). */
int sgnx_set_depth(sgnx_solver* s, const double* theta, size_t n, int mem);
/* Ubar = (dR/dU (U, t))^T W and, if theta_bar is given, theta_bar =
 * (dR/dtheta)^T W (Nd values); n = 5*Nd; mem as sgnx_rhs (device: Ubar
 * distinct from U and W). */
int sgnx_rhs_vjp(sgnx_solver* s, const double* U, const double* W, double* Ubar,
                 double* theta_bar, size_t n, double t, int mem);
/* Tape of adaptive runs: between tape_begin and tape_end, sgnx_integrate
 * (and sgnx_batch_integrate, for each member) records every accepted step (t, dt, gauge log row) and snapshots of step
 * input states, at most max_snapshots held (every K-th step; K doubles when
 * the budget is full). tape_info: steps recorded, stride K (0: planned
 * snapshots, sgnx_tape_begin_plan), snapshots held.
 * tape_replay: the recorded steps (fixed t, dt) on the resident state, with
 * gauge recording -- the map the adjoint differentiates. */
int sgnx_tape_begin(sgnx_solver* s, long max_snapshots);
/* Starts a tape like sgnx_tape_begin for a run of `steps` accepted steps
 * (known in advance, e.g. a max_accepted bound) whose reverse sweep may hold
 * `states` states at once, snapshots included: the snapshots are taken at the
 * steps of the first sweep of the binomial schedule (Griewank's revolve) for
 * that budget, and the sweep reverses segment j with states - j - 1 slots,
 * which recomputes the fewest steps the budget allows (sgnx_tape_set_slots is
 * not used). A run of fewer steps is reversed correctly with fewer
 * snapshots; the steps of a longer run past the plan form one segment.
 * steps <= 2^20 (SGNX_EINVAL beyond). No reference counterpart (the reference
 * has no adjoint). */
int sgnx_tape_begin_plan(sgnx_solver* s, long steps, long states);
/* The snapshot steps sgnx_tape_begin_plan takes for (steps, states), in
 * ascending order from 0 (up to cap of them into at; n = their count). Host
 * only, no device needed. */
int sgnx_tape_plan_positions(long steps, long states, long* at, long cap, long* n);
/* States the reverse sweep may hold besides the snapshots while it recomputes
 * a segment (-1, the default: all K - 1 of it). Fewer slots select a binomial
 * schedule (Griewank's revolve bound) inside each segment, plus two states of
 * scratch; the result is the same bit for bit. */
int sgnx_tape_set_slots(sgnx_solver* s, long slots);
/* on != 0 (set before the run; set before sgnx_tape_begin, it also keeps
 * the buffers of an earlier run for reuse): the taped run of the fused BS3 path also
 * keeps each accepted step's stage-2 and stage-3 inputs, which its stage
 * launches write as they form them (two more states per step), so the reverse
 * sweep recomputes no stage. They are dropped, and the sweep recomputes, once
 * the snapshot budget is full. tape_stages: the steps whose stages are kept.
 * The adjoint is the same bit for bit either way. */
int sgnx_tape_keep_stages(sgnx_solver* s, int on);
int sgnx_tape_stages(const sgnx_solver* s, long* steps);
/* of the last sgnx_adjoint: recorded steps it recomputed, slots it held at once */
int sgnx_adjoint_stats(const sgnx_solver* s, long* recomputed_steps, long* peak_slots);
int sgnx_tape_end(sgnx_solver* s);
int sgnx_tape_info(const sgnx_solver* s, long* steps, long* stride, long* snapshots);
int sgnx_tape_replay(sgnx_solver* s);
/* the recorded steps' start times and sizes (up to cap of each); steps = count */
int sgnx_tape_steps(const sgnx_solver* s, double* t, double* dt, long cap, long* steps);
/* Reverse sweep of the taped run (its step sizes held fixed): with Ubar_final
 * the cotangent of the final state (or NULL) and gauge_bar that of the gauge
 * log (rows x gauges as sgnx_gauge_log, or NULL), U0bar = the cotangent of
 * the initial state (5*Nd) and theta_bar (Nd, optional) that of the still
 * depth. The states inside a snapshot segment are recomputed (device memory:
 * snapshots + the slots of sgnx_tape_set_slots + 7 states of scratch). mem as
 * sgnx_rhs. */
int sgnx_adjoint(sgnx_solver* s, const double* Ubar_final, const double* gauge_bar,
                 double* U0bar, double* theta_bar, int mem);

/* ---- host setup (no device needed) ---------------------------------------- */

/* Replaces sgn::FROperators1D(P, GaussLobatto) (src/core/fr_operators.cpp:134-200):
 * nodes x, D, lL, lR, gL', gR', quadrature weights, each P+1 (D (P+1)^2). */
int sgnx_fr_operators(int P, double* x, double* D, double* lL, double* lR, double* gLp,
                      double* gRp, double* qw);
/* Replaces SGN2D::nodeCoords (src/solver/sgn2d.cpp:89-102): Nd each. */
int sgnx_node_coords(int P, int Nx, int Ny, double Lx, double Ly, double* x, double* y);
/* Sample points of HfaceX / HfaceY (layouts above), as sgn2d.cpp:75-83; with
 * wrapx / wrapy the last interface line uses the wrapped coordinate 0. */
int sgnx_face_coords(int P, int Nx, int Ny, double Lx, double Ly, int wrapx, int wrapy,
                     double* fxx, double* fxy, double* fyx, double* fyy);

#ifdef __cplusplus
}
#endif

#endif /* SGNX_H */
