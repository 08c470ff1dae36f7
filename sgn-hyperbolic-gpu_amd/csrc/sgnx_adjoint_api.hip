// =============================================================================
// sgnx_adjoint_api.hip -- C-ABI of the discrete adjoint: the RHS
// vector-Jacobian product, the still depth from one nodal field, and the
// reverse sweep over a recorded adaptive run. No counterpart in the reference.
//
// A tape records, while sgnx_integrate runs, every accepted step (t, dt, its
// gauge log row) and snapshots of the step input states: every K-th step,
// with K doubled (every other snapshot dropped) whenever the snapshot budget
// is full, so at most `max_snapshots` states are held whatever the run's
// length; or, for a run whose step count is given in advance (a planned
// tape), at the steps of the binomial schedule's first forward sweep, so that
// the reversal recomputes the fewest steps `states` states allow (Griewank's
// bound). The reverse sweep walks the snapshot segments backwards; inside a
// segment it recomputes the states with the forward's own stage kernels
// (bitwise the recorded trajectory), holding at most `slots` of them at once
// (a binomial schedule, Griewank's revolve bound, when that is fewer than the
// segment's K - 1), and then runs, per step n from the last,
//   k1 = R(u), y2 = u + dt/2 k1, k2 = R(y2), y3 = u + 3dt/4 k2    (rk.cpp:99-108)
//   y3b = J(y3)^T (4/9 dt lam)
//   y2b = J(y2)^T (1/3 dt lam + 3/4 dt y3b)
//   lam = lam + y3b + y2b + J(u)^T (2/9 dt lam + 1/2 dt y2b)
// with the gauge cotangents added to lam (and subtracted from theta's, the
// gauge value being h - theta) before the step that produced their row. The
// step sizes are those of the recording (the controller's dependence of dt on
// the state is not differentiated), which makes the result the exact gradient
// of the replayed run, sgnx_tape_replay.
// =============================================================================

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <new>
#include <string>
#include <vector>

#include "sgnx_tape.h"

void chain_coef(VjpArgs& a, VjpArgs& b, double* buf) {
  a.coefNext = buf;
  a.Un = b.U;
  a.An = b.A;
  a.can = b.ca;
  a.cbn = b.cb;
  b.coef = buf;
  b.coefReady = true;
}

namespace {

int fail(sgnx_solver* s, int code, const std::string& msg) {
  s->err = msg;
  return code;
}

#define HIPCHK(s, call)                                                               \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail((s), e_ == hipErrorOutOfMemory ? SGNX_ENOMEM : SGNX_EDEVICE,        \
                  std::string(#call) + ": " + hipGetErrorString(e_));                 \
  } while (0)

int alloc_state(sgnx_solver* s, double** p, size_t doubles) {
  if (*p) return SGNX_OK;
  HIPCHK(s, hipMalloc((void**)p, doubles * sizeof(double)));
  return SGNX_OK;
}

int check_adjoint_support(sgnx_solver* s) {
  if (s->ph.nManning != 0.0)
    return fail(s, SGNX_EUNSUPPORTED, "adjoint: the Manning friction term is not differentiated");
  return SGNX_OK;
}

// one RHS vector-Jacobian product on the solver's stream
int vjp(sgnx_solver* s, const VjpArgs& va, double t) {
  HIPCHK(s, launch_rhs_vjp(s->g, s->ph, va, ghost_time(s, t), s->stream));
  return SGNX_OK;
}

VjpArgs vjp_args(AdjTape* tp, const double* U, const double* A, double ca, const double* B,
                 double cb, double* out, double* thetaBar) {
  VjpArgs va{};
  va.U = U;
  va.A = A;
  va.ca = ca;
  va.B = B;
  va.cb = cb;
  va.out = out;
  va.thetaBar = thetaBar;
  va.coef = tp->coef;
  return va;
}

// u -> uNew by one BS3 step of size dt with the forward's kernels (fused stage
// launches on the row sweep, separate combinations otherwise): bitwise the
// state the adaptive loop accepted. Uses d_k1, d_k2, d_k3, d_tmp.
int bs3_fixed(sgnx_solver* s, const double* u, double* uNew, double t, double dt) {
  const size_t n = s->n();
  HIPCHK(s, launch_rhs(s->g, s->ph, u, s->d_k1, ghost_time(s, t), s->stream));
  if (s->g.variant == 1) {
    Stage a{};
    a.u = u;
    a.k1 = s->d_k1;
    a.a = 0.5 * dt;
    HIPCHK(s, launch_rhs_stage(s->g, s->ph, kAxpy, a, s->d_k2, ghost_time(s, t + 0.5 * dt),
                               s->stream));
    Stage b{};
    b.u = u;
    b.k1 = s->d_k2;
    b.a = 0.75 * dt;
    b.kc = s->d_k1;
    b.dt = dt;
    b.uNew = uNew;
    b.zOut = s->d_k3;
    HIPCHK(s, launch_rhs_stage(s->g, s->ph, kCand, b, nullptr, ghost_time(s, t + 0.75 * dt),
                               s->stream));
  } else {
    HIPCHK(s, launch_comb1(s->d_tmp, u, 0.5 * dt, s->d_k1, n, s->stream));
    HIPCHK(s, launch_rhs(s->g, s->ph, s->d_tmp, s->d_k2, ghost_time(s, t + 0.5 * dt), s->stream));
    HIPCHK(s, launch_comb1(s->d_tmp, u, 0.75 * dt, s->d_k2, n, s->stream));
    HIPCHK(s, launch_rhs(s->g, s->ph, s->d_tmp, s->d_k3, ghost_time(s, t + 0.75 * dt), s->stream));
    HIPCHK(s, launch_bs3_update(uNew, u, dt, s->d_k1, s->d_k2, s->d_k3, n, s->stream));
  }
  return SGNX_OK;
}

// lam <- (d u_{n+1} / d u_n)^T lam for the step u -> u_{n+1} of size dt at t,
// whose stage inputs are y2, y3; thetaBar += the still-depth part. The three
// products are chained: each of the first two also writes the next one's face
// coefficients (tp->coef2, then tp->coef).
//   uNew = u + dt ((2/9 k1 + 1/3 k2) + 4/9 k3)   (rk.cpp:106-108)
int step_adjoint_stored(sgnx_solver* s, AdjTape* tp, const double* u, const double* y2,
                        const double* y3, double t, double dt, double* thetaBar) {
  int rc;
  VjpArgs a3 = vjp_args(tp, y3, tp->lam, dt * (4.0 / 9.0), nullptr, 0.0, tp->y3b, thetaBar);
  VjpArgs a2 = vjp_args(tp, y2, tp->lam, dt * (1.0 / 3.0), tp->y3b, 0.75 * dt, tp->y2b, thetaBar);
  VjpArgs a1 = vjp_args(tp, u, tp->lam, dt * (2.0 / 9.0), tp->y2b, 0.5 * dt, tp->lam, thetaBar);
  a1.acc0 = tp->lam;
  a1.acc1 = tp->y3b;
  a1.acc2 = tp->y2b;
  chain_coef(a3, a2, tp->coef2);
  chain_coef(a2, a1, tp->coef);
  if ((rc = vjp(s, a3, t + 0.75 * dt))) return rc;
  if ((rc = vjp(s, a2, t + 0.5 * dt))) return rc;
  return vjp(s, a1, t);
}

// step_adjoint_stored with the stage inputs recomputed from u.
// Uses d_k1, d_k2, d_k3, d_tmp.
int step_adjoint(sgnx_solver* s, AdjTape* tp, const double* u, double t, double dt,
                 double* thetaBar) {
  const size_t n = s->n();
  double* y2 = s->d_tmp;
  double* y3 = s->d_k3;
  HIPCHK(s, launch_rhs(s->g, s->ph, u, s->d_k1, ghost_time(s, t), s->stream));
  if (s->g.variant == 1) {
    // the forward's stage-2 launch, keeping its input and forming the next
    // one: y2, k2 = R(y2) and y3 in one pass
    Stage a{};
    a.u = u;
    a.k1 = s->d_k1;
    a.a = 0.5 * dt;
    a.yOut = y2;
    a.y3Out = y3;
    a.a3 = 0.75 * dt;
    HIPCHK(s, launch_rhs_stage(s->g, s->ph, kAxpy, a, s->d_k2, ghost_time(s, t + 0.5 * dt),
                               s->stream));
  } else {
    HIPCHK(s, launch_comb1(y2, u, 0.5 * dt, s->d_k1, n, s->stream));
    HIPCHK(s, launch_rhs(s->g, s->ph, y2, s->d_k2, ghost_time(s, t + 0.5 * dt), s->stream));
    HIPCHK(s, launch_comb1(y3, u, 0.75 * dt, s->d_k2, n, s->stream));
  }
  return step_adjoint_stored(s, tp, u, y2, y3, t, dt, thetaBar);
}

int ensure_vjp_buffers(sgnx_solver* s, AdjTape* tp) {
  int rc;
  if ((rc = alloc_state(s, &tp->coef, (size_t)vjp_coef_doubles(s->g)))) return rc;
  if ((rc = alloc_state(s, &tp->coef2, (size_t)vjp_coef_doubles(s->g)))) return rc;
  if ((rc = alloc_state(s, &tp->thb, (size_t)s->g.Nd))) return rc;
  return SGNX_OK;
}

// ---- recomputation inside a snapshot segment ------------------------------------
// the largest step count a tape plan covers
constexpr long kMaxPlanSteps = 1L << 20;

// F(L, c): forward steps needed to reverse L steps from a held start state
// with c free state slots, F(1, c) = 0, F(L, 0) = L (L - 1) / 2, and
// F(L, c) = min_m m + F(L - m, c - 1) + F(m, c): Griewank's binomial bound
// r L - binom(C + r, C + 1) with C = c + 1 checkpoints (the start state's
// included) and r the least repetition number with binom(C + r, C) >= L.
double binom_capped(long n, long k) {
  k = std::min(k, n - k);  // binom(n, k) = binom(n, n - k): the shorter product
  double r = 1.0;
  for (long i = 1; i <= k; ++i) {
    r = r * (double)(n - k + i) / (double)i;
    if (r > 1e30) return 1e30;
  }
  return r;
}

long fwd_cost(long L, long c) {
  if (L <= 1) return 0;
  if (c <= 0) return L * (L - 1) / 2;
  const long C = c + 1;
  long r = 1;
  while (binom_capped(C + r, C) < (double)L) ++r;
  return r * L - (long)binom_capped(C + r, C + 1);
}

// the first checkpoint of an optimal schedule: advance m steps and keep the
// state. g(m) = m + F(L - m, c - 1) + F(m, c) is convex in m (F(., c) has
// nondecreasing increments, the repetition numbers), so its smallest
// minimiser is the first m whose forward difference is not negative: a
// binary search, the same m a scan over 1 .. L - 1 picks first.
long best_split(long L, long c) {
  auto g = [&](long m) { return m + fwd_cost(L - m, c - 1) + fwd_cost(m, c); };
  long lo = 1, hi = L - 1;
  while (lo < hi) {
    const long mid = lo + (hi - lo) / 2;
    if (g(mid + 1) - g(mid) >= 0) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// The snapshot steps of a planned tape: the first forward sweep of the
// binomial schedule for `steps` steps with `states` states held at once (the
// start state included): p_0 = 0, p_{j+1} = p_j + best_split(steps - p_j, c - j)
// with c = states - 1, until no slot is left or the remaining range fits its
// free slots, whose every step is then a snapshot too (the taped run keeps
// them for free). Reversing segment j with c - j free slots, last segment
// first, then recomputes at most F(steps, c) - (steps - 1) steps.
std::vector<long> plan_positions(long steps, long states) {
  std::vector<long> at{0};
  const long c = states - 1;
  for (;;) {
    const long j = (long)at.size() - 1, rem = steps - at.back();
    if (rem < 2) break;
    if (c - j >= rem - 1) {
      for (long k = 1; k < rem; ++k) at.push_back(at.back() + 1);
      break;
    }
    if (c - j < 1) break;
    at.push_back(at.back() + best_split(rem, c - j));
  }
  return at;
}

int take_slot(sgnx_solver* s, AdjTape* tp, double** out) {
  if (tp->freeSeg.empty()) {
    double* p = nullptr;
    HIPCHK(s, hipMalloc((void**)&p, s->n() * sizeof(double)));
    tp->seg.push_back(p);
    tp->freeSeg.push_back(p);
  }
  *out = tp->freeSeg.back();
  tp->freeSeg.pop_back();
  tp->peakSlots = std::max(tp->peakSlots, ++tp->inUse);
  return SGNX_OK;
}

void give_slot(AdjTape* tp, double* p) {
  tp->freeSeg.push_back(p);
  --tp->inUse;
}

// u_{a+m} from u (holding u_a) by m recorded steps, into dst (or, dst null,
// into one of the ping-pong pair, returned in *out)
int advance(sgnx_solver* s, AdjTape* tp, const double* u, long a, long m, double* dst,
            const double** out) {
  const double* src = u;
  int rc;
  for (long k = 0; k < m; ++k) {
    double* to = (k == m - 1 && dst) ? dst : tp->pp[k & 1];
    if ((rc = bs3_fixed(s, src, to, tp->t[a + k], tp->dt[a + k]))) return rc;
    ++tp->recomputed;
    src = to;
  }
  if (out) *out = src;
  return SGNX_OK;
}

int reverse_step(sgnx_solver* s, AdjTape* tp, const double* u, long m, const double* gb) {
  if (gb && tp->row[m] >= 0 && tp->row[m] < s->grows)
    HIPCHK(s, launch_gauge_bar(gb + (size_t)tp->row[m] * s->nG, s->d_gnodes, s->nG, tp->lam,
                               tp->thb, s->stream));
  if (m < (long)tp->y2s.size())
    return step_adjoint_stored(s, tp, u, tp->y2s[m], tp->y3s[m], tp->t[m], tp->dt[m], tp->thb);
  return step_adjoint(s, tp, u, tp->t[m], tp->dt[m], tp->thb);
}

// reverse the recorded steps a .. a+L-1 from u (holding u_a) with c free
// slots; lam holds the cotangent of u_{a+L} and leaves with that of u_a
int reverse_range(sgnx_solver* s, AdjTape* tp, const double* u, long a, long L, long c,
                  const double* gb) {
  int rc;
  if (L >= 2 && c >= L - 1) {
    // every state of the range fits: one forward pass, then the reverse
    std::vector<double*> st(L - 1);
    const double* prev = u;
    for (long k = 0; k < L - 1; ++k) {
      if ((rc = take_slot(s, tp, &st[k]))) return rc;
      if ((rc = advance(s, tp, prev, a + k, 1, st[k], nullptr))) return rc;
      prev = st[k];
    }
    for (long k = L - 1; k >= 1; --k) {
      if ((rc = reverse_step(s, tp, st[k - 1], a + k, gb))) return rc;
      give_slot(tp, st[k - 1]);
    }
    return reverse_step(s, tp, u, a, gb);
  }
  while (L > 0) {
    if (L == 1) return reverse_step(s, tp, u, a, gb);
    if (c <= 0) {
      // no slot: recompute the last state of the range through the pair
      const double* v = nullptr;
      if ((rc = advance(s, tp, u, a, L - 1, nullptr, &v))) return rc;
      if ((rc = reverse_step(s, tp, v, a + L - 1, gb))) return rc;
      --L;
      continue;
    }
    const long m = best_split(L, c);
    double* slot = nullptr;
    if ((rc = take_slot(s, tp, &slot))) return rc;
    if ((rc = advance(s, tp, u, a, m, slot, nullptr))) return rc;
    if ((rc = reverse_range(s, tp, slot, a + m, L - m, c - 1, gb))) return rc;
    give_slot(tp, slot);
    L = m;
  }
  return SGNX_OK;
}

AdjTape* tape_of(sgnx_solver* s) {
  if (!s->tape) s->tape = new (std::nothrow) AdjTape();
  return s->tape;
}

void release_snapshots(AdjTape* tp) {
  for (double* p : tp->snap) tp->spare.push_back(p);
  if (tp->lastIn && !tp->lastInSnap) tp->spare.push_back(tp->lastIn);
  tp->lastIn = nullptr;
  tp->lastInSnap = false;
  tp->snap.clear();
  tp->snapAt.clear();
}

void release_stages(AdjTape* tp) {
  for (double* p : tp->y2s) tp->spare.push_back(p);
  for (double* p : tp->y3s) tp->spare.push_back(p);
  tp->y2s.clear();
  tp->y3s.clear();
}

int take_spare(sgnx_solver* s, AdjTape* tp, double** out) {
  if (!tp->spare.empty()) {
    *out = tp->spare.back();
    tp->spare.pop_back();
    return SGNX_OK;
  }
  HIPCHK(s, hipMalloc((void**)out, s->n() * sizeof(double)));
  return SGNX_OK;
}

}  // namespace

int adjoint_begin(sgnx_solver* s, bool storeAll, AdjTape** out) {
  AdjTape* tp = s->tape;
  if (!tp || tp->recording) return fail(s, SGNX_EINVAL, "adjoint: no finished tape");
  int rc;
  if ((rc = check_adjoint_support(s))) return rc;
  const long N = (long)tp->t.size();
  if (N > 0 && (tp->snap.empty() || tp->snapAt.size() != tp->snap.size() || tp->snapAt[0] != 0))
    return fail(s, SGNX_EINVAL, "adjoint: the tape lacks snapshots");
  if (storeAll && N > 0 && (long)tp->snap.size() != N)
    return fail(s, SGNX_EINVAL,
                "adjoint: a batched sweep needs every step's snapshot (max_snapshots >= steps)");
  if ((rc = ensure_vjp_buffers(s, tp))) return rc;
  for (double** p : {&tp->lam, &tp->y3b, &tp->y2b})
    if ((rc = alloc_state(s, p, s->n()))) return rc;
  *out = tp;
  return SGNX_OK;
}

int adjoint_inputs(sgnx_solver* s, AdjTape* tp, const double* Ubar_final, const double* gauge_bar,
                   int mem, hipStream_t st, const double** gb) {
  const size_t n = s->n(), Nd = (size_t)s->g.Nd;
  if (gauge_bar && tp->gnodes != s->gnodes)
    return fail(s, SGNX_EINVAL, "adjoint: the gauge nodes differ from those of the recording");
  const hipMemcpyKind in = mem == SGNX_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  if (Ubar_final) HIPCHK(s, hipMemcpyAsync(tp->lam, Ubar_final, n * sizeof(double), in, st));
  else HIPCHK(s, hipMemsetAsync(tp->lam, 0, n * sizeof(double), st));
  HIPCHK(s, hipMemsetAsync(tp->thb, 0, Nd * sizeof(double), st));
  *gb = nullptr;
  const size_t gdoubles = (size_t)s->grows * s->nG;
  if (gauge_bar && gdoubles > 0) {
    if (mem == SGNX_MEM_HOST) {
      if (tp->gbCap < gdoubles) {
        if (tp->gb) (void)hipFree(tp->gb);
        tp->gb = nullptr;
        tp->gbCap = 0;
        HIPCHK(s, hipMalloc((void**)&tp->gb, gdoubles * sizeof(double)));
        tp->gbCap = gdoubles;
      }
      HIPCHK(s, hipMemcpyAsync(tp->gb, gauge_bar, gdoubles * sizeof(double), in, st));
      *gb = tp->gb;
    } else {
      *gb = gauge_bar;
    }
  }
  return SGNX_OK;
}

int adjoint_outputs(sgnx_solver* s, AdjTape* tp, double* U0bar, double* theta_bar, int mem,
                    hipStream_t st) {
  const hipMemcpyKind outk = mem == SGNX_MEM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  HIPCHK(s, hipMemcpyAsync(U0bar, tp->lam, s->n() * sizeof(double), outk, st));
  if (theta_bar)
    HIPCHK(s, hipMemcpyAsync(theta_bar, tp->thb, (size_t)s->g.Nd * sizeof(double), outk, st));
  return SGNX_OK;
}

bool tape_recording(const sgnx_solver* s) { return s->tape && s->tape->recording; }

bool tape_stage_outputs(const sgnx_solver* s, double** y2, double** y3) {
  const AdjTape* tp = s->tape;
  if (!tp || !tp->recording || !tp->keepStages || tp->stagesOff) return false;
  if (tp->y2s.size() != tp->t.size() + 1) return false;
  *y2 = tp->y2s.back();
  *y3 = tp->y3s.back();
  return true;
}

int tape_step_begin(sgnx_solver* s) {
  AdjTape* tp = s->tape;
  const long n = (long)tp->t.size();
  int rc;
  if (tp->keepStages && !tp->stagesOff && s->g.variant == 1 && (long)tp->y2s.size() == n) {
    double *a = nullptr, *b = nullptr;
    if ((rc = take_spare(s, tp, &a))) return rc;
    if ((rc = take_spare(s, tp, &b))) {
      tp->spare.push_back(a);
      return rc;
    }
    tp->y2s.push_back(a);
    tp->y3s.push_back(b);
  }
  if (tp->planStates > 0) {
    // planned mode: the snapshot steps are fixed in advance
    tp->pendingSnap = tp->nextPlan < tp->planAt.size() && tp->planAt[tp->nextPlan] == n;
    return SGNX_OK;
  }
  if (n % tp->K != 0 || (long)tp->snap.size() > n / tp->K) return SGNX_OK;
  while ((long)tp->snap.size() >= tp->maxSnap) {
    // the stage states are kept only while every step's snapshot is
    tp->stagesOff = true;
    release_stages(tp);
    // budget full: keep every other snapshot, double the stride
    std::vector<double*> keep;
    std::vector<long> keepAt;
    for (size_t j = 0; j < tp->snap.size(); ++j) {
      if (j % 2 == 0) {
        keep.push_back(tp->snap[j]);
        keepAt.push_back(tp->snapAt[j]);
      } else {
        tp->spare.push_back(tp->snap[j]);
      }
    }
    tp->snap.swap(keep);
    tp->snapAt.swap(keepAt);
    tp->K *= 2;
    if (n % tp->K != 0) return SGNX_OK;
  }
  tp->pendingSnap = true;
  return SGNX_OK;
}

int tape_accept(sgnx_solver* s, double t, double dt, long row) {
  AdjTape* tp = s->tape;
  tp->t.push_back(t);
  tp->dt.push_back(dt);
  tp->row.push_back(row);
  if (tp->pendingSnap || tp->planStates > 0) {
    // no copy: the buffer holding the accepted step's input becomes the
    // snapshot, and the solver takes a spare one for its next candidate. A
    // planned tape also holds every accepted step's input until the next
    // acceptance (the last one is where the sweep starts).
    double* fresh = nullptr;
    const int rc = take_spare(s, tp, &fresh);
    if (rc) return rc;
    if (tp->planStates > 0) {
      if (tp->lastIn && !tp->lastInSnap) tp->spare.push_back(tp->lastIn);
      tp->lastIn = s->d_uNew;
      tp->lastInSnap = tp->pendingSnap;
    }
    if (tp->pendingSnap) {
      tp->snap.push_back(s->d_uNew);
      tp->snapAt.push_back((long)tp->t.size() - 1);
      if (tp->planStates > 0) ++tp->nextPlan;
    }
    s->d_uNew = fresh;
    tp->pendingSnap = false;
  }
  return SGNX_OK;
}

void tape_free(sgnx_solver* s) {
  AdjTape* tp = s->tape;
  if (!tp) return;
  release_snapshots(tp);
  release_stages(tp);
  for (double* p : tp->spare) (void)hipFree(p);
  for (double* p : tp->seg) (void)hipFree(p);
  for (double* p : {tp->lam, tp->y3b, tp->y2b, tp->coef, tp->coef2, tp->thb, tp->gb, tp->pp[0],
                    tp->pp[1]})
    if (p) (void)hipFree(p);
  delete tp;
  s->tape = nullptr;
}

extern "C" {

int sgnx_set_depth(sgnx_solver* s, const double* theta, size_t n, int mem) {
  if (!s || !theta) return SGNX_EINVAL;
  if (n != (size_t)s->g.Nd) return fail(s, SGNX_EINVAL, "set_depth: size mismatch");
  if (mem != SGNX_MEM_HOST && mem != SGNX_MEM_DEVICE)
    return fail(s, SGNX_EINVAL, "set_depth: bad memory kind");
  HIPCHK(s, hipMemcpyAsync(s->d_Hnod, theta, n * sizeof(double),
                           mem == SGNX_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                           s->stream));
  HIPCHK(s, launch_bathy_grad(s->g, s->d_Hnod, s->d_Hx, s->d_Hy, s->stream));
  HIPCHK(s, launch_face_depth(s->g, s->d_Hnod, s->d_HfX, s->d_HfY, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  // Hx, Hy are now the arithmetic the row sweep can repeat from Hnod
  const char* v = std::getenv("SGNX_SWEEP_HXR");
  s->g.hxr = !v || std::atoi(v) != 0;
  return SGNX_OK;
}

int sgnx_rhs_vjp(sgnx_solver* s, const double* U, const double* W, double* Ubar,
                 double* theta_bar, size_t n, double t, int mem) {
  if (!s || !U || !W || !Ubar) return SGNX_EINVAL;
  if (n != s->n()) return fail(s, SGNX_EINVAL, "rhs_vjp: size mismatch");
  if (mem != SGNX_MEM_HOST && mem != SGNX_MEM_DEVICE)
    return fail(s, SGNX_EINVAL, "rhs_vjp: bad memory kind");
  int rc;
  if ((rc = check_adjoint_support(s))) return rc;
  AdjTape* tp = tape_of(s);
  if (!tp) return fail(s, SGNX_ENOMEM, "host allocation failed");
  if ((rc = ensure_vjp_buffers(s, tp))) return rc;
  const size_t Nd = (size_t)s->g.Nd;
  const double *dU = U, *dW = W;
  double* dOut = Ubar;
  double* dTh = theta_bar;
  if (mem == SGNX_MEM_HOST) {
    HIPCHK(s, hipMemcpyAsync(s->d_tmp, U, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, hipMemcpyAsync(s->d_k2, W, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
    dU = s->d_tmp;
    dW = s->d_k2;
    dOut = s->d_k3;
    dTh = theta_bar ? tp->thb : nullptr;
  } else if (Ubar == U || Ubar == W) {
    return fail(s, SGNX_EINVAL, "rhs_vjp: Ubar must not overlap U or W");
  }
  if (dTh) HIPCHK(s, hipMemsetAsync(dTh, 0, Nd * sizeof(double), s->stream));
  VjpArgs va = vjp_args(tp, dU, dW, 1.0, nullptr, 0.0, dOut, dTh);
  if ((rc = vjp(s, va, t))) return rc;
  if (mem == SGNX_MEM_HOST) {
    HIPCHK(s, hipMemcpyAsync(Ubar, dOut, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    if (theta_bar)
      HIPCHK(s, hipMemcpyAsync(theta_bar, dTh, Nd * sizeof(double), hipMemcpyDeviceToHost,
                               s->stream));
  }
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return SGNX_OK;
}

int sgnx_tape_begin(sgnx_solver* s, long max_snapshots) {
  if (!s) return SGNX_EINVAL;
  if (max_snapshots < 1) return fail(s, SGNX_EINVAL, "tape: max_snapshots must be >= 1");
  AdjTape* tp = tape_of(s);
  if (!tp) return fail(s, SGNX_ENOMEM, "host allocation failed");
  HIPCHK(s, hipStreamSynchronize(s->stream));
  release_snapshots(tp);
  release_stages(tp);
  tp->stagesOff = false;
  tp->pendingSnap = false;
  // buffers beyond the new budget (three states a step while stages are
  // kept) are returned to the device
  const long cap = tp->keepStages ? 3 * max_snapshots : max_snapshots;
  while ((long)tp->spare.size() > cap) {
    (void)hipFree(tp->spare.back());
    tp->spare.pop_back();
  }
  tp->recording = true;
  tp->maxSnap = max_snapshots;
  tp->K = 1;
  tp->planStates = 0;
  tp->planAt.clear();
  tp->nextPlan = 0;
  tp->t.clear();
  tp->dt.clear();
  tp->row.clear();
  tp->gnodes = s->gnodes;
  return SGNX_OK;
}

int sgnx_tape_begin_plan(sgnx_solver* s, long steps, long states) {
  if (!s) return SGNX_EINVAL;
  if (steps < 1 || states < 1 || steps > kMaxPlanSteps)
    return fail(s, SGNX_EINVAL, "tape: the plan needs 1 <= steps <= 2^20 and states >= 1");
  std::vector<long> at = plan_positions(steps, states);
  int rc;
  if ((rc = sgnx_tape_begin(s, (long)at.size()))) return rc;
  AdjTape* tp = s->tape;
  tp->planStates = states;
  tp->planAt.swap(at);
  // stage inputs are kept only if every step has its snapshot
  tp->stagesOff = (long)tp->planAt.size() < steps;
  return SGNX_OK;
}

int sgnx_tape_plan_positions(long steps, long states, long* at, long cap, long* n) {
  if (!n || steps < 1 || steps > kMaxPlanSteps || states < 1 || cap < 0 || (cap > 0 && !at))
    return SGNX_EINVAL;
  const std::vector<long> p = plan_positions(steps, states);
  for (long k = 0; k < std::min((long)p.size(), cap); ++k) at[k] = p[k];
  *n = (long)p.size();
  return SGNX_OK;
}

int sgnx_tape_set_slots(sgnx_solver* s, long slots) {
  if (!s) return SGNX_EINVAL;
  if (slots < -1) return fail(s, SGNX_EINVAL, "tape: slots must be >= -1");
  AdjTape* tp = tape_of(s);
  if (!tp) return fail(s, SGNX_ENOMEM, "host allocation failed");
  tp->slots = slots;
  return SGNX_OK;
}

int sgnx_tape_keep_stages(sgnx_solver* s, int on) {
  if (!s) return SGNX_EINVAL;
  AdjTape* tp = tape_of(s);
  if (!tp) return fail(s, SGNX_ENOMEM, "host allocation failed");
  tp->keepStages = on != 0;
  return SGNX_OK;
}

int sgnx_tape_stages(const sgnx_solver* s, long* steps) {
  if (!s || !steps) return SGNX_EINVAL;
  const AdjTape* tp = s->tape;
  *steps = tp ? std::min((long)tp->y2s.size(), (long)tp->t.size()) : 0;
  return SGNX_OK;
}

int sgnx_adjoint_stats(const sgnx_solver* s, long* recomputed_steps, long* peak_slots) {
  if (!s || !recomputed_steps || !peak_slots) return SGNX_EINVAL;
  *recomputed_steps = s->tape ? s->tape->recomputed : 0;
  *peak_slots = s->tape ? s->tape->peakSlots : 0;
  return SGNX_OK;
}

int sgnx_tape_end(sgnx_solver* s) {
  if (!s) return SGNX_EINVAL;
  if (s->tape) {
    s->tape->recording = false;
    s->tape->pendingSnap = false;
  }
  return SGNX_OK;
}

int sgnx_tape_info(const sgnx_solver* s, long* steps, long* stride, long* snapshots) {
  if (!s || !steps || !stride || !snapshots) return SGNX_EINVAL;
  const AdjTape* tp = s->tape;
  *steps = tp ? (long)tp->t.size() : 0;
  *stride = tp ? (tp->planStates > 0 ? 0 : tp->K) : 1;
  *snapshots = tp ? (long)tp->snap.size() : 0;
  return SGNX_OK;
}

int sgnx_tape_steps(const sgnx_solver* s, double* t, double* dt, long cap, long* steps) {
  if (!s || !steps || (cap > 0 && (!t || !dt))) return SGNX_EINVAL;
  const AdjTape* tp = s->tape;
  const long N = tp ? (long)tp->t.size() : 0;
  for (long k = 0; k < std::min(N, cap); ++k) {
    t[k] = tp->t[k];
    dt[k] = tp->dt[k];
  }
  *steps = N;
  return SGNX_OK;
}

int sgnx_tape_replay(sgnx_solver* s) {
  if (!s) return SGNX_EINVAL;
  AdjTape* tp = s->tape;
  if (!tp || tp->t.empty()) return fail(s, SGNX_EINVAL, "tape_replay: no recorded steps");
  if (tp->recording) return fail(s, SGNX_EINVAL, "tape_replay: the tape is still recording");
  int rc;
  for (size_t k = 0; k < tp->t.size(); ++k) {
    if ((rc = bs3_fixed(s, s->d_U, s->d_uNew, tp->t[k], tp->dt[k]))) return rc;
    std::swap(s->d_U, s->d_uNew);
    HIPCHK(s, record_gauges(s, s->stream, tp->t[k] + tp->dt[k]));
  }
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return SGNX_OK;
}

int sgnx_adjoint(sgnx_solver* s, const double* Ubar_final, const double* gauge_bar,
                 double* U0bar, double* theta_bar, int mem) {
  if (!s || !U0bar) return SGNX_EINVAL;
  if (mem != SGNX_MEM_HOST && mem != SGNX_MEM_DEVICE)
    return fail(s, SGNX_EINVAL, "adjoint: bad memory kind");
  AdjTape* tp = nullptr;
  int rc;
  if ((rc = adjoint_begin(s, false, &tp))) return rc;
  const double* gb = nullptr;
  if ((rc = adjoint_inputs(s, tp, Ubar_final, gauge_bar, mem, s->stream, &gb))) return rc;
  const long N = (long)tp->t.size();
  const size_t n = s->n();
  // snapshot segments, last first; inside a segment the states are recomputed
  // with at most c_j held at once (binomial schedule when that is fewer than
  // the segment needs): the slots of sgnx_tape_set_slots, or in planned mode
  // the planned states less the snapshots still held (segment j: j + 1)
  const long S = (long)tp->snap.size();
  auto seg_end = [&](long j) { return j + 1 < S ? tp->snapAt[j + 1] : N; };
  auto free_slots = [&](long j) {
    if (tp->planStates > 0) return std::max(0L, tp->planStates - (j + 1));
    return tp->slots < 0 ? seg_end(j) - tp->snapAt[j] : tp->slots;
  };
  bool pairNeeded = false;
  for (long j = 0; j < S; ++j)
    if (free_slots(j) < seg_end(j) - tp->snapAt[j] - 1) pairNeeded = true;
  if (pairNeeded)
    for (double** p : {&tp->pp[0], &tp->pp[1]})
      if ((rc = alloc_state(s, p, n))) return rc;
  tp->recomputed = 0;
  tp->peakSlots = 0;
  tp->inUse = 0;
  for (long j = S - 1; j >= 0 && N > 0; --j) {
    const long n0 = tp->snapAt[j];
    long n1 = seg_end(j);
    if (j == S - 1 && tp->lastIn && !tp->lastInSnap && N - 1 > n0) {
      // the last step from its held input: where the schedule's first sweep
      // ends (that sweep is the taped run itself)
      if ((rc = reverse_step(s, tp, tp->lastIn, N - 1, gb))) return rc;
      --n1;
    }
    if ((rc = reverse_range(s, tp, tp->snap[j], n0, n1 - n0, free_slots(j), gb))) return rc;
  }
  if ((rc = adjoint_outputs(s, tp, U0bar, theta_bar, mem, s->stream))) return rc;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return SGNX_OK;
}

}  // extern "C"
