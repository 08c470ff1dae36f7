// =============================================================================
// sgnx_handle.h -- the solver handle behind the C-ABI and the pieces of the
// adaptive loop shared by the single-grid and the batched drivers.
// =============================================================================
#ifndef SGNX_HANDLE_H
#define SGNX_HANDLE_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/sgnx.h"
#include "sgnx_internal.h"

using namespace sgnx;

struct sgnx_solver {
  int P = 4, Nx = 0, Ny = 0;
  double Lx = 0.0, Ly = 0.0, hx = 0.0, hy = 0.0;
  Geometry g{};
  Phys ph{};
  hipStream_t stream = nullptr;
  int device = 0;
  // device buffers
  OpTables* d_ops = nullptr;
  double *d_Hnod = nullptr, *d_Hx = nullptr, *d_Hy = nullptr;
  double *d_HfX = nullptr, *d_HfY = nullptr, *d_inflow = nullptr;
  double *d_U = nullptr, *d_k1 = nullptr, *d_k2 = nullptr, *d_k3 = nullptr, *d_k4 = nullptr;
  double *d_tmp = nullptr, *d_uNew = nullptr;
  double *d_part = nullptr, *d_out = nullptr;
  double* h_out = nullptr;   // pinned, 2 doubles
  // the adaptive loop's per-attempt result, written by the device straight
  // into host memory: (sum, max, sequence number); h_post is the host view,
  // d_post the device address of the same memory
  double* h_post = nullptr;
  double* d_post = nullptr;
  unsigned long long postSeq = 0;
  double inflowOmega = 0.0;
  bool hasInflow = false;
  std::string err;
  // launch timing: event pairs around RHS launches while enabled, by stage mode
  bool timing = false;
  std::vector<hipEvent_t> events;   // 2 per recorded launch, reused
  std::vector<int> timedMode;       // stage mode of each recorded launch
  long timed = 0;
  long long partCap = 0;            // doubles in d_part
  // gauges recorded by sgnx_integrate after every accepted step
  long long* d_gnodes = nullptr;
  double* d_glog = nullptr;         // gcap rows of nG values
  int nG = 0;
  long gcap = 0, grows = 0;
  std::vector<double> gtimes;       // time of every recorded row
  std::vector<long long> gnodes;    // host copy of the gauge nodes
  // adjoint tape (accepted steps and state snapshots of the recorded runs)
  struct AdjTape* tape = nullptr;
  // this GPU's copy of the static scratch array of sgnx_gauges_noalloc (its
  // device address, looked up once per device by sgnx_create)
  void* scratch = nullptr;

  size_t n() const { return (size_t)NV * (size_t)g.Nd; }
};

// Adjoint tape hooks of the adaptive loops (sgnx_adjoint_api.hip): before the
// first attempt of every step (whether the step's input state is to be a
// snapshot; buffers for kept stage inputs), after every accepted step, once
// d_U and d_uNew are swapped (its time, step size and gauge log row; a due
// snapshot takes over d_uNew, which then holds the step's input, and the
// solver gets a spare buffer in its place), and the release of the tape's
// buffers.
bool tape_recording(const sgnx_solver* s);
// While a tape keeps stage states: the buffers the current step's stage 2 and
// stage 3 launches write their stage inputs to (false: none).
bool tape_stage_outputs(const sgnx_solver* s, double** y2, double** y3);
int tape_step_begin(sgnx_solver* s);
int tape_accept(sgnx_solver* s, double t, double dt, long row);
void tape_free(sgnx_solver* s);

// eta at the gauges of s into its next log row (after an accepted step at t)
inline hipError_t record_gauges(sgnx_solver* s, hipStream_t st, double t) {
  if (s->nG < 1 || s->grows >= s->gcap) return hipSuccess;
  const hipError_t e = launch_gather_eta(s->g, s->d_U, s->d_gnodes, s->nG,
                                         s->d_glog + (size_t)s->grows * s->nG, st);
  if (e == hipSuccess) {
    ++s->grows;
    s->gtimes.push_back(t);
  }
  return e;
}

inline GhostTime ghost_time(const sgnx_solver* s, double t) {
  GhostTime gt{0.0, 0.0};
  if (s->hasInflow) {
    gt.s = std::sin(s->inflowOmega * t);
    gt.c = std::cos(s->inflowOmega * t);
  }
  return gt;
}

// The three row-sweep launches of one BS3 attempt (rk.cpp:99-131) on a
// solver's buffers: stage 2 -> k2; stage 3 -> uNew, and the partial embedded
// sum into k3's buffer; k4 = R(uNew) -> k4 with the error partials in d_part.
struct BS3Launch {
  int mode;
  Stage sa;
  double* dU;
  double t;
};
inline void bs3_launches(const sgnx_solver* s, double t, double dt, double atol, double rtol,
                         BS3Launch L[3]) {
  Stage a{};
  a.u = s->d_U;
  a.k1 = s->d_k1;
  a.a = 0.5 * dt;
  L[0] = {kAxpy, a, s->d_k2, t + 0.5 * dt};
  Stage b{};
  b.u = s->d_U;
  b.k1 = s->d_k2;
  b.a = 0.75 * dt;
  b.kc = s->d_k1;
  b.dt = dt;
  b.uNew = s->d_uNew;
  b.zOut = s->d_k3;
  L[1] = {kCand, b, nullptr, t + 0.75 * dt};
  Stage c{};
  c.u = s->d_uNew;
  c.u0 = s->d_U;
  c.z = s->d_k3;
  c.dt = dt;
  c.atol = atol;
  c.rtol = rtol;
  c.part = s->d_part;
  L[2] = {kErr, c, s->d_k4, t + dt};
  double *y2 = nullptr, *y3 = nullptr;
  if (tape_stage_outputs(s, &y2, &y3)) {
    L[0].sa.yOut = y2;
    L[1].sa.yOut = y3;
  }
}

// The step-size controller of rk.cpp:67-164 for one run: the loop condition
// and end clipping before an attempt (next), the error norm, acceptance and
// PI update after it (finish). The caller launches the stages between the two
// and, when finish returns true, adopts uNew and k4 (FSAL).
struct RkCtl {
  const sgnx_rk_opts* o;
  double alpha, beta;
  double t, tEnd, dt, errPrev = 1.0;
  size_t n;
  double* log;
  long cap, attempt = 0;
  sgnx_rk_stats S{};
  bool done = false;

  RkCtl(const sgnx_rk_opts* o_, double t0, double tEnd_, double dt0, size_t n_, double* log_,
        long cap_)
      : o(o_), t(t0), tEnd(tEnd_), n(n_), log(log_), cap(cap_) {
    const int orderLow = 2;
    const double expo = 1.0 / (orderLow + 1.0);
    alpha = 0.7 * expo;
    beta = 0.4 * expo;
    S.tReached = t0;
    dt = std::min(dt0, o->dtMax);
  }
  bool next() {
    if (done) return false;
    if (!(t < tEnd - 1e-15 * std::max(1.0, std::abs(tEnd)))) return false;
    if (o->max_accepted > 0 && S.accepted >= o->max_accepted) return false;
    if (t + dt > tEnd) dt = tEnd - t;
    return true;
  }
  bool finish(double errSum, double errMax) {
    const double errL2 = std::sqrt(errSum / static_cast<double>(n));
    const double errNorm = std::max(errL2, 1e-300);
    if (log && attempt < cap) {
      log[2 * attempt] = dt;
      log[2 * attempt + 1] = errNorm;
    }
    ++attempt;
    if (!(errNorm == errNorm)) {
      // a non-finite error norm would make the controller's dt NaN and the loop
      // endless; stop with the state at the last accepted step instead
      S.aborted = 1;
      S.dtLast = dt;
      done = true;
      return false;
    }
    bool accepted = false;
    if (errNorm <= 1.0) {
      t += dt;
      ++S.accepted;
      S.tReached = t;
      S.dtLast = dt;
      S.maxNormLast = errMax;
      errPrev = errNorm;
      accepted = true;
    } else {
      ++S.rejected;
    }
    double fac = o->safety * std::pow(errNorm, -alpha) * std::pow(errPrev, beta);
    fac = std::clamp(fac, o->facMin, o->facMax);
    double dtNext = dt * fac;
    dtNext = std::min(dtNext, o->dtMax);
    if (dtNext < o->dtMin) {
      S.aborted = 1;
      S.dtLast = dtNext;
      done = true;
      return accepted;
    }
    dt = dtNext;
    return accepted;
  }
};

#endif  // SGNX_HANDLE_H
