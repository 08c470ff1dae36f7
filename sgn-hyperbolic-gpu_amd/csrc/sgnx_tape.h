// =============================================================================
// sgnx_tape.h -- the adjoint tape of a solver handle: the accepted steps and
// state snapshots of recorded runs and the buffers of the reverse sweep,
// shared by the single-grid sweep (sgnx_adjoint_api.hip) and the batched one
// (sgnx_batch.hip).
// =============================================================================
#ifndef SGNX_TAPE_H
#define SGNX_TAPE_H

#include <vector>

#include "sgnx_adjoint.h"
#include "sgnx_handle.h"

struct AdjTape {
  bool recording = false;
  long maxSnap = 0;
  long K = 1;                      // snapshot stride in steps (doubling mode)
  std::vector<double*> snap;       // snap[j] = input state of step snapAt[j]
  std::vector<long> snapAt;        // ascending, snapAt[0] = 0
  // planned mode (planStates > 0): snapshots at the steps of planAt, the
  // first sweep of the binomial schedule for the planned step count; the
  // sweep then holds at most planStates states (snapshots included)
  long planStates = 0;
  std::vector<long> planAt;
  size_t nextPlan = 0;
  // planned mode: the input state of the last accepted step (a snapshot, or
  // held besides them), where the schedule's first sweep ends
  double* lastIn = nullptr;
  bool lastInSnap = false;
  bool pendingSnap = false;        // the step under way is to be snapshot on acceptance
  std::vector<double*> spare;      // allocated, unused snapshot buffers
  std::vector<double> t, dt;       // accepted steps
  std::vector<long> row;           // gauge log row written after step n, or -1
  std::vector<long long> gnodes;   // gauge nodes while recording
  // reverse-sweep buffers
  double *lam = nullptr, *y3b = nullptr, *y2b = nullptr, *coef = nullptr, *thb = nullptr;
  double* coef2 = nullptr;         // second face-coefficient buffer (chained products)
  double* gb = nullptr;            // gauge cotangents staged from the host
  size_t gbCap = 0;
  long slots = -1;                 // states the sweep may hold besides the snapshots (-1: any)
  std::vector<double*> seg;        // recomputation slots (allocated)
  std::vector<double*> freeSeg;    // of those, the unused ones
  double* pp[2] = {nullptr, nullptr};   // ping-pong pair for slot-free recomputation
  long recomputed = 0, peakSlots = 0, inUse = 0;   // of the last sweep
  // stage states kept by the taped run (every step's snapshot held): the
  // stage-2 and stage-3 inputs of step n, written by its accepted attempt
  bool keepStages = false, stagesOff = false;
  std::vector<double*> y2s, y3s;
};


// The tape of s checked for a reverse sweep (finished, snapshots present,
// with storeAll every step's snapshot, the Manning term off) and its sweep
// buffers allocated; errors go to s->err.
int adjoint_begin(sgnx_solver* s, bool storeAll, AdjTape** out);
// On stream st: lam = Ubar_final (or 0), thetaBar = 0, and the gauge-log
// cotangents staged on the device (*gb, null without them); mem as sgnx_rhs.
int adjoint_inputs(sgnx_solver* s, AdjTape* tp, const double* Ubar_final, const double* gauge_bar,
                   int mem, hipStream_t st, const double** gb);
// On stream st: U0bar = lam, theta_bar = thetaBar (when given).
int adjoint_outputs(sgnx_solver* s, AdjTape* tp, double* U0bar, double* theta_bar, int mem,
                    hipStream_t st);
// Chains product a to the step's next product b (b's cotangent is
// b.ca*b.A + b.cb*a.out): a's launch also writes b's face coefficients into
// buf, and b's launch skips the coefficient kernel.
void chain_coef(VjpArgs& a, VjpArgs& b, double* buf);

#endif  // SGNX_TAPE_H
