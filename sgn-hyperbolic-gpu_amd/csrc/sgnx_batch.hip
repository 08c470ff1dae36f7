// =============================================================================
// sgnx_batch.hip -- many independent solvers of uneven sizes advanced together.
//
// Every RHS launch of an adaptive step covers all running members: the
// row-sweep workgroups of every member are packed back to back (a per-member
// tile offset table) into one launch per stage kind, followed by one launch
// that folds each member's error partials. Each member keeps its own state
// buffers, time, step size and controller (the same RkCtl as sgnx_integrate)
// and its own tile decomposition, so its results are bit for bit those of
// sgnx_integrate on it alone. The reference's counterpart is a parameter sweep
// run as separate jobs (scripts/shoal_sweep.sbatch).
//
// Members recording an adjoint tape (sgnx_tape_begin) record it here as in
// sgnx_integrate: snapshot copies on the batch stream ahead of a step, the
// accepted steps after it.
//
// A step is: member arguments (pinned host -> device), three stage launches,
// the fold, the (sum, max) pairs (device -> pinned host), one synchronisation;
// optionally captured once into a hipGraph and replayed while the set of
// running members is unchanged (the copies read and write the same pinned
// buffers, which the host refills before each replay).
// =============================================================================

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sgnx_handle.h"
#include "sgnx_tape.h"

struct sgnx_batch {
  std::vector<sgnx_solver*> m;
  int n = 0;
  int device = 0;
  int waves3 = 0x2;                    // bit k: launch kind k built for 3 waves/SIMD
  hipStream_t stream = nullptr;
  // argument tables, [4 launch kinds][n slots], and the tile offsets
  BatchMember* h_mem = nullptr;        // pinned
  BatchMember* d_mem = nullptr;
  long long* h_ts = nullptr;           // pinned, n + 1
  long long* d_ts = nullptr;
  double* h_out = nullptr;             // pinned, 2 n
  double* d_out = nullptr;
  // the captured step and the running set it was captured for
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<int> graphSet;
  bool graphKeep = false;   // the captured step keeps stage inputs
  // the launch shape baked into the captured step: its grid (the members'
  // tiles in all) and each running member's tile height, so that a change of
  // either (sgnx_set_rhs_kernel or sgnx_autotune on a member) forces a capture
  long long graphTotal = 0;
  std::vector<int> graphRows;
  // argument arena of the batched reverse sweep (pinned host, device)
  char* h_adj = nullptr;
  char* d_adj = nullptr;
  size_t adjCap = 0;
  std::string err;
};

namespace {

thread_local std::string g_batch_create_error;

int bfail(sgnx_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  else g_batch_create_error = msg;
  return code;
}

#define BCHK(b, call)                                                                 \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return bfail((b), e_ == hipErrorOutOfMemory ? SGNX_ENOMEM : SGNX_EDEVICE,       \
                   std::string(#call) + ": " + hipGetErrorString(e_));                \
  } while (0)

void drop_graph(sgnx_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graphSet.clear();
  b->graphRows.clear();
  b->graphTotal = 0;
}

BatchMember member_args(const sgnx_solver* s, const double* U, double* dU, const Stage& sa,
                        double t) {
  BatchMember bm{};
  bm.U = U;
  bm.dU = dU;
  bm.g = s->g;
  bm.ph = s->ph;
  bm.gt = ghost_time(s, t);
  bm.sa = sa;
  bm.nStrips = sweep_strips(s->g);
  bm.nTiles = (int)sweep_tiles(s->g);
  return bm;
}

// tile offsets of the members in `set` (slot order) into h_ts; returns the total
long long fill_offsets(sgnx_batch* b, const std::vector<int>& set) {
  long long acc = 0;
  for (size_t i = 0; i < set.size(); ++i) {
    b->h_ts[i] = acc;
    acc += sweep_tiles(b->m[set[i]]->g);
  }
  b->h_ts[set.size()] = acc;
  return acc;
}

// the device work of one attempted step for the members in `set` (slots
// 0..A-1 of kinds 1..3 of h_mem filled by the caller); `keep`: stages 2 and 3
// also write their inputs where a member's stage names a buffer for them
// whether every member's Hx, Hy are (D Hnod) / J bit for bit, so that the
// batched sweep may form them from Hnod (otherwise it reads them, which is
// right for every member)
bool batch_hxr(const sgnx_batch* b) {
  for (const sgnx_solver* s : b->m)
    if (!s->g.hxr) return false;
  return true;
}

int enqueue_step(sgnx_batch* b, const std::vector<int>& set, long long total, bool keep) {
  const int A = (int)set.size();
  const size_t n = (size_t)b->n;
  BCHK(b, hipMemcpyAsync(b->d_mem, b->h_mem, 4 * n * sizeof(BatchMember),
                         hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipMemcpyAsync(b->d_ts, b->h_ts, (n + 1) * sizeof(long long), hipMemcpyHostToDevice,
                         b->stream));
  for (int k = 1; k <= 3; ++k)
    BCHK(b, launch_rhs_stage_batch(b->d_mem + k * n, b->d_ts, A, total, k,
                                   (b->waves3 >> k) & 1, b->stream, keep && k < 3, batch_hxr(b)));
  BCHK(b, launch_finish_batch(b->d_mem + 3 * n, A, b->d_out, b->stream));
  BCHK(b, hipMemcpyAsync(b->h_out, b->d_out, 2 * (size_t)A * sizeof(double),
                         hipMemcpyDeviceToHost, b->stream));
  return SGNX_OK;
}

// tile heights tried by the autotuners (rows of elements per workgroup)
constexpr int kTuneRows[] = {2, 4, 8, 12, 16, 24, 32};

// a step size for the probe launches: any finite value times the same work
constexpr double kProbeDt = 1e-9;

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

}  // namespace

extern "C" {

const char* sgnx_batch_last_error(const sgnx_batch* b) {
  return b ? b->err.c_str() : g_batch_create_error.c_str();
}

int sgnx_batch_create(sgnx_solver* const* members, int n, sgnx_batch** out) {
  if (!out) return SGNX_EINVAL;
  *out = nullptr;
  if (!members || n < 1) return bfail(nullptr, SGNX_EINVAL, "batch: no members");
  int dev = -1;
  for (int i = 0; i < n; ++i) {
    const sgnx_solver* s = members[i];
    if (!s) return bfail(nullptr, SGNX_EINVAL, "batch: null member");
    if (s->g.variant != 1)
      return bfail(nullptr, SGNX_EINVAL,
                   s->ph.nManning != 0.0
                       ? "batch: members with Manning friction are not supported (no row sweep)"
                       : "batch: members must use the row-sweep kernel");
    if (dev < 0) dev = s->device;
    if (s->device != dev) return bfail(nullptr, SGNX_EINVAL, "batch: members on different devices");
    for (int j = 0; j < i; ++j)
      if (members[j] == s) return bfail(nullptr, SGNX_EINVAL, "batch: a member appears twice");
  }
  sgnx_batch* b = new (std::nothrow) sgnx_batch();
  if (!b) return bfail(nullptr, SGNX_ENOMEM, "batch: out of host memory");
  b->m.assign(members, members + n);
  b->n = n;
  b->device = dev;
  int rc = SGNX_OK;
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == SGNX_OK) {
      rc = (e == hipErrorOutOfMemory) ? SGNX_ENOMEM : SGNX_EDEVICE;
      g_batch_create_error = std::string(what) + ": " + hipGetErrorString(e);
    }
  };
  chk(hipSetDevice(dev), "hipSetDevice");
  chk(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
  chk(hipHostMalloc((void**)&b->h_mem, 4 * (size_t)n * sizeof(BatchMember)), "hipHostMalloc");
  chk(hipMalloc((void**)&b->d_mem, 4 * (size_t)n * sizeof(BatchMember)), "hipMalloc");
  chk(hipHostMalloc((void**)&b->h_ts, ((size_t)n + 1) * sizeof(long long)), "hipHostMalloc");
  chk(hipMalloc((void**)&b->d_ts, ((size_t)n + 1) * sizeof(long long)), "hipMalloc");
  chk(hipHostMalloc((void**)&b->h_out, 2 * (size_t)n * sizeof(double)), "hipHostMalloc");
  chk(hipMalloc((void**)&b->d_out, 2 * (size_t)n * sizeof(double)), "hipMalloc");
  if (rc != SGNX_OK) {
    sgnx_batch_destroy(b);
    return rc;
  }
  if (const char* w = std::getenv("SGNX_SWEEP_WAVES3")) b->waves3 = waves_mask_from_text(w);
  else b->waves3 = members[0]->g.waves3;
  *out = b;
  return SGNX_OK;
}

int sgnx_batch_destroy(sgnx_batch* b) {
  if (!b) return SGNX_EINVAL;
  drop_graph(b);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->h_mem) (void)hipHostFree(b->h_mem);
  if (b->h_ts) (void)hipHostFree(b->h_ts);
  if (b->h_out) (void)hipHostFree(b->h_out);
  if (b->h_adj) (void)hipHostFree(b->h_adj);
  for (void* p : {(void*)b->d_mem, (void*)b->d_ts, (void*)b->d_out, (void*)b->d_adj})
    if (p) (void)hipFree(p);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return SGNX_OK;
}

int sgnx_batch_set_waves(sgnx_batch* b, int waves3) {
  if (!b || waves3 < 0 || waves3 > 0xf) return SGNX_EINVAL;
  b->waves3 = waves3;
  drop_graph(b);
  return SGNX_OK;
}

int sgnx_batch_rhs(sgnx_batch* b, const double* const* U, double* const* dU, const double* t) {
  if (!b || !U || !dU || !t) return SGNX_EINVAL;
  const int n = b->n;
  BCHK(b, hipSetDevice(b->device));
  for (int i = 0; i < n; ++i) {
    if (!U[i] || !dU[i] || U[i] == dU[i]) return bfail(b, SGNX_EINVAL, "batch rhs: bad buffers");
    BCHK(b, hipStreamSynchronize(b->m[i]->stream));
  }
  std::vector<int> all(n);
  for (int i = 0; i < n; ++i) {
    all[i] = i;
    Stage sa{};
    sa.u = U[i];
    b->h_mem[i] = member_args(b->m[i], U[i], dU[i], sa, t[i]);
  }
  const long long total = fill_offsets(b, all);
  BCHK(b, hipMemcpyAsync(b->d_mem, b->h_mem, (size_t)n * sizeof(BatchMember),
                         hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipMemcpyAsync(b->d_ts, b->h_ts, ((size_t)n + 1) * sizeof(long long),
                         hipMemcpyHostToDevice, b->stream));
  BCHK(b, launch_rhs_stage_batch(b->d_mem, b->d_ts, n, total, kPlain, b->waves3 & 1, b->stream,
                                 false, batch_hxr(b)));
  BCHK(b, hipStreamSynchronize(b->stream));
  return SGNX_OK;
}

int sgnx_batch_integrate(sgnx_batch* b, const double* t0, const double* tEnd, const double* dt0,
                         const sgnx_rk_opts* opts, sgnx_rk_stats* stats, int flags) {
  if (!b || !t0 || !tEnd || !dt0 || !opts || !stats) return SGNX_EINVAL;
  const int n = b->n;
  const bool useGraph = (flags & SGNX_BATCH_GRAPH) != 0;
  BCHK(b, hipSetDevice(b->device));
  for (int i = 0; i < n; ++i) BCHK(b, hipStreamSynchronize(b->m[i]->stream));
  std::vector<RkCtl> ctl;
  ctl.reserve(n);
  for (int i = 0; i < n; ++i) ctl.emplace_back(&opts[i], t0[i], tEnd[i], dt0[i], b->m[i]->n(),
                                               nullptr, 0);
  // FSAL heads: k1 = R(U) of every member, one launch
  {
    std::vector<int> all(n);
    for (int i = 0; i < n; ++i) {
      all[i] = i;
      sgnx_solver* s = b->m[i];
      Stage sa{};
      sa.u = s->d_U;
      b->h_mem[i] = member_args(s, s->d_U, s->d_k1, sa, ctl[i].t);
      ++ctl[i].S.rhsEvals;
    }
    const long long total = fill_offsets(b, all);
    BCHK(b, hipMemcpyAsync(b->d_mem, b->h_mem, (size_t)n * sizeof(BatchMember),
                           hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipMemcpyAsync(b->d_ts, b->h_ts, ((size_t)n + 1) * sizeof(long long),
                           hipMemcpyHostToDevice, b->stream));
    BCHK(b, launch_rhs_stage_batch(b->d_mem, b->d_ts, n, total, kPlain, b->waves3 & 1,
                                   b->stream, false, batch_hxr(b)));
    // the copies above read the pinned tables asynchronously; the steps below
    // refill them
    BCHK(b, hipStreamSynchronize(b->stream));
  }
  std::vector<int> set;
  int rc = SGNX_OK;
  for (;;) {
    set.clear();
    for (int i = 0; i < n; ++i)
      if (ctl[i].next()) set.push_back(i);
    if (set.empty()) break;
    const int A = (int)set.size();
    // members recording a tape: whether the step's input state is due as a
    // snapshot, and the buffers its stage launches keep their inputs in
    for (int slot = 0; slot < A; ++slot) {
      sgnx_solver* s = b->m[set[slot]];
      if (tape_recording(s) && tape_step_begin(s))
        return bfail(b, SGNX_EDEVICE, "batch: tape snapshot: " + s->err);
    }
    bool keep = false;
    for (int slot = 0; slot < A; ++slot) {
      const int i = set[slot];
      sgnx_solver* s = b->m[i];
      BS3Launch L[3];
      bs3_launches(s, ctl[i].t, ctl[i].dt, opts[i].atol, opts[i].rtol, L);
      keep = keep || L[0].sa.yOut != nullptr;
      for (int k = 0; k < 3; ++k)
        b->h_mem[(size_t)(k + 1) * n + slot] = member_args(s, L[k].sa.u, L[k].dU, L[k].sa, L[k].t);
    }
    const long long total = fill_offsets(b, set);
    if (useGraph) {
      std::vector<int> rowsNow(A);
      for (int slot = 0; slot < A; ++slot) rowsNow[slot] = b->m[set[slot]]->g.rows;
      if (!b->exec || b->graphSet != set || b->graphKeep != keep || b->graphTotal != total ||
          b->graphRows != rowsNow) {
        drop_graph(b);
        BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_step(b, set, total, keep);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(b->stream, &g);
        if (rc) {
          if (g) (void)hipGraphDestroy(g);
          return rc;
        }
        BCHK(b, ec);
        b->graph = g;
        BCHK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graphSet = set;
        b->graphKeep = keep;
        b->graphTotal = total;
        b->graphRows = rowsNow;
      }
      BCHK(b, hipGraphLaunch(b->exec, b->stream));
    } else {
      if ((rc = enqueue_step(b, set, total, keep))) return rc;
    }
    BCHK(b, hipStreamSynchronize(b->stream));
    for (int slot = 0; slot < A; ++slot) {
      const int i = set[slot];
      sgnx_solver* s = b->m[i];
      ctl[i].S.rhsEvals += 3;
      const double tStep = ctl[i].t, dtStep = ctl[i].dt;
      if (ctl[i].finish(b->h_out[2 * slot], b->h_out[2 * slot + 1])) {
        std::swap(s->d_U, s->d_uNew);
        std::swap(s->d_k1, s->d_k4);
        const long rows = s->grows;
        BCHK(b, record_gauges(s, b->stream, ctl[i].t));
        if (tape_recording(s) && tape_accept(s, tStep, dtStep, s->grows > rows ? rows : -1))
          return bfail(b, SGNX_EDEVICE, "batch: tape: " + s->err);
      }
    }
  }
  BCHK(b, hipStreamSynchronize(b->stream));
  for (int i = 0; i < n; ++i) stats[i] = ctl[i].S;
  return SGNX_OK;
}

// ---- autotuning (no counterpart in the reference) --------------------------

int sgnx_autotune(sgnx_solver* s, int reps, int* rows_out, int* waves3_out) {
  if (!s || reps < 1) return SGNX_EINVAL;
  if (s->ph.nManning != 0.0) {
    s->err = "autotune: the row sweep has no Manning friction term";
    return SGNX_EUNSUPPORTED;
  }
  if (hipSetDevice(s->device) != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess)
    return SGNX_EDEVICE;
  EventPair ev;
  if (hipEventCreate(&ev.a) != hipSuccess || hipEventCreate(&ev.b) != hipSuccess)
    return SGNX_EDEVICE;
  const Geometry keep = s->g;
  s->g.variant = 1;
  BS3Launch L[3];
  // ms of `reps` launches of kind k (0: R(U) into k2's buffer; 1..3: a BS3
  // stage at the probe step) with the kind built for three waves or two
  auto timed = [&](int k, bool three, float* ms) -> hipError_t {
    s->g.waves3 = three ? (1 << k) : 0;
    Stage sa{};
    sa.u = s->d_U;
    const Stage& st = (k == 0) ? sa : L[k - 1].sa;
    double* dU = (k == 0) ? s->d_k2 : L[k - 1].dU;
    const GhostTime gt = ghost_time(s, 0.0);
    hipError_t e = launch_rhs_stage(s->g, s->ph, k, st, dU, gt, s->stream);   // warm
    if (e != hipSuccess) return e;
    if ((e = hipEventRecord(ev.a, s->stream)) != hipSuccess) return e;
    for (int r = 0; r < reps; ++r)
      if ((e = launch_rhs_stage(s->g, s->ph, k, st, dU, gt, s->stream)) != hipSuccess) return e;
    if ((e = hipEventRecord(ev.b, s->stream)) != hipSuccess) return e;
    if ((e = hipEventSynchronize(ev.b)) != hipSuccess) return e;
    return hipEventElapsedTime(ms, ev.a, ev.b);
  };
  double best = 1e300;
  int bestRows = keep.rows, bestMask = keep.waves3;
  for (int rows : kTuneRows) {
    if (rows > s->g.Ny && rows != kTuneRows[0]) continue;
    s->g.rows = rows;
    bs3_launches(s, 0.0, kProbeDt, 1.0, 0.0, L);
    double tot = 0.0;
    int mask = 0;
    for (int k = 1; k <= 3; ++k) {
      float t2 = 0.0f, t3 = 0.0f;
      if (timed(k, false, &t2) != hipSuccess || timed(k, true, &t3) != hipSuccess) {
        s->g = keep;
        s->err = "autotune: probe launch failed";
        return SGNX_EDEVICE;
      }
      tot += std::min(t2, t3);
      if (t3 < t2) mask |= 1 << k;
    }
    if (tot < best) {
      best = tot;
      bestRows = rows;
      bestMask = mask;
    }
  }
  s->g.rows = bestRows;
  float p2 = 0.0f, p3 = 0.0f;
  if (timed(0, false, &p2) != hipSuccess || timed(0, true, &p3) != hipSuccess) {
    s->g = keep;
    s->err = "autotune: probe launch failed";
    return SGNX_EDEVICE;
  }
  s->g.waves3 = bestMask | (p3 < p2 ? 1 : 0);
  if (rows_out) *rows_out = s->g.rows;
  if (waves3_out) *waves3_out = s->g.waves3;
  return SGNX_OK;
}

int sgnx_batch_autotune(sgnx_batch* b, int reps, int* rows_out, int* waves3_out) {
  if (!b || reps < 1) return SGNX_EINVAL;
  const int n = b->n;
  BCHK(b, hipSetDevice(b->device));
  for (int i = 0; i < n; ++i) BCHK(b, hipStreamSynchronize(b->m[i]->stream));
  drop_graph(b);
  EventPair ev;
  BCHK(b, hipEventCreate(&ev.a));
  BCHK(b, hipEventCreate(&ev.b));
  std::vector<int> all(n);
  for (int i = 0; i < n; ++i) all[i] = i;
  int maxNy = 1;
  for (sgnx_solver* s : b->m) maxNy = std::max(maxNy, s->g.Ny);
  const size_t nn = (size_t)n;
  // fill the tables of every kind for the members' current tile height
  auto fill = [&]() -> long long {
    for (int i = 0; i < n; ++i) {
      sgnx_solver* s = b->m[i];
      BS3Launch L[3];
      bs3_launches(s, 0.0, kProbeDt, 1.0, 0.0, L);
      Stage sa{};
      sa.u = s->d_U;
      b->h_mem[i] = member_args(s, s->d_U, s->d_k2, sa, 0.0);
      for (int k = 0; k < 3; ++k)
        b->h_mem[(size_t)(k + 1) * nn + i] = member_args(s, L[k].sa.u, L[k].dU, L[k].sa, L[k].t);
    }
    return fill_offsets(b, all);
  };
  auto timed = [&](int k, bool three, long long total, float* ms) -> hipError_t {
    hipError_t e = launch_rhs_stage_batch(b->d_mem + k * nn, b->d_ts, n, total, k, three,
                                          b->stream, false, batch_hxr(b));
    if (e != hipSuccess) return e;
    if ((e = hipEventRecord(ev.a, b->stream)) != hipSuccess) return e;
    for (int r = 0; r < reps; ++r)
      if ((e = launch_rhs_stage_batch(b->d_mem + k * nn, b->d_ts, n, total, k, three,
                                      b->stream, false, batch_hxr(b))) != hipSuccess)
        return e;
    if ((e = hipEventRecord(ev.b, b->stream)) != hipSuccess) return e;
    if ((e = hipEventSynchronize(ev.b)) != hipSuccess) return e;
    return hipEventElapsedTime(ms, ev.a, ev.b);
  };
  std::vector<int> keepRows(n);
  for (int i = 0; i < n; ++i) keepRows[i] = b->m[i]->g.rows;
  double best = 1e300;
  int bestRows = keepRows[0], bestMask = b->waves3;
  for (int rows : kTuneRows) {
    if (rows > maxNy && rows != kTuneRows[0]) continue;
    for (sgnx_solver* s : b->m) s->g.rows = rows;
    const long long total = fill();
    BCHK(b, hipMemcpyAsync(b->d_mem, b->h_mem, 4 * nn * sizeof(BatchMember),
                           hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipMemcpyAsync(b->d_ts, b->h_ts, (nn + 1) * sizeof(long long),
                           hipMemcpyHostToDevice, b->stream));
    double tot = 0.0;
    int mask = 0;
    for (int k = 1; k <= 3; ++k) {
      float t2 = 0.0f, t3 = 0.0f;
      BCHK(b, timed(k, false, total, &t2));
      BCHK(b, timed(k, true, total, &t3));
      tot += std::min(t2, t3);
      if (t3 < t2) mask |= 1 << k;
    }
    BCHK(b, hipStreamSynchronize(b->stream));
    if (tot < best) {
      best = tot;
      bestRows = rows;
      bestMask = mask;
    }
  }
  for (sgnx_solver* s : b->m) s->g.rows = bestRows;
  const long long total = fill();
  BCHK(b, hipMemcpyAsync(b->d_mem, b->h_mem, 4 * nn * sizeof(BatchMember),
                         hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipMemcpyAsync(b->d_ts, b->h_ts, (nn + 1) * sizeof(long long), hipMemcpyHostToDevice,
                         b->stream));
  float p2 = 0.0f, p3 = 0.0f;
  BCHK(b, timed(0, false, total, &p2));
  BCHK(b, timed(0, true, total, &p3));
  BCHK(b, hipStreamSynchronize(b->stream));
  b->waves3 = bestMask | (p3 < p2 ? 1 : 0);
  for (sgnx_solver* s : b->m) s->g.waves3 = b->waves3;
  if (rows_out) *rows_out = bestRows;
  if (waves3_out) *waves3_out = b->waves3;
  return SGNX_OK;
}

// The reverse sweeps of every member's tape in lockstep, aligned at the last
// step: iteration k reverses step N_m - 1 - k of each member that has it, with
// one launch per kind over all of them (the gauge cotangents, the stage
// recomputation of members whose tape keeps no stage inputs: R(u), then the
// stage-2 launch that writes y2 and forms y3; the three products), each
// member on its own buffers. The arithmetic is that of sgnx_adjoint on each
// member alone, so the cotangents are bit for bit the same. Every member's
// tape must hold every step's snapshot. When every member alone gives each
// compute unit a full round of product workgroups (6 of 128 work-items), the
// members' own sweeps run back to back instead: the products are bound by
// their traffic, so the lockstep's launches gain nothing there, and its
// per-step argument tables and member lookups made it slower (same bits).
int sgnx_batch_adjoint(sgnx_batch* b, const double* const* Ubar_final,
                       const double* const* gauge_bar, double* const* U0bar,
                       double* const* theta_bar, int mem) {
  if (!b || !U0bar) return SGNX_EINVAL;
  if (mem != SGNX_MEM_HOST && mem != SGNX_MEM_DEVICE)
    return bfail(b, SGNX_EINVAL, "batch adjoint: bad memory kind");
  const int n = b->n;
  BCHK(b, hipSetDevice(b->device));
  int cus = 0;
  BCHK(b, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
  bool each_fills = n > 0;
  for (int i = 0; i < n; ++i) each_fills = each_fills && vjp_tiles(b->m[i]->g) >= 6LL * cus;
  if (each_fills) {
    for (int i = 0; i < n; ++i) {
      sgnx_solver* s = b->m[i];
      if (!U0bar[i]) return bfail(b, SGNX_EINVAL, "batch adjoint: U0bar missing for a member");
      AdjTape* t = nullptr;
      int rc = adjoint_begin(s, true, &t);
      if (!rc)
        rc = sgnx_adjoint(s, Ubar_final ? Ubar_final[i] : nullptr,
                          gauge_bar ? gauge_bar[i] : nullptr, U0bar[i],
                          theta_bar ? theta_bar[i] : nullptr, mem);
      if (rc) return bfail(b, rc, "batch adjoint, member " + std::to_string(i) + ": " + s->err);
    }
    return SGNX_OK;
  }
  std::vector<AdjTape*> tp(n, nullptr);
  std::vector<const double*> gb(n, nullptr);
  long Nmax = 0;
  for (int i = 0; i < n; ++i) {
    sgnx_solver* s = b->m[i];
    if (!U0bar[i]) return bfail(b, SGNX_EINVAL, "batch adjoint: U0bar missing for a member");
    BCHK(b, hipStreamSynchronize(s->stream));
    int rc = adjoint_begin(s, true, &tp[i]);
    if (!rc)
      rc = adjoint_inputs(s, tp[i], Ubar_final ? Ubar_final[i] : nullptr,
                          gauge_bar ? gauge_bar[i] : nullptr, mem, b->stream, &gb[i]);
    if (rc) return bfail(b, rc, "batch adjoint, member " + std::to_string(i) + ": " + s->err);
    Nmax = std::max(Nmax, (long)tp[i]->t.size());
  }
  // argument arena: gauges, two RHS tables (R(u); the stage-2 launch keeping
  // y2 and forming y3) + offsets, three products + offsets
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t nn = (size_t)n;
  const size_t oG = 0, oR1 = up(oG + nn * sizeof(GaugeMember));
  const size_t oR2 = up(oR1 + nn * sizeof(BatchMember)), oRT = up(oR2 + nn * sizeof(BatchMember));
  const size_t oV3 = up(oRT + (nn + 1) * sizeof(long long));
  const size_t oV2 = up(oV3 + nn * sizeof(VjpMember)), oV1 = up(oV2 + nn * sizeof(VjpMember));
  const size_t oVT = up(oV1 + nn * sizeof(VjpMember));
  const size_t total = up(oVT + (nn + 1) * sizeof(long long));
  if (b->adjCap < total) {
    if (b->h_adj) (void)hipHostFree(b->h_adj);
    if (b->d_adj) (void)hipFree(b->d_adj);
    b->h_adj = nullptr;
    b->d_adj = nullptr;
    b->adjCap = 0;
    BCHK(b, hipHostMalloc((void**)&b->h_adj, total, hipHostMallocDefault));
    BCHK(b, hipMalloc((void**)&b->d_adj, total));
    b->adjCap = total;
  }
  auto* hG = reinterpret_cast<GaugeMember*>(b->h_adj + oG);
  auto* hR1 = reinterpret_cast<BatchMember*>(b->h_adj + oR1);
  auto* hR2 = reinterpret_cast<BatchMember*>(b->h_adj + oR2);
  auto* hRT = reinterpret_cast<long long*>(b->h_adj + oRT);
  auto* hV3 = reinterpret_cast<VjpMember*>(b->h_adj + oV3);
  auto* hV2 = reinterpret_cast<VjpMember*>(b->h_adj + oV2);
  auto* hV1 = reinterpret_cast<VjpMember*>(b->h_adj + oV1);
  auto* hVT = reinterpret_cast<long long*>(b->h_adj + oVT);
  auto dptr = [&](size_t off) { return b->d_adj + off; };
  for (long k = 0; k < Nmax; ++k) {
    int A = 0, Rn = 0;   // members in this step; of those, the ones that recompute stages
    long long rAcc = 0, vAcc = 0;
    for (int i = 0; i < n; ++i) {
      AdjTape* T = tp[i];
      const long m = (long)T->t.size() - 1 - k;
      if (m < 0) continue;
      sgnx_solver* s = b->m[i];
      const double t = T->t[m], dt = T->dt[m];
      const double* u = T->snap[m];
      GaugeMember gq{};
      if (gb[i] && T->row[m] >= 0 && T->row[m] < s->grows) {
        gq.Gb = gb[i] + (size_t)T->row[m] * s->nG;
        gq.nodes = s->d_gnodes;
        gq.n = s->nG;
        gq.lam = T->lam;
        gq.thetaBar = T->thb;
      }
      hG[A] = gq;
      // stage inputs y2, y3: kept by the taped run, or recomputed here
      const bool kept = m < (long)T->y2s.size();
      const double* y2 = kept ? T->y2s[m] : s->d_tmp;
      const double* y3 = kept ? T->y3s[m] : s->d_k3;
      if (!kept) {
        Stage s1{};
        s1.u = u;
        hR1[Rn] = member_args(s, u, s->d_k1, s1, t);
        // the forward's stage-2 launch, keeping its input and forming the
        // next one: y2, k2 = R(y2) and y3 = u + 3dt/4 k2
        Stage s2{};
        s2.u = u;
        s2.k1 = s->d_k1;
        s2.a = 0.5 * dt;
        s2.yOut = s->d_tmp;
        s2.y3Out = s->d_k3;
        s2.a3 = 0.75 * dt;
        hR2[Rn] = member_args(s, u, s->d_k2, s2, t + 0.5 * dt);
        hRT[Rn] = rAcc;
        rAcc += sweep_tiles(s->g);
        ++Rn;
      }
      // uNew = u + dt ((2/9 k1 + 1/3 k2) + 4/9 k3)   (rk.cpp:106-108)
      VjpArgs a3{};
      a3.U = y3;
      a3.A = T->lam;
      a3.ca = dt * (4.0 / 9.0);
      a3.out = T->y3b;
      a3.thetaBar = T->thb;
      a3.coef = T->coef;
      VjpArgs a2 = a3;
      a2.U = y2;
      a2.ca = dt * (1.0 / 3.0);
      a2.B = T->y3b;
      a2.cb = 0.75 * dt;
      a2.out = T->y2b;
      VjpArgs a1 = a3;
      a1.U = u;
      a1.ca = dt * (2.0 / 9.0);
      a1.B = T->y2b;
      a1.cb = 0.5 * dt;
      a1.out = T->lam;
      a1.acc0 = T->lam;
      a1.acc1 = T->y3b;
      a1.acc2 = T->y2b;
      chain_coef(a3, a2, T->coef2);
      chain_coef(a2, a1, T->coef);
      hV3[A] = VjpMember{a3, s->g, s->ph, ghost_time(s, t + 0.75 * dt)};
      hV2[A] = VjpMember{a2, s->g, s->ph, ghost_time(s, t + 0.5 * dt)};
      hV1[A] = VjpMember{a1, s->g, s->ph, ghost_time(s, t)};
      hVT[A] = vAcc;
      vAcc += vjp_tiles(s->g);
      ++A;
    }
    hRT[Rn] = rAcc;
    hVT[A] = vAcc;
    BCHK(b, hipMemcpyAsync(b->d_adj, b->h_adj, total, hipMemcpyHostToDevice, b->stream));
    const auto* dR1 = reinterpret_cast<const BatchMember*>(dptr(oR1));
    const auto* dR2 = reinterpret_cast<const BatchMember*>(dptr(oR2));
    const auto* dRT = reinterpret_cast<const long long*>(dptr(oRT));
    const auto* dVT = reinterpret_cast<const long long*>(dptr(oVT));
    BCHK(b, launch_gauge_bar_batch(reinterpret_cast<const GaugeMember*>(dptr(oG)), A, b->stream));
    if (Rn > 0) {
      BCHK(b, launch_rhs_stage_batch(dR1, dRT, Rn, rAcc, kPlain, b->waves3 & 1, b->stream, false,
                                     batch_hxr(b)));
      BCHK(b, launch_rhs_stage_batch(dR2, dRT, Rn, rAcc, kAxpy, (b->waves3 >> 1) & 1, b->stream,
                                     true, batch_hxr(b)));
    }
    // the products chained: the first launches the coefficient kernel, each
    // later one finds its coefficients written by the one before
    for (size_t o : {oV3, oV2, oV1})
      BCHK(b, launch_rhs_vjp_batch(reinterpret_cast<const VjpMember*>(dptr(o)), dVT, A, vAcc,
                                   o != oV3, b->stream));
    // the arena is refilled for the next iteration
    BCHK(b, hipStreamSynchronize(b->stream));
  }
  for (int i = 0; i < n; ++i) {
    int rc = adjoint_outputs(b->m[i], tp[i], U0bar[i], theta_bar ? theta_bar[i] : nullptr, mem,
                             b->stream);
    if (rc) return bfail(b, rc, "batch adjoint: " + b->m[i]->err);
  }
  BCHK(b, hipStreamSynchronize(b->stream));
  return SGNX_OK;
}

}  // extern "C"
