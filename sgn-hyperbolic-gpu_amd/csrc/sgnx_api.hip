// =============================================================================
// sgnx_api.hip -- the C-ABI of libsgnx: solver handles, device memory, and the
// device-resident adaptive Bogacki-Shampine 3(2) loop.
//
// The loop is the arithmetic of src/solver/rk.cpp:67-164 on device buffers:
// stage inputs, the third-order candidate and the scaled embedded error are
// formed by kernels with the CPU's association (not the reference GPU
// kComb3's), the error partials are folded on the device in a fixed order, and
// only the (sum, max) pair crosses to the host, where the PI controller runs
// with the C library's pow as in rk.cpp:149-161. The RHS and stage arithmetic
// are bitwise equal to the CPU's; the error sum differs from the CPU's
// sequential sum only in rounding, so accept/reject decisions agree unless an
// error norm lies within rounding of 1.
// =============================================================================

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "sgnx_handle.h"

// Scratch of sgnx_gauges_noalloc: the node list (SGNX_GAUGES_NOALLOC_MAX
// 64-bit indices) followed by the gathered values (as many doubles), 64 KiB.
// Every GPU the library runs on holds its own copy until the process ends, and
// every call on that GPU shares it.
static __device__ __align__(256) unsigned char
    scratch[SGNX_GAUGES_NOALLOC_MAX * (sizeof(long long) + sizeof(double))];



namespace {

thread_local std::string g_create_error;

// Device address of `scratch` on each GPU, looked up once per device with that
// device current. In host code the array's name is only a stand-in for the
// symbol: kernels and copies must be given this address.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

bool hxr_allowed() {
  const char* v = std::getenv("SGNX_SWEEP_HXR");
  return !v || std::atoi(v) != 0;
}

hipError_t scratch_address(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

int fail(sgnx_solver* s, int code, const std::string& msg) {
  if (s) s->err = msg;
  else g_create_error = msg;
  return code;
}

#define HIPCHK(s, call)                                                               \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail((s), e_ == hipErrorOutOfMemory ? SGNX_ENOMEM : SGNX_EDEVICE,        \
                  std::string(#call) + ": " + hipGetErrorString(e_));                 \
  } while (0)

void free_all(sgnx_solver* s) {
  for (void* p : {(void*)s->d_ops, (void*)s->d_Hnod, (void*)s->d_Hx, (void*)s->d_Hy,
                  (void*)s->d_HfX, (void*)s->d_HfY, (void*)s->d_inflow, (void*)s->d_U,
                  (void*)s->d_k1, (void*)s->d_k2, (void*)s->d_k3, (void*)s->d_k4,
                  (void*)s->d_tmp, (void*)s->d_uNew, (void*)s->d_part, (void*)s->d_out,
                  (void*)s->d_gnodes, (void*)s->d_glog})
    if (p) (void)hipFree(p);
  if (s->h_out) (void)hipHostFree(s->h_out);
  if (s->h_post) (void)hipHostFree(s->h_post);
  for (hipEvent_t e : s->events) (void)hipEventDestroy(e);
  s->events.clear();
  tape_free(s);
}


// One RHS launch: plain input (mode kPlain, either kernel variant) or a fused
// stage input (row sweep only), bracketed by timing events when enabled.
int rhs_launch(sgnx_solver* s, int mode, const Stage& sa, double* dU, double t) {
  const GhostTime gt = ghost_time(s, t);
  if (s->timing) {
    const size_t need = 2 * (size_t)(s->timed + 1);
    while (s->events.size() < need) {
      hipEvent_t e;
      HIPCHK(s, hipEventCreate(&e));
      s->events.push_back(e);
    }
    if (s->timedMode.size() < (size_t)(s->timed + 1)) s->timedMode.resize(s->timed + 1);
    s->timedMode[s->timed] = mode;
    HIPCHK(s, hipEventRecord(s->events[2 * s->timed], s->stream));
  }
  if (mode == kPlain) HIPCHK(s, launch_rhs(s->g, s->ph, sa.u, dU, gt, s->stream));
  else HIPCHK(s, launch_rhs_stage(s->g, s->ph, mode, sa, dU, gt, s->stream));
  if (s->timing) {
    HIPCHK(s, hipEventRecord(s->events[2 * s->timed + 1], s->stream));
    ++s->timed;
  }
  return SGNX_OK;
}

int rhs_dev(sgnx_solver* s, const double* U, double* dU, double t) {
  Stage sa{};
  sa.u = U;
  return rhs_launch(s, kPlain, sa, dU, t);
}

int sync(sgnx_solver* s) {
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return SGNX_OK;
}

}  // namespace

extern "C" {

const char* sgnx_last_error(const sgnx_solver* s) {
  return s ? s->err.c_str() : g_create_error.c_str();
}

int sgnx_create(const sgnx_desc* d, sgnx_solver** out) {
  if (!d || !out) return fail(nullptr, SGNX_EINVAL, "null descriptor or output");
  *out = nullptr;
  if (d->P != 4) return fail(nullptr, SGNX_EUNSUPPORTED, "the kernels implement P = 4 only");
  if (d->Nx < 2 || d->Ny < 2) return fail(nullptr, SGNX_EINVAL, "SGN2D: Nx,Ny >= 2");
  if (!(d->Lx > 0.0) || !(d->Ly > 0.0)) return fail(nullptr, SGNX_EINVAL, "Lx, Ly must be > 0");
  if (!d->D || !d->lL || !d->lR || !d->gLp || !d->gRp || !d->qw)
    return fail(nullptr, SGNX_EINVAL, "operator tables missing");
  for (int k = 0; k < NP; ++k) {
    const double eL = (k == 0) ? 1.0 : 0.0, eR = (k == NP - 1) ? 1.0 : 0.0;
    if (d->lL[k] != eL || d->lR[k] != eR || std::signbit(d->lL[k]) || std::signbit(d->lR[k]))
      return fail(nullptr, SGNX_EUNSUPPORTED,
                  "face rows are not the Gauss-Lobatto unit rows (only GLL points are supported)");
  }
  for (int k = 0; k < 4; ++k)
    if (d->bc[k] < 0 || d->bc[k] > 2) return fail(nullptr, SGNX_EINVAL, "bad side code");
  if ((d->bc[0] == SGNX_BC_PERIODIC) != (d->bc[1] == SGNX_BC_PERIODIC) ||
      (d->bc[2] == SGNX_BC_PERIODIC) != (d->bc[3] == SGNX_BC_PERIODIC))
    return fail(nullptr, SGNX_EINVAL, "periodic sides must come in pairs");
  if (d->bc[1] == SGNX_BC_INFLOW || d->bc[2] == SGNX_BC_INFLOW || d->bc[3] == SGNX_BC_INFLOW)
    return fail(nullptr, SGNX_EUNSUPPORTED, "inflow is supported on the left side only");
  if (d->bc[0] == SGNX_BC_INFLOW && !d->inflow)
    return fail(nullptr, SGNX_EINVAL, "inflow side without an inflow table");
  if (d->Hnod && (!d->HfaceX || !d->HfaceY))
    return fail(nullptr, SGNX_EINVAL, "face still depths required with a nodal still depth");
  // node offsets within one variable's array are 32-bit byte offsets in the kernels
  if ((long long)d->Nx * d->Ny * NPP > (1LL << 28))
    return fail(nullptr, SGNX_EUNSUPPORTED, "more than 2^28 nodes per variable");

  auto* s = new (std::nothrow) sgnx_solver();
  if (!s) return fail(nullptr, SGNX_ENOMEM, "host allocation failed");
  s->P = d->P; s->Nx = d->Nx; s->Ny = d->Ny; s->Lx = d->Lx; s->Ly = d->Ly;
  s->hx = d->Lx / d->Nx;
  s->hy = d->Ly / d->Ny;
  Geometry& g = s->g;
  g.Nx = d->Nx; g.Ny = d->Ny;
  g.Nd = (long long)d->Nx * d->Ny * NPP;
  g.Jx = 0.5 * s->hx;                    // mesh2d.h:49-50
  g.Jy = 0.5 * s->hy;
  g.dx = make_const_div(g.Jx);
  g.dy = make_const_div(g.Jy);
  g.variant = 1;
  g.rows = 16;
  if (const char* v = std::getenv("SGNX_RHS_VARIANT")) g.variant = std::atoi(v) == 0 ? 0 : 1;
  if (const char* r = std::getenv("SGNX_SWEEP_ROWS")) g.rows = std::max(1, std::atoi(r));
  g.waves3 = 0x2;   // stage 2 at 3 waves/SIMD, the other kinds at 2 (fastest at 1024x512)
  if (const char* w = std::getenv("SGNX_SWEEP_WAVES3")) g.waves3 = waves_mask_from_text(w);
  for (int k = 0; k < 4; ++k) g.bc[k] = d->bc[k];
  Phys& ph = s->ph;
  ph.g = d->par.g; ph.c = d->par.c; ph.gamma = d->par.gamma; ph.nManning = d->par.nManning;
  ph.halfg = 0.5 * ph.g;
  ph.c2 = ph.c * ph.c;
  ph.mc2 = -ph.c * ph.c;
  ph.m2c2 = -2.0 * ph.c2;
  s->hasInflow = d->bc[0] == SGNX_BC_INFLOW;
  s->inflowOmega = d->inflow_omega;
  // the row sweep leaves out the Manning friction term: with friction the
  // element-row tiles and their separate stage passes evaluate the RHS
  if (ph.nManning != 0.0) g.variant = 0;

  auto cleanup_fail = [&](int code) {
    std::string msg = s->err;
    free_all(s);
    delete s;
    return fail(nullptr, code, msg);
  };
#define CK(call)                                                                   \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      s->err = std::string(#call) + ": " + hipGetErrorString(e_);                  \
      return cleanup_fail(e_ == hipErrorOutOfMemory ? SGNX_ENOMEM : SGNX_EDEVICE); \
    }                                                                              \
  } while (0)

  CK(hipGetDevice(&s->device));
  CK(scratch_address(s->device, &s->scratch));
  if (!std::getenv("SGNX_SWEEP_ROWS")) {
    // default tile height: the tallest that still gives every compute unit
    // about three workgroups (small grids would leave most of the chip idle
    // at 16 rows). The RHS does not depend on it; the adaptive step's error
    // norm is summed per workgroup, so its last bits can.
    int cus = 0;
    CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device));
    const long long strips = (d->Nx + kStripW - 1) / kStripW, want = 3LL * std::max(cus, 1);
    for (int r : {16, 12, 8, 6, 4, 3, 2}) {
      g.rows = r;
      if (strips * ((d->Ny + r - 1) / r) >= want) break;
    }
  }
  const size_t Nd = (size_t)g.Nd, nst = (size_t)NV * Nd;
  const size_t nfx = (size_t)(d->Nx + 1) * d->Ny * NP, nfy = (size_t)d->Nx * (d->Ny + 1) * NP;
  OpTables T;
  std::memcpy(T.D, d->D, sizeof(T.D));
  std::memcpy(T.gLp, d->gLp, sizeof(T.gLp));
  std::memcpy(T.gRp, d->gRp, sizeof(T.gRp));
  std::memcpy(T.qw, d->qw, sizeof(T.qw));
  CK(hipMalloc((void**)&s->d_ops, sizeof(OpTables)));
  CK(hipMemcpy(s->d_ops, &T, sizeof(OpTables), hipMemcpyHostToDevice));
  CK(hipMalloc((void**)&s->d_Hnod, Nd * sizeof(double)));
  CK(hipMalloc((void**)&s->d_Hx, Nd * sizeof(double)));
  CK(hipMalloc((void**)&s->d_Hy, Nd * sizeof(double)));
  CK(hipMalloc((void**)&s->d_HfX, nfx * sizeof(double)));
  CK(hipMalloc((void**)&s->d_HfY, nfy * sizeof(double)));
  g.ops = s->d_ops;
  g.Hnod = s->d_Hnod; g.Hx = s->d_Hx; g.Hy = s->d_Hy; g.HfX = s->d_HfX; g.HfY = s->d_HfY;
  if (d->Hnod) {
    CK(hipMemcpy(s->d_Hnod, d->Hnod, Nd * sizeof(double), hipMemcpyHostToDevice));
    CK(hipMemcpy(s->d_HfX, d->HfaceX, nfx * sizeof(double), hipMemcpyHostToDevice));
    CK(hipMemcpy(s->d_HfY, d->HfaceY, nfy * sizeof(double), hipMemcpyHostToDevice));
    if (d->Hx && d->Hy) {
      CK(hipMemcpy(s->d_Hx, d->Hx, Nd * sizeof(double), hipMemcpyHostToDevice));
      CK(hipMemcpy(s->d_Hy, d->Hy, Nd * sizeof(double), hipMemcpyHostToDevice));
    } else {
      CK(launch_bathy_grad(g, s->d_Hnod, s->d_Hx, s->d_Hy, nullptr));
      CK(hipDeviceSynchronize());
    }
  } else {
    CK(hipMemset(s->d_Hnod, 0, Nd * sizeof(double)));
    CK(hipMemset(s->d_Hx, 0, Nd * sizeof(double)));
    CK(hipMemset(s->d_Hy, 0, Nd * sizeof(double)));
    CK(hipMemset(s->d_HfX, 0, nfx * sizeof(double)));
    CK(hipMemset(s->d_HfY, 0, nfy * sizeof(double)));
  }
  if (s->hasInflow) {
    const size_t ni = (size_t)d->Ny * NP * NV * 3;
    CK(hipMalloc((void**)&s->d_inflow, ni * sizeof(double)));
    CK(hipMemcpy(s->d_inflow, d->inflow, ni * sizeof(double), hipMemcpyHostToDevice));
  }
  g.inflow = s->d_inflow;
  for (double** p : {&s->d_U, &s->d_k1, &s->d_k2, &s->d_k3, &s->d_k4, &s->d_tmp, &s->d_uNew})
    CK(hipMalloc((void**)p, nst * sizeof(double)));
  CK(hipMemset(s->d_U, 0, nst * sizeof(double)));
  {
    // partial triples: kReduceBlocks for the elementwise reductions, one per
    // workgroup of the fused candidate stage (at most strips * Ny)
    const long long strips = (d->Nx + kStripW - 1) / kStripW;
    s->partCap = 3 * std::max<long long>(kReduceBlocks, strips * d->Ny);
    CK(hipMalloc((void**)&s->d_part, (size_t)s->partCap * sizeof(double)));
  }
  CK(hipMalloc((void**)&s->d_out, 2 * sizeof(double)));
  CK(hipHostMalloc((void**)&s->h_out, 2 * sizeof(double), hipHostMallocDefault));
  CK(hipHostMalloc((void**)&s->h_post, 4 * sizeof(double),
                   hipHostMallocMapped | hipHostMallocCoherent));
  CK(hipHostGetDevicePointer((void**)&s->d_post, s->h_post, 0));
  std::memset(s->h_post, 0, 4 * sizeof(double));
  {
    // whether the row sweep may form Hx, Hy from Hnod: they were computed
    // here, or the caller's equal that arithmetic bit for bit
    // (SGNX_SWEEP_HXR=0 makes it read them regardless)
    int* bad = reinterpret_cast<int*>(s->d_out);
    CK(hipMemset(bad, 0, sizeof(int)));
    CK(launch_bathy_check(g, s->d_Hnod, s->d_Hx, s->d_Hy, bad, nullptr));
    int hb = 1;
    CK(hipMemcpy(&hb, bad, sizeof(int), hipMemcpyDeviceToHost));
    g.hxr = hb == 0 && hxr_allowed();
  }
#undef CK
  *out = s;
  return SGNX_OK;
}

int sgnx_destroy(sgnx_solver* s) {
  if (!s) return SGNX_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  else (void)hipDeviceSynchronize();
  free_all(s);
  delete s;
  return SGNX_OK;
}

int sgnx_set_stream(sgnx_solver* s, void* stream) {
  if (!s) return SGNX_EINVAL;
  s->stream = static_cast<hipStream_t>(stream);
  return SGNX_OK;
}

long long sgnx_ndof(const sgnx_solver* s) { return s ? s->g.Nd : 0; }

void* sgnx_state_ptr(sgnx_solver* s) { return s ? s->d_U : nullptr; }

int sgnx_upload(sgnx_solver* s, const double* U, size_t n) {
  if (!s || !U) return SGNX_EINVAL;
  if (n != s->n()) return fail(s, SGNX_EINVAL, "upload: size mismatch");
  HIPCHK(s, hipMemcpyAsync(s->d_U, U, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
  return sync(s);
}

int sgnx_download(const sgnx_solver* cs, double* U, size_t n) {
  auto* s = const_cast<sgnx_solver*>(cs);
  if (!s || !U) return SGNX_EINVAL;
  if (n != s->n()) return fail(s, SGNX_EINVAL, "download: size mismatch");
  HIPCHK(s, hipMemcpyAsync(U, s->d_U, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  return sync(s);
}

int sgnx_rhs(sgnx_solver* s, const double* U, double* dU, size_t n, double t, int mem) {
  if (!s || !U || !dU) return SGNX_EINVAL;
  if (n != s->n()) return fail(s, SGNX_EINVAL, "SGN2D::rhs: size mismatch");
  if (mem == SGNX_MEM_DEVICE) {
    if (U == dU) return fail(s, SGNX_EINVAL, "rhs: U and dU must be distinct");
    return rhs_dev(s, U, dU, t);
  }
  if (mem != SGNX_MEM_HOST) return fail(s, SGNX_EINVAL, "rhs: bad memory kind");
  HIPCHK(s, hipMemcpyAsync(s->d_tmp, U, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
  int rc = rhs_dev(s, s->d_tmp, s->d_k2, t);
  if (rc) return rc;
  HIPCHK(s, hipMemcpyAsync(dU, s->d_k2, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  return sync(s);
}

int sgnx_rhs_once(sgnx_solver* s, double* dU, size_t n, double t) {
  if (!s || !dU) return SGNX_EINVAL;
  if (n != s->n()) return fail(s, SGNX_EINVAL, "rhsOnce: size mismatch");
  int rc = rhs_dev(s, s->d_U, s->d_k1, t);
  if (rc) return rc;
  HIPCHK(s, hipMemcpyAsync(dU, s->d_k1, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  return sync(s);
}

// The attempt's (sum, max) error pair as k_finish_post leaves it in host
// memory: poll its sequence number for a while (a small grid's attempt ends
// within tens of microseconds, sooner than a stream synchronisation returns),
// then synchronise the stream, which also reports a failed launch.
static int wait_post(sgnx_solver* s, double* err) {
  const volatile unsigned long long* seq =
      reinterpret_cast<const volatile unsigned long long*>(s->h_post + 2);
  const volatile double* val = s->h_post;
  const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(300);
  bool seen = false;
  for (int k = 0; !seen; ++k) {
    seen = *seq == s->postSeq;
    if (!seen && (k & 63) == 63 && std::chrono::steady_clock::now() > until) break;
  }
  if (!seen) {
    const int rc = sync(s);
    if (rc) return rc;
    if (*seq != s->postSeq) return fail(s, SGNX_EDEVICE, "adaptive step: error norm not posted");
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  err[0] = val[0];
  err[1] = val[1];
  return SGNX_OK;
}

// The BS3 stages with the row-sweep kernel (rk.cpp:99-131): k2, then the k3
// launch writing the candidate d_uNew and the partial embedded sum (into d_k3),
// then, when `withErr`, k4 = R(uNew) into d_k4 with the scaled error partials
// in d_part.
static int fused_bs3(sgnx_solver* s, double t, double dt, double atol, double rtol,
                     bool withErr) {
  BS3Launch L[3];
  bs3_launches(s, t, dt, atol, rtol, L);
  int rc;
  for (int k = 0; k < (withErr ? 3 : 2); ++k)
    if ((rc = rhs_launch(s, L[k].mode, L[k].sa, L[k].dU, L[k].t))) return rc;
  return SGNX_OK;
}

int sgnx_integrate(sgnx_solver* s, double t0, double tEnd, double dt0, const sgnx_rk_opts* o,
                   sgnx_rk_stats* st, double* step_log, long cap) {
  if (!s || !o || !st) return SGNX_EINVAL;
  const size_t n = s->n();
  RkCtl c(o, t0, tEnd, dt0, n, step_log, cap);
  int rc = rhs_dev(s, s->d_U, s->d_k1, c.t);   // FSAL head
  if (rc) return rc;
  ++c.S.rhsEvals;
  const bool taping = tape_recording(s);
  long stepBegun = -1;
  while (c.next()) {
    const double t = c.t, dt = c.dt;
    if (taping && stepBegun != c.S.accepted) {
      if ((rc = tape_step_begin(s))) return rc;
      stepBegun = c.S.accepted;
    }
    if (s->g.variant == 1) {
      // stage inputs, the candidate and the embedded error fused into the RHS
      if ((rc = fused_bs3(s, t, dt, o->atol, o->rtol, true))) return rc;
      HIPCHK(s, launch_finish_post(s->d_part, sweep_tiles(s->g), s->d_post, ++s->postSeq,
                                   s->stream));
    } else {
      HIPCHK(s, launch_comb1(s->d_tmp, s->d_U, 0.5 * dt, s->d_k1, n, s->stream));
      if ((rc = rhs_dev(s, s->d_tmp, s->d_k2, t + 0.5 * dt))) return rc;
      HIPCHK(s, launch_comb1(s->d_tmp, s->d_U, 0.75 * dt, s->d_k2, n, s->stream));
      if ((rc = rhs_dev(s, s->d_tmp, s->d_k3, t + 0.75 * dt))) return rc;
      HIPCHK(s, launch_bs3_update(s->d_uNew, s->d_U, dt, s->d_k1, s->d_k2, s->d_k3, n,
                                  s->stream));
      if ((rc = rhs_dev(s, s->d_uNew, s->d_k4, t + dt))) return rc;
      HIPCHK(s, launch_err_norm(s->d_U, s->d_uNew, s->d_k1, s->d_k2, s->d_k3, s->d_k4, dt,
                                o->atol, o->rtol, n, s->d_part, s->d_out, s->stream));
    }
    c.S.rhsEvals += 3;
    double err[2];
    if (s->g.variant == 1) {
      if ((rc = wait_post(s, err))) return rc;
    } else {
      HIPCHK(s, hipMemcpyAsync(s->h_out, s->d_out, 2 * sizeof(double), hipMemcpyDeviceToHost,
                               s->stream));
      if ((rc = sync(s))) return rc;
      err[0] = s->h_out[0];
      err[1] = s->h_out[1];
    }
    if (c.finish(err[0], err[1])) {
      std::swap(s->d_U, s->d_uNew);
      std::swap(s->d_k1, s->d_k4);
      const long rows = s->grows;
      HIPCHK(s, record_gauges(s, s->stream, c.t));
      if (taping && (rc = tape_accept(s, t, dt, s->grows > rows ? rows : -1))) return rc;
    }
    if (c.done) break;
  }
  if ((rc = sync(s))) return rc;
  *st = c.S;
  return SGNX_OK;
}

int sgnx_rk4_step(sgnx_solver* s, double t, double dt) {
  if (!s) return SGNX_EINVAL;
  // rk.cpp:23-38
  const size_t n = s->n();
  int rc;
  if ((rc = rhs_dev(s, s->d_U, s->d_k1, t))) return rc;
  HIPCHK(s, launch_comb1(s->d_tmp, s->d_U, 0.5 * dt, s->d_k1, n, s->stream));
  if ((rc = rhs_dev(s, s->d_tmp, s->d_k2, t + 0.5 * dt))) return rc;
  HIPCHK(s, launch_comb1(s->d_tmp, s->d_U, 0.5 * dt, s->d_k2, n, s->stream));
  if ((rc = rhs_dev(s, s->d_tmp, s->d_k3, t + 0.5 * dt))) return rc;
  HIPCHK(s, launch_comb1(s->d_tmp, s->d_U, dt, s->d_k3, n, s->stream));
  if ((rc = rhs_dev(s, s->d_tmp, s->d_k4, t + dt))) return rc;
  HIPCHK(s, launch_rk4_update(s->d_U, dt, s->d_k1, s->d_k2, s->d_k3, s->d_k4, n, s->stream));
  return sync(s);
}

int sgnx_bs3_step(sgnx_solver* s, double t, double dt) {
  if (!s) return SGNX_EINVAL;
  // rk.cpp:41-54
  const size_t n = s->n();
  int rc;
  if ((rc = rhs_dev(s, s->d_U, s->d_k1, t))) return rc;
  if (s->g.variant == 1) {
    // the fused stage kernels of the adaptive loop, without the last RHS
    // (a fixed step has no use for k4 or the error)
    if ((rc = fused_bs3(s, t, dt, 1.0, 0.0, false))) return rc;
  } else {
    HIPCHK(s, launch_comb1(s->d_tmp, s->d_U, 0.5 * dt, s->d_k1, n, s->stream));
    if ((rc = rhs_dev(s, s->d_tmp, s->d_k2, t + 0.5 * dt))) return rc;
    HIPCHK(s, launch_comb1(s->d_tmp, s->d_U, 0.75 * dt, s->d_k2, n, s->stream));
    if ((rc = rhs_dev(s, s->d_tmp, s->d_k3, t + 0.75 * dt))) return rc;
    HIPCHK(s, launch_bs3_update(s->d_uNew, s->d_U, dt, s->d_k1, s->d_k2, s->d_k3, n,
                                s->stream));
  }
  std::swap(s->d_U, s->d_uNew);
  return sync(s);
}

int sgnx_eta_max(sgnx_solver* s, double x0, double x1, double y0, double y1, double* eta,
                 long long* node) {
  if (!s || !eta || !node) return SGNX_EINVAL;
  const int np = s->P + 1;
  std::vector<double> xg(np);
  if (sgnx_fr_operators(s->P, xg.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) !=
      SGNX_OK)
    return fail(s, SGNX_EINVAL, "eta_max: operator tables");
  EtaWindow win{x0, x1, y0, y1, {}, {}};
  for (int a = 0; a < np; ++a) {
    win.xo[a] = 0.5 * s->hx * (xg[a] + 1.0);      // the host's node coordinates
    win.yo[a] = 0.5 * s->hy * (xg[a] + 1.0);
  }
  HIPCHK(s, launch_eta_max(s->g, s->d_U, s->hx, s->hy, win, s->d_part, s->stream));
  std::vector<double> part(2 * kReduceBlocks);
  HIPCHK(s, hipMemcpyAsync(part.data(), s->d_part, part.size() * sizeof(double),
                           hipMemcpyDeviceToHost, s->stream));
  int rc = sync(s);
  if (rc) return rc;
  double best = -1e300;
  long long bi = -1;
  for (int k = 0; k < kReduceBlocks; ++k) {
    const double v = part[2 * k];
    const long long i = (long long)part[2 * k + 1];
    if (v > best || (v == best && i >= 0 && (bi < 0 || i < bi))) {
      best = v;
      bi = i;
    }
  }
  *eta = best;
  *node = bi;
  return SGNX_OK;
}

int sgnx_gauges(sgnx_solver* s, const long long* nodes, int n, double* eta) {
  if (!s || (n > 0 && (!nodes || !eta)) || n < 0) return SGNX_EINVAL;
  if (n == 0) return SGNX_OK;
  for (int k = 0; k < n; ++k)
    if (nodes[k] < 0 || nodes[k] >= s->g.Nd) return fail(s, SGNX_EINVAL, "gauges: node out of range");
  long long* dn = nullptr;
  double* dv = nullptr;
  HIPCHK(s, hipMalloc((void**)&dn, n * sizeof(long long)));
  hipError_t e = hipMalloc((void**)&dv, n * sizeof(double));
  if (e == hipSuccess) e = hipMemcpyAsync(dn, nodes, n * sizeof(long long), hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) e = launch_gather_eta(s->g, s->d_U, dn, n, dv, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(eta, dv, n * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(dn);
  if (dv) (void)hipFree(dv);
  HIPCHK(s, e);
  return SGNX_OK;
}

int sgnx_gauges_noalloc(sgnx_solver* s, const long long* nodes, int n, double* eta) {
  if (!s || (n > 0 && (!nodes || !eta)) || n < 0) return SGNX_EINVAL;
  if (n > SGNX_GAUGES_NOALLOC_MAX)
    return fail(s, SGNX_EINVAL, "gauges_noalloc: " + std::to_string(n) + " gauges, at most " +
                                    std::to_string(SGNX_GAUGES_NOALLOC_MAX) + " per call");
  if (n == 0) return SGNX_OK;
  for (int k = 0; k < n; ++k)
    if (nodes[k] < 0 || nodes[k] >= s->g.Nd) return fail(s, SGNX_EINVAL, "gauges: node out of range");
  if (!s->scratch) return fail(s, SGNX_EDEVICE, "gauges_noalloc: no scratch address");
  // the node list, then the values: every entry read is written first by
  // this call, so nothing needs clearing
  long long* dn = static_cast<long long*>(s->scratch);
  double* dv = reinterpret_cast<double*>(dn + SGNX_GAUGES_NOALLOC_MAX);
  HIPCHK(s, hipMemcpyAsync(dn, nodes, n * sizeof(long long), hipMemcpyHostToDevice, s->stream));
  HIPCHK(s, launch_gather_eta(s->g, s->d_U, dn, n, dv, s->stream));
  HIPCHK(s, hipMemcpyAsync(eta, dv, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return SGNX_OK;
}

int sgnx_element_means(sgnx_solver* s, double* out) {
  if (!s || !out) return SGNX_EINVAL;
  const size_t ne = (size_t)s->Nx * s->Ny;
  double* d = nullptr;
  HIPCHK(s, hipMalloc((void**)&d, 5 * ne * sizeof(double)));
  hipError_t e = launch_elem_means(s->g, s->d_U, d, s->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, d, 5 * ne * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  HIPCHK(s, e);
  return SGNX_OK;
}

int sgnx_set_gauges(sgnx_solver* s, const long long* nodes, int n, long cap) {
  if (!s || n < 0 || cap < 0 || (n > 0 && !nodes)) return SGNX_EINVAL;
  for (int k = 0; k < n; ++k)
    if (nodes[k] < 0 || nodes[k] >= s->g.Nd) return fail(s, SGNX_EINVAL, "gauges: node out of range");
  HIPCHK(s, hipStreamSynchronize(s->stream));
  if (s->d_gnodes) (void)hipFree(s->d_gnodes);
  if (s->d_glog) (void)hipFree(s->d_glog);
  s->d_gnodes = nullptr;
  s->d_glog = nullptr;
  s->nG = 0;
  s->gcap = 0;
  s->grows = 0;
  s->gtimes.clear();
  s->gnodes.clear();
  if (n == 0 || cap == 0) return SGNX_OK;
  s->gnodes.assign(nodes, nodes + n);
  HIPCHK(s, hipMalloc((void**)&s->d_gnodes, n * sizeof(long long)));
  HIPCHK(s, hipMalloc((void**)&s->d_glog, (size_t)n * cap * sizeof(double)));
  HIPCHK(s, hipMemcpy(s->d_gnodes, nodes, n * sizeof(long long), hipMemcpyHostToDevice));
  s->nG = n;
  s->gcap = cap;
  return SGNX_OK;
}

int sgnx_gauge_log(sgnx_solver* s, double* eta, double* t, long cap, long* rows) {
  if (!s || !rows) return SGNX_EINVAL;
  const long r = std::min(cap, s->grows);
  if (r > 0) {
    if (!eta) return SGNX_EINVAL;
    HIPCHK(s, hipMemcpyAsync(eta, s->d_glog, (size_t)r * s->nG * sizeof(double),
                             hipMemcpyDeviceToHost, s->stream));
    int rc = sync(s);
    if (rc) return rc;
    if (t) std::copy(s->gtimes.begin(), s->gtimes.begin() + r, t);
  }
  *rows = r;
  return SGNX_OK;
}

static int diag(sgnx_solver* s, int which, double* out) {
  if (!s || !out) return SGNX_EINVAL;
  HIPCHK(s, launch_diag(s->g, s->ph, s->d_U, which, s->d_part, s->d_out, s->stream));
  HIPCHK(s, hipMemcpyAsync(s->h_out, s->d_out, 2 * sizeof(double), hipMemcpyDeviceToHost,
                           s->stream));
  int rc = sync(s);
  if (rc) return rc;
  *out = (which == 0) ? s->h_out[1] : s->h_out[0];
  return SGNX_OK;
}

int sgnx_max_wave_speed(sgnx_solver* s, double* out) { return diag(s, 0, out); }

int sgnx_cfl_time_step(sgnx_solver* s, double cfl, double* out) {
  double lam = 0.0;
  int rc = diag(s, 0, &lam);
  if (rc) return rc;
  // sgn2d.cpp:310-317
  const double dxEff = std::fmin(s->hx, s->hy) / (2.0 * s->P + 1.0);
  *out = cfl * dxEff / (lam > 0.0 ? lam : 1e-300);
  return SGNX_OK;
}

int sgnx_mass(sgnx_solver* s, double* out) {
  double m = 0.0;
  int rc = diag(s, 1, &m);
  if (rc) return rc;
  *out = m * (s->g.Jx * s->g.Jy);
  return SGNX_OK;
}

int sgnx_energy(sgnx_solver* s, double* out) {
  double E = 0.0;
  int rc = diag(s, 2, &E);
  if (rc) return rc;
  *out = E * (s->g.Jx * s->g.Jy);
  return SGNX_OK;
}

int sgnx_set_rhs_kernel(sgnx_solver* s, int variant, int rows) {
  if (!s || variant < 0 || variant > 1 || rows < 1) return SGNX_EINVAL;
  if (variant == 1 && s->ph.nManning != 0.0)
    return fail(s, SGNX_EUNSUPPORTED, "the row sweep has no Manning friction term (variant 0 has)");
  s->g.variant = variant;
  s->g.rows = rows;
  return SGNX_OK;
}

int sgnx_get_rhs_kernel(const sgnx_solver* s, int* variant, int* rows, int* waves3) {
  if (!s) return SGNX_EINVAL;
  if (variant) *variant = s->g.variant;
  if (rows) *rows = s->g.rows;
  if (waves3) *waves3 = s->g.waves3;
  return SGNX_OK;
}

int sgnx_enable_timing(sgnx_solver* s, int on) {
  if (!s) return SGNX_EINVAL;
  s->timing = on != 0;
  s->timed = 0;
  return SGNX_OK;
}

int sgnx_rhs_timing(sgnx_solver* s, long* launches, double* total_ms) {
  if (!s || !launches || !total_ms) return SGNX_EINVAL;
  int rc = sync(s);
  if (rc) return rc;
  for (int m = 0; m < SGNX_TIMING_MODES; ++m) { launches[m] = 0; total_ms[m] = 0.0; }
  for (long i = 0; i < s->timed; ++i) {
    float ms = 0.0f;
    HIPCHK(s, hipEventElapsedTime(&ms, s->events[2 * i], s->events[2 * i + 1]));
    const int m = s->timedMode[i];
    ++launches[m];
    total_ms[m] += ms;
  }
  s->timed = 0;
  return SGNX_OK;
}

int sgnx_bathymetry(const sgnx_solver* cs, double* Hx, double* Hy) {
  auto* s = const_cast<sgnx_solver*>(cs);
  if (!s || !Hx || !Hy) return SGNX_EINVAL;
  const size_t Nd = (size_t)s->g.Nd;
  HIPCHK(s, hipMemcpyAsync(Hx, s->d_Hx, Nd * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipMemcpyAsync(Hy, s->d_Hy, Nd * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  return sync(s);
}

}  // extern "C"
