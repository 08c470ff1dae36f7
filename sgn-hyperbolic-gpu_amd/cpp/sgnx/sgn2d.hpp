// =============================================================================
// sgnx/sgn2d.hpp -- C++ host classes over the libsgnx C-ABI (include/sgnx.h).
//
//   sgnx::SGN2DHip  drop-in for the reference device class sgn::SGN2DCuda
//                   (src/cuda/sgn2d_cuda.h:30-79): constructed from a host
//                   solver object, same methods (upload, download, rhsOnce,
//                   integrateAdaptive). The constructor is a template over the
//                   host type, so it accepts the reference's sgn::SGN2D as is
//                   (it reads only the accessors src/solver/sgn2d.h:61-98
//                   declares) as well as sgnx::SGN2D below.
//   sgnx::SGN2D     the reference host API of sgn::SGN2D (src/solver/sgn2d.h:
//                   49-113: constructor arguments, nDof, nTotal, nodeCoords,
//                   makeState, rhs, rhsFn, maxWaveSpeed, cflTimeStep, mass,
//                   energy, accessors) with every evaluation on the GPU, plus
//                   the side kinds the reference lacks (walls, inflow).
//
// Header-only; link with libsgnx.so. Errors: the reference aborts on a device
// error (sgn2d_cuda.cu:29-37) and throws std::invalid_argument on bad setup
// (sgn2d.cpp:36,47,129); here setup errors throw std::invalid_argument and
// device errors throw sgnx::Error with the library's message. An adaptive run
// that stops at dtMin is data (stats.aborted), as in rk.cpp:155-160.
// =============================================================================
#ifndef SGNX_SGN2D_HPP
#define SGNX_SGN2D_HPP

#include <algorithm>
#include <cstddef>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/sgnx.h"

namespace sgnx {

using State = std::vector<double>;                                       // rk.h:36
using RhsFn = std::function<void(double t, const State& u, State& dudt)>;  // rk.h:39

// AdaptiveRKOptions (rk.h:53-61) plus maxAccepted (0 = off).
struct AdaptiveRKOptions {
  double atol = 1e-8, rtol = 1e-6, dtMin = 1e-13, dtMax = 1e30;
  double safety = 0.9, facMin = 0.2, facMax = 5.0;
  long maxAccepted = 0;
};

// AdaptiveRKStats (rk.h:64-72).
struct AdaptiveRKStats {
  long accepted = 0, rejected = 0, rhsEvals = 0;
  double tReached = 0.0, dtLast = 0.0, maxNormLast = 0.0;
  bool aborted = false;
};

// sgn2d::Params (sgn_flux2d.h:57-62).
struct Params {
  double g = 9.81, c = 0.0, gamma = 1.5, nManning = 0.0;
};

enum class Side { Periodic = SGNX_BC_PERIODIC, Wall = SGNX_BC_WALL, Inflow = SGNX_BC_INFLOW };

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

namespace detail {

inline void check(int rc, const sgnx_solver* s) {
  if (rc == SGNX_OK) return;
  const std::string msg = sgnx_last_error(s);
  if (rc == SGNX_EINVAL) throw std::invalid_argument(msg);
  throw Error(rc, msg);
}

inline sgnx_rk_opts opts(const AdaptiveRKOptions& o) {
  return sgnx_rk_opts{o.atol, o.rtol, o.dtMin, o.dtMax, o.safety, o.facMin, o.facMax,
                      o.maxAccepted};
}

inline AdaptiveRKStats stats(const sgnx_rk_stats& s) {
  AdaptiveRKStats r;
  r.accepted = s.accepted; r.rejected = s.rejected; r.rhsEvals = s.rhsEvals;
  r.tReached = s.tReached; r.dtLast = s.dtLast; r.maxNormLast = s.maxNormLast;
  r.aborted = s.aborted != 0;
  return r;
}

// Left / bottom edge depths of the reference layout (ne*np, sgn2d.h:108-112)
// extended by the periodic wrap of the last line (the reference samples the
// right / top edge of the last column / row at the wrapped coordinate).
inline std::vector<double> interfaces_x(const std::vector<double>& left, int Nx, int Ny, int np) {
  std::vector<double> out((size_t)(Nx + 1) * Ny * np);
  for (int ey = 0; ey < Ny; ++ey)
    for (int i = 0; i <= Nx; ++i)
      for (int t = 0; t < np; ++t)
        out[((size_t)ey * (Nx + 1) + i) * np + t] = left[((size_t)ey * Nx + (i % Nx)) * np + t];
  return out;
}
inline std::vector<double> interfaces_y(const std::vector<double>& bottom, int Nx, int Ny, int np) {
  std::vector<double> out((size_t)Nx * (Ny + 1) * np);
  for (int j = 0; j <= Ny; ++j)
    for (int ex = 0; ex < Nx; ++ex)
      for (int t = 0; t < np; ++t)
        out[((size_t)j * Nx + ex) * np + t] = bottom[((size_t)(j % Ny) * Nx + ex) * np + t];
  return out;
}

}  // namespace detail

// ---------------------------------------------------------------------------
// Device solver from a host solver (drop-in for sgn::SGN2DCuda).
// ---------------------------------------------------------------------------
class SGN2DHip {
 public:
  // Copies tables, bathymetry and parameters out of `host` (sgn2d_cuda.cu:318-400);
  // holds no reference to it afterwards. All four sides periodic.
  template <class HostSolver>
  explicit SGN2DHip(const HostSolver& host) {
    const auto& mesh = host.mesh();
    const auto& ops = host.ops();
    const auto& par = host.params();
    const int np = host.order() + 1;
    const std::vector<double> HfX =
        detail::interfaces_x(host.stillDepthFaceX(), mesh.Nx, mesh.Ny, np);
    const std::vector<double> HfY =
        detail::interfaces_y(host.stillDepthFaceY(), mesh.Nx, mesh.Ny, np);
    sgnx_desc d{};
    d.P = host.order();
    d.Nx = mesh.Nx;
    d.Ny = mesh.Ny;
    d.Lx = mesh.Lx;
    d.Ly = mesh.Ly;
    d.par = sgnx_params{par.g, par.c, par.gamma, par.nManning};
    d.D = ops.D().data();
    d.lL = ops.lL().data();
    d.lR = ops.lR().data();
    d.gLp = ops.gLprime().data();
    d.gRp = ops.gRprime().data();
    d.qw = ops.weights().data();
    d.Hnod = host.stillDepthNodal().data();
    d.Hx = host.stillDepthGradX().data();
    d.Hy = host.stillDepthGradY().data();
    d.HfaceX = HfX.data();
    d.HfaceY = HfY.data();
    for (int k = 0; k < 4; ++k) d.bc[k] = SGNX_BC_PERIODIC;
    detail::check(sgnx_create(&d, &s_), nullptr);
  }
  explicit SGN2DHip(const sgnx_desc& d) { detail::check(sgnx_create(&d, &s_), nullptr); }
  ~SGN2DHip() { sgnx_destroy(s_); }
  SGN2DHip(const SGN2DHip&) = delete;
  SGN2DHip& operator=(const SGN2DHip&) = delete;

  int nDof() const { return (int)sgnx_ndof(s_); }
  int nTotal() const { return 5 * nDof(); }

  void upload(const State& U) { detail::check(sgnx_upload(s_, U.data(), U.size()), s_); }
  void download(State& U) const {
    U.resize((size_t)nTotal());
    detail::check(sgnx_download(s_, U.data(), U.size()), s_);
  }
  // One RHS of the uploaded state, downloaded (t enters inflow ghosts only).
  void rhsOnce(State& dU, double t = 0.0) {
    dU.resize((size_t)nTotal());
    detail::check(sgnx_rhs_once(s_, dU.data(), dU.size(), t), s_);
  }
  // Device-resident BS3(2) on the uploaded state; the result stays on the
  // device until download().
  AdaptiveRKStats integrateAdaptive(double t0, double tEnd, double dt0,
                                    const AdaptiveRKOptions& opt = {}) {
    const sgnx_rk_opts o = detail::opts(opt);
    sgnx_rk_stats st{};
    detail::check(sgnx_integrate(s_, t0, tEnd, dt0, &o, &st, nullptr, 0), s_);
    return detail::stats(st);
  }
  void rk4Step(double t, double dt) { detail::check(sgnx_rk4_step(s_, t, dt), s_); }
  void bs3Step(double t, double dt) { detail::check(sgnx_bs3_step(s_, t, dt), s_); }

  // ---- gauges and the discrete adjoint (no counterpart in the reference) ----
  // eta = h - H at `nodes` after every accepted step, up to cap rows.
  void setGauges(const std::vector<long long>& nodes, long cap) {
    detail::check(sgnx_set_gauges(s_, nodes.data(), (int)nodes.size(), cap), s_);
    nG_ = (int)nodes.size();
    gcap_ = cap;
  }
  // the recorded rows, row-major (rows x gauges)
  std::vector<double> gaugeLog() {
    std::vector<double> eta((size_t)std::max(1L, gcap_) * std::max(1, nG_));
    long rows = 0;
    detail::check(sgnx_gauge_log(s_, eta.data(), nullptr, gcap_, &rows), s_);
    eta.resize((size_t)rows * nG_);
    return eta;
  }
  // still depth from the nodal field theta (face depths: the mean of the two
  // edge values meeting at a line node)
  void setDepth(const std::vector<double>& theta) {
    detail::check(sgnx_set_depth(s_, theta.data(), theta.size(), SGNX_MEM_HOST), s_);
  }
  // Ubar = (dR/dU)^T W, thetaBar = (dR/dtheta)^T W (when given)
  void rhsVjp(const State& U, const State& W, State& Ubar, std::vector<double>* thetaBar = nullptr,
              double t = 0.0) {
    Ubar.resize(U.size());
    if (thetaBar) thetaBar->resize((size_t)nDof());
    detail::check(sgnx_rhs_vjp(s_, U.data(), W.data(), Ubar.data(),
                               thetaBar ? thetaBar->data() : nullptr, U.size(), t, SGNX_MEM_HOST),
                  s_);
  }
  // record the following integrateAdaptive calls (at most maxSnapshots states)
  void tapeBegin(long maxSnapshots) { detail::check(sgnx_tape_begin(s_, maxSnapshots), s_); }
  void tapeBeginPlan(long steps, long states) {
    detail::check(sgnx_tape_begin_plan(s_, steps, states), s_);
  }
  void tapeEnd() { detail::check(sgnx_tape_end(s_), s_); }
  void tapeSlots(long slots) { detail::check(sgnx_tape_set_slots(s_, slots), s_); }
  // keep each taped step's stage inputs (two more states per step) so the
  // reverse sweep recomputes no stage; set before the run
  void tapeKeepStages(bool on) { detail::check(sgnx_tape_keep_stages(s_, on ? 1 : 0), s_); }
  // steps whose stage inputs the tape keeps
  long tapeStages() const {
    long n = 0;
    detail::check(sgnx_tape_stages(s_, &n), s_);
    return n;
  }
  // the recorded steps again (same t, dt) from the resident state
  void tapeReplay() { detail::check(sgnx_tape_replay(s_), s_); }
  // cotangents of the recorded run's initial state and still depth
  void adjoint(const State* UbarFinal, const std::vector<double>* gaugeBar, State& U0bar,
               std::vector<double>* thetaBar = nullptr) {
    U0bar.resize((size_t)nTotal());
    if (thetaBar) thetaBar->resize((size_t)nDof());
    detail::check(sgnx_adjoint(s_, UbarFinal ? UbarFinal->data() : nullptr,
                               gaugeBar ? gaugeBar->data() : nullptr, U0bar.data(),
                               thetaBar ? thetaBar->data() : nullptr, SGNX_MEM_HOST),
                  s_);
  }

  sgnx_solver* handle() { return s_; }

 private:
  sgnx_solver* s_ = nullptr;
  int nG_ = 0;
  long gcap_ = 0;
};

// ---------------------------------------------------------------------------
// The reference host API of sgn::SGN2D with GPU evaluations.
// ---------------------------------------------------------------------------
class SGN2D {
 public:
  // P, Nx, Ny, Lx, Ly, par, stillDepth as sgn::SGN2D (sgn2d.h:57-59; Gauss-
  // Lobatto points). sides = {left, right, bottom, top}; an Inflow left side
  // takes `inflow` = per (ey, t) face node and variable the coefficients
  // {base, sin, cos} of (base + s*sin(omega t)) + c*cos(omega t).
  SGN2D(int P, int Nx, int Ny, double Lx, double Ly, const Params& par,
        std::function<double(double, double)> stillDepth = {},
        const Side (&sides)[4] = kPeriodic, std::vector<double> inflow = {},
        double omega = 0.0)
      : P_(P), Nx_(Nx), Ny_(Ny), Lx_(Lx), Ly_(Ly), par_(par) {
    if (Nx < 2 || Ny < 2) throw std::invalid_argument("SGN2D: Nx,Ny >= 2");
    const int np = P + 1;
    x_.assign(np, 0.0); D_.assign((size_t)np * np, 0.0); lL_.assign(np, 0.0);
    lR_.assign(np, 0.0); gLp_.assign(np, 0.0); gRp_.assign(np, 0.0); qw_.assign(np, 0.0);
    detail::check(sgnx_fr_operators(P, x_.data(), D_.data(), lL_.data(), lR_.data(),
                                     gLp_.data(), gRp_.data(), qw_.data()), nullptr);
    const size_t Nd = (size_t)Nx * Ny * np * np;
    const size_t nfx = (size_t)(Nx + 1) * Ny * np, nfy = (size_t)Nx * (Ny + 1) * np;
    sgnx_desc d{};
    d.P = P; d.Nx = Nx; d.Ny = Ny; d.Lx = Lx; d.Ly = Ly;
    d.par = sgnx_params{par.g, par.c, par.gamma, par.nManning};
    d.D = D_.data(); d.lL = lL_.data(); d.lR = lR_.data();
    d.gLp = gLp_.data(); d.gRp = gRp_.data(); d.qw = qw_.data();
    for (int k = 0; k < 4; ++k) d.bc[k] = static_cast<int>(sides[k]);
    if (stillDepth) {
      std::vector<double> x(Nd), y(Nd), fxx(nfx), fxy(nfx), fyx(nfy), fyy(nfy);
      detail::check(sgnx_node_coords(P, Nx, Ny, Lx, Ly, x.data(), y.data()), nullptr);
      detail::check(sgnx_face_coords(P, Nx, Ny, Lx, Ly, sides[0] == Side::Periodic,
                                     sides[2] == Side::Periodic, fxx.data(), fxy.data(),
                                     fyx.data(), fyy.data()), nullptr);
      Hnod_.resize(Nd); HfX_.resize(nfx); HfY_.resize(nfy);
      for (size_t i = 0; i < Nd; ++i) Hnod_[i] = stillDepth(x[i], y[i]);
      for (size_t i = 0; i < nfx; ++i) HfX_[i] = stillDepth(fxx[i], fxy[i]);
      for (size_t i = 0; i < nfy; ++i) HfY_[i] = stillDepth(fyx[i], fyy[i]);
      d.Hnod = Hnod_.data(); d.HfaceX = HfX_.data(); d.HfaceY = HfY_.data();
    }
    inflow_ = std::move(inflow);
    if (!inflow_.empty()) { d.inflow = inflow_.data(); d.inflow_omega = omega; }
    dev_ = new SGN2DHip(d);
  }
  ~SGN2D() { delete dev_; }
  SGN2D(const SGN2D&) = delete;
  SGN2D& operator=(const SGN2D&) = delete;

  int order() const { return P_; }
  int nElem() const { return Nx_ * Ny_; }
  int nodesPerElem() const { return (P_ + 1) * (P_ + 1); }
  int nDof() const { return nElem() * nodesPerElem(); }
  int nTotal() const { return 5 * nDof(); }
  const Params& params() const { return par_; }

  void nodeCoords(std::vector<double>& x, std::vector<double>& y) const {
    x.resize((size_t)nDof()); y.resize((size_t)nDof());
    detail::check(sgnx_node_coords(P_, Nx_, Ny_, Lx_, Ly_, x.data(), y.data()), nullptr);
  }
  // (h, u, v, w, p) -> var-major conserved state (sgn2d.cpp:107-120).
  State makeState(const std::vector<double>& h, const std::vector<double>& u,
                  const std::vector<double>& v, const std::vector<double>& w,
                  const std::vector<double>& p) const {
    const size_t Nd = (size_t)nDof();
    State U(5 * Nd);
    for (size_t i = 0; i < Nd; ++i) {
      U[i] = h[i]; U[Nd + i] = h[i] * u[i]; U[2 * Nd + i] = h[i] * v[i];
      U[3 * Nd + i] = h[i] * w[i]; U[4 * Nd + i] = h[i] * p[i];
    }
    return U;
  }

  // R(U) on the GPU, bitwise equal to the reference CPU rhs() (sgn2d.cpp:125-290).
  void rhs(const State& U, State& dUdt, double t = 0.0) const {
    if ((int)U.size() != nTotal()) throw std::invalid_argument("SGN2D::rhs: size mismatch");
    dUdt.resize(U.size());
    detail::check(sgnx_rhs(dev_->handle(), U.data(), dUdt.data(), U.size(), t, SGNX_MEM_HOST),
                  dev_->handle());
  }
  RhsFn rhsFn() const {
    return [this](double t, const State& U, State& dU) { rhs(U, dU, t); };
  }
  double maxWaveSpeed(const State& U) { return diag(U, &sgnx_max_wave_speed); }
  double cflTimeStep(const State& U, double cfl) {
    dev_->upload(U);
    double out = 0.0;
    detail::check(sgnx_cfl_time_step(dev_->handle(), cfl, &out), dev_->handle());
    return out;
  }
  double mass(const State& U) { return diag(U, &sgnx_mass); }
  double energy(const State& U) { return diag(U, &sgnx_energy); }

  // Device-class methods on the resident state.
  SGN2DHip& device() { return *dev_; }

  const std::vector<double>& nodes() const { return x_; }
  const std::vector<double>& stillDepthNodal() const { return Hnod_; }

 private:
  static constexpr Side kPeriodic[4] = {Side::Periodic, Side::Periodic, Side::Periodic,
                                        Side::Periodic};
  double diag(const State& U, int (*fn)(sgnx_solver*, double*)) {
    dev_->upload(U);
    double out = 0.0;
    detail::check(fn(dev_->handle(), &out), dev_->handle());
    return out;
  }

  int P_, Nx_, Ny_;
  double Lx_, Ly_;
  Params par_;
  std::vector<double> x_, D_, lL_, lR_, gLp_, gRp_, qw_;
  std::vector<double> Hnod_, HfX_, HfY_, inflow_;
  SGN2DHip* dev_ = nullptr;
};

}  // namespace sgnx

#endif  // SGNX_SGN2D_HPP
