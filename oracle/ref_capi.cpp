// =============================================================================
// ref_capi.cpp -- C-ABI over the reference CPU solver (TEST INFRASTRUCTURE ONLY).
//
// Compiled together with the reference's own sources (src/core/fr_operators.cpp,
// src/solver/rk.cpp, src/solver/sgn2d.cpp, read where they lie) into
// oracle/_ref/libsgnref.so by oracle/Makefile. It is the reference build that pins the
// oracle restatement: golden fixtures are written from it and, where it is
// present, tests compare the oracle against it directly. Nothing here is part
// of the product, and no reference source is copied into this repository.
// =============================================================================

#include "src/core/fr_operators.h"
#include "src/physics/sgn_flux2d.h"
#include "src/solver/rk.h"
#include "src/solver/sgn2d.h"

#include <cstdio>
#include <cstring>
#include <exception>
#include <vector>

namespace {

sgn2d::Params parOf(const double* p4) {
  sgn2d::Params p;
  p.g = p4[0]; p.c = p4[1]; p.gamma = p4[2]; p.nManning = p4[3];
  return p;
}

unsigned long long fnv64(const double* v, size_t n) {
  unsigned long long h = 1469598103934665603ull;
  const unsigned char* p = reinterpret_cast<const unsigned char*>(v);
  for (size_t i = 0; i < n * sizeof(double); ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

struct StopAfter {};  // thrown by the logging RHS wrapper to end a run early

}  // namespace

struct ref_solver {
  sgn::SGN2D* s;
};

typedef double (*ref_depth_fn)(double, double);

extern "C" {

int ref_fr_operators(int P, int gauss, double* x, double* D, double* lL, double* lR,
                     double* gLp, double* gRp, double* qw) {
  try {
    sgn::FROperators1D op(P, gauss ? sgn::PointSet::GaussLegendre : sgn::PointSet::GaussLobatto);
    const int n = P + 1;
    std::memcpy(x, op.nodes().data(), n * sizeof(double));
    std::memcpy(D, op.D().data(), (size_t)n * n * sizeof(double));
    std::memcpy(lL, op.lL().data(), n * sizeof(double));
    std::memcpy(lR, op.lR().data(), n * sizeof(double));
    std::memcpy(gLp, op.gLprime().data(), n * sizeof(double));
    std::memcpy(gRp, op.gRprime().data(), n * sizeof(double));
    std::memcpy(qw, op.weights().data(), n * sizeof(double));
    return 0;
  } catch (...) {
    return 1;
  }
}

void ref_flux_x(const double* U, double H, const double* p4, double* F) {
  sgn2d::fluxX(U, H, parOf(p4), F);
}
void ref_flux_y(const double* U, double H, const double* p4, double* G) {
  sgn2d::fluxY(U, H, parOf(p4), G);
}
void ref_source(const double* U, double H, double Hx, double Hy, const double* p4, double* S) {
  sgn2d::source(U, H, Hx, Hy, parOf(p4), S);
}
void ref_ncp_x(const double* U, double hx, const double* p4, double* B) {
  sgn2d::ncpX(U, hx, parOf(p4), B);
}
void ref_ncp_y(const double* U, double hy, const double* p4, double* B) {
  sgn2d::ncpY(U, hy, parOf(p4), B);
}
double ref_signal_speed(const double* U, const double* p4) {
  return sgn2d::signalSpeed(U, parOf(p4));
}
double ref_max_wave_speed(const double* U, int dir, const double* p4) {
  return sgn2d::maxWaveSpeed(U, dir, parOf(p4));
}
void ref_rusanov(const double* UL, const double* UR, int dir, double Hf, const double* p4,
                 double* Fn) {
  sgn2d::rusanov(UL, UR, dir, Hf, parOf(p4), Fn);
}
double ref_energy_density(const double* U, double H, const double* p4) {
  return sgn2d::energyDensity(U, H, parOf(p4));
}

// depth == NULL -> flat bottom (the reference's empty stillDepth).
ref_solver* ref_create(int P, int Nx, int Ny, double Lx, double Ly, const double* p4,
                       ref_depth_fn depth, int gauss, char* err, size_t errlen) {
  try {
    std::function<double(double, double)> f;
    if (depth) f = [depth](double x, double y) { return depth(x, y); };
    auto* s = new sgn::SGN2D(P, Nx, Ny, Lx, Ly, parOf(p4), f,
                             gauss ? sgn::PointSet::GaussLegendre : sgn::PointSet::GaussLobatto);
    return new ref_solver{s};
  } catch (const std::exception& ex) {
    if (err && errlen) std::snprintf(err, errlen, "%s", ex.what());
    return nullptr;
  }
}
void ref_destroy(ref_solver* s) {
  if (s) { delete s->s; delete s; }
}
long ref_ndof(const ref_solver* s) { return s->s->nDof(); }
void ref_node_coords(const ref_solver* s, double* x, double* y) {
  std::vector<double> X, Y;
  s->s->nodeCoords(X, Y);
  std::memcpy(x, X.data(), X.size() * sizeof(double));
  std::memcpy(y, Y.data(), Y.size() * sizeof(double));
}
// Hnod, Hx, Hy: nDof each; HfaceX, HfaceY: nElem*(P+1) each (left / bottom edges).
void ref_bathymetry(const ref_solver* s, double* Hnod, double* Hx, double* Hy, double* HfX,
                    double* HfY) {
  const auto& S = *s->s;
  std::memcpy(Hnod, S.stillDepthNodal().data(), S.stillDepthNodal().size() * sizeof(double));
  std::memcpy(Hx, S.stillDepthGradX().data(), S.stillDepthGradX().size() * sizeof(double));
  std::memcpy(Hy, S.stillDepthGradY().data(), S.stillDepthGradY().size() * sizeof(double));
  std::memcpy(HfX, S.stillDepthFaceX().data(), S.stillDepthFaceX().size() * sizeof(double));
  std::memcpy(HfY, S.stillDepthFaceY().data(), S.stillDepthFaceY().size() * sizeof(double));
}
int ref_rhs(const ref_solver* s, const double* U, double* dU) {
  try {
    sgn::State u(U, U + s->s->nTotal()), du;
    s->s->rhs(u, du);
    std::memcpy(dU, du.data(), du.size() * sizeof(double));
    return 0;
  } catch (...) {
    return 1;
  }
}
double ref_max_wave_speed_state(const ref_solver* s, const double* U) {
  sgn::State u(U, U + s->s->nTotal());
  return s->s->maxWaveSpeed(u);
}
double ref_cfl_time_step(const ref_solver* s, const double* U, double cfl) {
  sgn::State u(U, U + s->s->nTotal());
  return s->s->cflTimeStep(u, cfl);
}
double ref_mass(const ref_solver* s, const double* U) {
  sgn::State u(U, U + s->s->nTotal());
  return s->s->mass(u);
}
double ref_energy(const ref_solver* s, const double* U) {
  sgn::State u(U, U + s->s->nTotal());
  return s->s->energy(u);
}
void ref_rk4_step(const ref_solver* s, double t, double dt, double* U) {
  sgn::State u(U, U + s->s->nTotal());
  sgn::rk4Step(s->s->rhsFn(), t, dt, u);
  std::memcpy(U, u.data(), u.size() * sizeof(double));
}
void ref_bs3_step(const ref_solver* s, double t, double dt, double* U) {
  sgn::State u(U, U + s->s->nTotal());
  sgn::bs3Step(s->s->rhsFn(), t, dt, u);
  std::memcpy(U, u.data(), u.size() * sizeof(double));
}

// Adaptive run. Every RHS call logs (t, digest of its output); with
// max_calls > 0 the run is cut short when RHS call max_calls+1 begins, leaving
// U at the last accepted state. stats = {accepted, rejected, rhsEvals} longs
// and {tReached, dtLast, maxNormLast} doubles (only valid when not cut short);
// returns 0 finished, 2 cut short, 1 error.
int ref_integrate(const ref_solver* s, double* U, double t0, double tEnd, double dt0,
                  const double* opt7, long max_calls, double* call_t,
                  unsigned long long* call_h, long call_cap, long* stats_l, double* stats_d,
                  int* aborted) {
  sgn::State u(U, U + s->s->nTotal());
  sgn::AdaptiveRKOptions opt;
  opt.atol = opt7[0]; opt.rtol = opt7[1]; opt.dtMin = opt7[2]; opt.dtMax = opt7[3];
  opt.safety = opt7[4]; opt.facMin = opt7[5]; opt.facMax = opt7[6];
  long calls = 0;
  const auto base = s->s->rhsFn();
  sgn::RhsFn f = [&](double t, const sgn::State& x, sgn::State& dx) {
    if (max_calls > 0 && calls >= max_calls) throw StopAfter{};
    base(t, x, dx);
    if (calls < call_cap) { call_t[calls] = t; call_h[calls] = fnv64(dx.data(), dx.size()); }
    ++calls;
  };
  int rc = 0;
  try {
    const auto st = sgn::integrateAdaptive(f, u, t0, tEnd, dt0, opt);
    stats_l[0] = st.accepted; stats_l[1] = st.rejected; stats_l[2] = st.rhsEvals;
    stats_d[0] = st.tReached; stats_d[1] = st.dtLast; stats_d[2] = st.maxNormLast;
    *aborted = st.aborted ? 1 : 0;
  } catch (const StopAfter&) {
    rc = 2;
  } catch (...) {
    return 1;
  }
  std::memcpy(U, u.data(), u.size() * sizeof(double));
  return rc;
}

// Reference integrators on the scalar ODEs of test/test_rk.cpp (kind 0: u'=-u,
// kind 1: oscillator, kind 2: u' = param*u).
static sgn::RhsFn odeFn(int kind, double param) {
  if (kind == 0) return [](double, const sgn::State& u, sgn::State& du) { du[0] = -u[0]; };
  if (kind == 1)
    return [](double, const sgn::State& u, sgn::State& du) { du[0] = u[1]; du[1] = -u[0]; };
  return [param](double, const sgn::State& u, sgn::State& du) { du[0] = param * u[0]; };
}
int ref_integrate_ode(int kind, double param, double* u, int n, double t0, double tEnd,
                      double dt0, const double* opt7, long* stats_l, double* stats_d,
                      int* aborted) {
  sgn::State v(u, u + n);
  sgn::AdaptiveRKOptions opt;
  opt.atol = opt7[0]; opt.rtol = opt7[1]; opt.dtMin = opt7[2]; opt.dtMax = opt7[3];
  opt.safety = opt7[4]; opt.facMin = opt7[5]; opt.facMax = opt7[6];
  const auto st = sgn::integrateAdaptive(odeFn(kind, param), v, t0, tEnd, dt0, opt);
  stats_l[0] = st.accepted; stats_l[1] = st.rejected; stats_l[2] = st.rhsEvals;
  stats_d[0] = st.tReached; stats_d[1] = st.dtLast; stats_d[2] = st.maxNormLast;
  *aborted = st.aborted ? 1 : 0;
  std::memcpy(u, v.data(), n * sizeof(double));
  return 0;
}
void ref_rk4_step_ode(int kind, double param, double* u, int n, double t, double dt) {
  sgn::State v(u, u + n);
  sgn::rk4Step(odeFn(kind, param), t, dt, v);
  std::memcpy(u, v.data(), n * sizeof(double));
}
void ref_bs3_step_ode(int kind, double param, double* u, int n, double t, double dt) {
  sgn::State v(u, u + n);
  sgn::bs3Step(odeFn(kind, param), t, dt, v);
  std::memcpy(u, v.data(), n * sizeof(double));
}

}  // extern "C"
