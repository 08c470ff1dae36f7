// Drop-in check against the reference's own classes: the reference CPU
// solver sgn::SGN2D (compiled from its sources into oracle/_ref/libsgnref.so)
// is handed to sgnx::SGN2DHip exactly as the reference hands it to
// sgn::SGN2DCuda (test/test_cuda_sgn2d.cu:52-80), and the GPU RHS must equal
// host.rhs() bit for bit; then the device-resident adaptive run of the
// reference gate (atol 1e-10, rtol 1e-8, tEnd 0.3) must take the same steps as
// sgn::integrateAdaptive and end within 1e-12 relative L2.
// Prints NO_DEVICE and exits 0 when no GPU is visible.
#include "src/solver/sgn2d.h"
#include "sgnx/sgn2d.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

int main() {
  const double g = 9.81, H0 = 1.0, Lx = 10.0, Ly = 8.0;
  sgn2d::Params par;
  par.g = g; par.gamma = 1.5; par.c = 3.0 * std::sqrt(g * H0);
  auto bump = [Lx, Ly](double x, double y) {
    const double dx = x - 0.5 * Lx, dy = y - 0.5 * Ly;
    return 1.0 - 0.4 * std::exp(-(dx * dx + dy * dy) / 6.0);
  };
  sgn::SGN2D host(4, 8, 6, Lx, Ly, par, bump);
  std::vector<double> x, y;
  host.nodeCoords(x, y);
  const int Nd = host.nDof();
  const double kx = 2.0 * M_PI / Lx, ky = 2.0 * M_PI / Ly;
  std::vector<double> h(Nd), u(Nd), v(Nd), w(Nd), p(Nd);
  for (int i = 0; i < Nd; ++i) {
    h[i] = bump(x[i], y[i]) + 0.04 * std::sin(kx * x[i]) * std::cos(ky * y[i]);
    u[i] = 0.08 * std::cos(kx * x[i]);
    v[i] = -0.05 * std::sin(ky * y[i]);
    w[i] = 0.02 * std::sin(kx * x[i]);
    p[i] = 0.01 * std::cos(ky * y[i]);
  }
  const sgn::State U = host.makeState(h, u, v, w, p);
  try {
    sgnx::SGN2DHip dev(host);
    sgn::State Rc;
    host.rhs(U, Rc);
    dev.upload(U);
    sgnx::State Rg;
    dev.rhsOnce(Rg);
    long bad = 0;
    for (size_t i = 0; i < Rc.size(); ++i) bad += std::memcmp(&Rc[i], &Rg[i], 8) != 0;
    std::printf("RHS entries differing from the reference: %ld of %zu\n", bad, Rc.size());
    sgn::AdaptiveRKOptions opt;
    opt.atol = 1e-10; opt.rtol = 1e-8;
    sgnx::AdaptiveRKOptions gopt;
    gopt.atol = 1e-10; gopt.rtol = 1e-8;
    const double dt0 = host.cflTimeStep(U, 0.5);
    sgn::State Uh = U;
    const auto sh = sgn::integrateAdaptive(host.rhsFn(), Uh, 0.0, 0.3, dt0, opt);
    dev.upload(U);
    const auto sg = dev.integrateAdaptive(0.0, 0.3, dt0, gopt);
    sgnx::State Ug;
    dev.download(Ug);
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < Uh.size(); ++i) {
      num += (Ug[i] - Uh[i]) * (Ug[i] - Uh[i]);
      den += Uh[i] * Uh[i];
    }
    const double rel = std::sqrt(num / den);
    std::printf("steps: reference acc=%ld rej=%ld rhs=%ld | sgnx acc=%ld rej=%ld rhs=%ld; rel L2 %.3e\n",
                sh.accepted, sh.rejected, sh.rhsEvals, sg.accepted, sg.rejected, sg.rhsEvals, rel);
    const bool ok = bad == 0 && sh.accepted == sg.accepted && sh.rejected == sg.rejected &&
                    rel <= 1e-12;
    std::printf(ok ? "DROPIN_OK\n" : "DROPIN_FAIL\n");
    return ok ? 0 : 1;
  } catch (const sgnx::Error& e) {
    if (e.code() == SGNX_EDEVICE || e.code() == SGNX_ENOMEM) {
      std::printf("NO_DEVICE (%s)\n", e.what());
      return 0;
    }
    std::printf("ERROR %d %s\n", e.code(), e.what());
    return 1;
  }
}
