// The gradient entry points of the C++ host class (sgnx::SGN2DHip): a taped
// adaptive run with gauges on walls in x / periodic in y, its reverse sweep,
// and the directional derivative of <ub, U_N> + <wg, G> with respect to the
// initial state and the nodal still depth against fourth-order central
// differences of the device replay of the taped steps. Prints ADJOINT_OK,
// or NO_DEVICE and exits 0 when no GPU is visible.
#include "sgnx/sgn2d.hpp"

#include <cmath>
#include <cstring>
#include <cstdio>
#include <vector>

int main() {
  const double g = 9.81, Lx = 6.0, Ly = 4.0;
  const int Nx = 12, Ny = 8;
  sgnx::Params par;
  par.g = g; par.gamma = 1.5; par.c = 3.0 * std::sqrt(g);
  auto depth = [](double x, double y) {
    return 1.0 - 0.2 * std::exp(-((x - 3) * (x - 3) + (y - 2) * (y - 2)));
  };
  const sgnx::Side sides[4] = {sgnx::Side::Wall, sgnx::Side::Wall, sgnx::Side::Periodic,
                               sgnx::Side::Periodic};
  try {
    sgnx::SGN2D sys(4, Nx, Ny, Lx, Ly, par, depth, sides);
    sgnx::SGN2DHip& dev = sys.device();
    std::vector<double> x, y;
    sys.nodeCoords(x, y);
    const int Nd = sys.nDof();
    std::vector<double> theta(Nd), h(Nd), z(Nd, 0.0), dth(Nd), dh(Nd);
    for (int i = 0; i < Nd; ++i) {
      theta[i] = depth(x[i], y[i]);
      h[i] = theta[i] + 0.02 * std::exp(-((x[i] - 2) * (x[i] - 2) + (y[i] - 1.5) * (y[i] - 1.5)));
      dth[i] = 0.01 * std::exp(-((x[i] - 2.5) * (x[i] - 2.5) + (y[i] - 2) * (y[i] - 2)) / 2.0);
      dh[i] = 0.01 * std::exp(-((x[i] - 3.5) * (x[i] - 3.5) + (y[i] - 2.5) * (y[i] - 2.5)) / 2.0);
    }
    const sgnx::State U0 = sys.makeState(h, z, z, z, z);
    const std::vector<long long> gauges = {40, 777, 1300, 1901};
    const long steps = 8;
    dev.setDepth(theta);
    dev.upload(U0);
    dev.setGauges(gauges, steps);
    const double dt0 = sys.cflTimeStep(U0, 0.5);
    dev.upload(U0);
    sgnx::AdaptiveRKOptions opt;
    opt.atol = 1e-10;
    opt.rtol = 1e-8;
    opt.maxAccepted = steps;
    dev.tapeBegin(1 << 20);
    const sgnx::AdaptiveRKStats st = dev.integrateAdaptive(0.0, 1e9, dt0, opt);
    dev.tapeEnd();
    const std::vector<double> G = dev.gaugeLog();
    sgnx::State UN;
    dev.download(UN);
    std::vector<double> wg(G.size()), ub(UN.size());
    for (size_t k = 0; k < wg.size(); ++k) wg[k] = std::cos(1.0 + 0.7 * (double)k);
    for (size_t k = 0; k < ub.size(); ++k) ub[k] = 1e-3 * std::sin(0.37 * (double)k);
    sgnx::State U0b;
    std::vector<double> thb;
    dev.adjoint(&ub, &wg, U0b, &thb);
    // the same run taped with its stage inputs kept: the same cotangents, bit for bit
    dev.upload(U0);
    dev.setGauges(gauges, steps);
    dev.tapeKeepStages(true);
    dev.tapeBegin(1 << 20);
    dev.integrateAdaptive(0.0, 1e9, dt0, opt);
    dev.tapeEnd();
    dev.tapeKeepStages(false);
    const long kept = dev.tapeStages();
    sgnx::State U0k;
    std::vector<double> thk;
    dev.adjoint(&ub, &wg, U0k, &thk);
    const bool keptSame =
        kept == steps && U0k.size() == U0b.size() && thk.size() == thb.size() &&
        std::memcmp(U0k.data(), U0b.data(), U0b.size() * sizeof(double)) == 0 &&
        std::memcmp(thk.data(), thb.data(), thb.size() * sizeof(double)) == 0;
    double pU = 0.0, pT = 0.0;
    for (int i = 0; i < Nd; ++i) {
      pU += U0b[i] * dh[i];
      pT += thb[i] * dth[i];
    }
    auto loss = [&](double e) {
      std::vector<double> th(Nd), hh(Nd);
      for (int i = 0; i < Nd; ++i) {
        th[i] = theta[i] + e * dth[i];
        hh[i] = h[i] + e * dh[i];
      }
      dev.setDepth(th);
      dev.upload(sys.makeState(hh, z, z, z, z));
      dev.setGauges(gauges, steps);
      dev.tapeReplay();
      const std::vector<double> Gr = dev.gaugeLog();
      sgnx::State Ur;
      dev.download(Ur);
      double L = 0.0;
      for (size_t k = 0; k < Gr.size(); ++k) L += wg[k] * Gr[k];
      for (size_t k = 0; k < Ur.size(); ++k) L += ub[k] * Ur[k];
      return L;
    };
    const double eps = 1e-3;
    const double fd = (8.0 * (loss(eps) - loss(-eps)) - (loss(2 * eps) - loss(-2 * eps))) /
                      (12.0 * eps);
    const double err = std::fabs(pU + pT - fd) / (std::fabs(pU) + std::fabs(pT));
    std::printf("steps %ld, gauge rows %zu; adjoint %.12e (state %.3e, depth %.3e), "
                "differences %.12e, rel err %.2e; kept stages same bits %d\n",
                st.accepted, G.size() / gauges.size(), pU + pT, pU, pT, fd, err, (int)keptSame);
    const bool ok = st.accepted == steps && G.size() == (size_t)steps * gauges.size() &&
                    err < 1e-6 && keptSame;
    std::printf(ok ? "ADJOINT_OK\n" : "ADJOINT_FAIL\n");
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
