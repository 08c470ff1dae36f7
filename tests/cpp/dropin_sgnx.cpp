// The standalone C++ host class sgnx::SGN2D (the reference's sgn::SGN2D API
// with GPU evaluations) against the CPU oracle: same constructor arguments,
// the RHS bitwise, walls and an adaptive run to identical step counts.
// Prints NO_DEVICE and exits 0 when no GPU is visible.
#include "sgnx/sgn2d.hpp"
#include "sgn_oracle.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

int main() {
  const double g = 9.81, Lx = 6.5, Ly = 4.1;
  const int Nx = 13, Ny = 7;
  sgnx::Params par;
  par.g = g; par.gamma = 1.5; par.c = 3.0 * std::sqrt(g);
  auto depth = [](double x, double y) { return 1.0 - 0.2 * std::exp(-((x - 3) * (x - 3) + (y - 2) * (y - 2))); };
  const sgnx::Side sides[4] = {sgnx::Side::Wall, sgnx::Side::Wall, sgnx::Side::Periodic,
                               sgnx::Side::Periodic};
  try {
    sgnx::SGN2D sys(4, Nx, Ny, Lx, Ly, par, depth, sides);
    std::vector<double> x, y;
    sys.nodeCoords(x, y);
    const int Nd = sys.nDof();
    std::vector<double> h(Nd), u(Nd), v(Nd), w(Nd), p(Nd);
    for (int i = 0; i < Nd; ++i) {
      h[i] = depth(x[i], y[i]) + 0.03 * std::sin(x[i]) * std::cos(2 * y[i]);
      u[i] = 0.1 * std::cos(x[i] + y[i]);
      v[i] = -0.07 * std::sin(y[i]);
      w[i] = 0.01 * std::cos(3 * x[i]);
      p[i] = 0.02 * std::sin(x[i] * y[i]);
    }
    const sgnx::State U = sys.makeState(h, u, v, w, p);
    sgnx::State R;
    sys.rhs(U, R);
    // the oracle on the same still-depth samples
    std::vector<double> fxx((Nx + 1) * Ny * 5), fxy(fxx.size()), fyx(Nx * (Ny + 1) * 5), fyy(fyx.size());
    sgnx_face_coords(4, Nx, Ny, Lx, Ly, 0, 1, fxx.data(), fxy.data(), fyx.data(), fyy.data());
    std::vector<double> Hn(Nd), HfX(fxx.size()), HfY(fyx.size());
    for (int i = 0; i < Nd; ++i) Hn[i] = depth(x[i], y[i]);
    for (size_t i = 0; i < HfX.size(); ++i) HfX[i] = depth(fxx[i], fxy[i]);
    for (size_t i = 0; i < HfY.size(); ++i) HfY[i] = depth(fyx[i], fyy[i]);
    orc_desc d{};
    d.P = 4; d.Nx = Nx; d.Ny = Ny; d.Lx = Lx; d.Ly = Ly;
    d.g = par.g; d.c = par.c; d.gamma = par.gamma; d.nManning = 0.0;
    d.bc[0] = d.bc[1] = ORC_BC_WALL; d.bc[2] = d.bc[3] = ORC_BC_PERIODIC;
    d.Hnod = Hn.data(); d.HfaceX = HfX.data(); d.HfaceY = HfY.data();
    char err[256];
    orc_solver* o = orc_create(&d, err, sizeof err);
    if (!o) { std::printf("oracle: %s\n", err); return 1; }
    std::vector<double> Ro(U.size());
    orc_rhs(o, U.data(), Ro.data(), 0.0);
    long bad = 0;
    for (size_t i = 0; i < Ro.size(); ++i) bad += std::memcmp(&Ro[i], &R[i], 8) != 0;
    std::printf("RHS entries differing from the oracle: %ld of %zu\n", bad, Ro.size());
    const double dt0 = sys.cflTimeStep(U, 0.5);
    const double dto = orc_cfl_time_step(o, U.data(), 0.5);
    sgnx::AdaptiveRKOptions opt;
    opt.maxAccepted = 40;
    sys.device().upload(U);
    const auto sg = sys.device().integrateAdaptive(0.0, 1e9, dt0, opt);
    sgnx::State Ug;
    sys.device().download(Ug);
    orc_rk_opts oo{1e-8, 1e-6, 1e-13, 1e30, 0.9, 0.2, 5.0, 40};
    orc_rk_stats so{};
    std::vector<double> Uo = U;
    orc_integrate(o, Uo.data(), 0.0, 1e9, dto, &oo, &so, nullptr, 0, nullptr, nullptr, 0);
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < Uo.size(); ++i) { num += (Ug[i] - Uo[i]) * (Ug[i] - Uo[i]); den += Uo[i] * Uo[i]; }
    const double rel = std::sqrt(num / den);
    std::printf("dt0 %s; steps oracle %ld/%ld sgnx %ld/%ld; rel L2 %.3e\n",
                dt0 == dto ? "equal" : "DIFFERENT", so.accepted, so.rejected, sg.accepted,
                sg.rejected, rel);
    orc_destroy(o);
    const bool ok = bad == 0 && dt0 == dto && so.accepted == sg.accepted &&
                    so.rejected == sg.rejected && rel <= 1e-12;
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
