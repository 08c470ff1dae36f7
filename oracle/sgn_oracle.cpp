// =============================================================================
// sgn_oracle.cpp -- CPU oracle: restatement of the reference 2D FR/SGN path.
//
// TEST INFRASTRUCTURE ONLY. Loaded by tests/, __graft_entry__.smoke() and the
// CPU-baseline leg of bench.py, to check results against; never linked by the product.
//
// Every floating-point expression below keeps the evaluation order of the
// reference line it cites: the oracle's bits are the contract the HIP kernels
// are held to. Build with -O2 -ffp-contract=off (oracle/Makefile).
//
// Extension over the reference (whose 2D mesh is periodic on every side,
// src/core/mesh2d.h:36-45): each side may instead be a reflective wall (ghost =
// interior face state with the normal momentum negated) or, on the left side,
// an inflow whose ghost state is (base + s*sin(w t)) + c*cos(w t) per face node.
// With all four sides periodic the oracle is the reference, bit for bit.
// =============================================================================

#include "sgn_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <vector>

namespace orc {

constexpr int NV = 5;             // conserved variables h, hu, hv, hw, hp
constexpr int IH = 0, IHU = 1, IHV = 2, IHW = 3, IHP = 4;
constexpr double kPi = 3.14159265358979323846;

struct Par { double g, c, gamma, nManning; };

static Par parFrom(const double* p4) { return Par{p4[0], p4[1], p4[2], p4[3]}; }

// ---- pointwise physics: src/physics/sgn_flux2d.h ----------------------------

// sgn_flux2d.h:73-78 (signalSpeed)
static double signalSpeed(const double* U, const Par& par) {
  const double h = U[IH];
  const double p = U[IHP] / h;
  const double arg = par.g * h + p + par.c * par.c;
  return std::sqrt(arg > 0.0 ? arg : 0.0);
}

// sgn_flux2d.h:82-86 (maxWaveSpeed)
static double waveSpeed(const double* U, int dir, const Par& par) {
  const double vn = (dir == 0 ? U[IHU] : U[IHV]) / U[IH];
  return std::fabs(vn) + signalSpeed(U, par);
}

// sgn_flux2d.h:89-101 (fluxX)
static void fluxX(const double* U, double Hl, const Par& par, double* F) {
  const double h = U[IH], hu = U[IHU];
  const double u = hu / h, v = U[IHV] / h, w = U[IHW] / h, p = U[IHP] / h;
  const double c2 = par.c * par.c;
  F[IH] = hu;
  F[IHU] = hu * u + 0.5 * par.g * (h * h - Hl * Hl) + h * p;
  F[IHV] = hu * v;
  F[IHW] = hu * w;
  F[IHP] = hu * (p + c2);
}

// sgn_flux2d.h:104-116 (fluxY)
static void fluxY(const double* U, double Hl, const Par& par, double* G) {
  const double h = U[IH], hv = U[IHV];
  const double u = U[IHU] / h, v = hv / h, w = U[IHW] / h, p = U[IHP] / h;
  const double c2 = par.c * par.c;
  G[IH] = hv;
  G[IHU] = hv * u;
  G[IHV] = hv * v + 0.5 * par.g * (h * h - Hl * Hl) + h * p;
  G[IHW] = hv * w;
  G[IHP] = hv * (p + c2);
}

// sgn_flux2d.h:122-143 (source, Manning branch :135-142)
static void source(const double* U, double Hl, double Hx, double Hy, const Par& par,
                   double* S) {
  const double h = U[IH];
  const double u = U[IHU] / h, v = U[IHV] / h, w = U[IHW] / h, p = U[IHP] / h;
  const double c2 = par.c * par.c;
  const double bath = par.g * (h - Hl) + par.gamma * p;
  S[IH] = 0.0;
  S[IHU] = bath * Hx;
  S[IHV] = bath * Hy;
  S[IHW] = par.gamma * p;
  S[IHP] = -2.0 * c2 * (w + u * Hx + v * Hy);
  if (par.nManning != 0.0) {
    const double n2 = par.nManning * par.nManning;
    const double vmag = std::sqrt(u * u + v * v);
    const double cf = n2 * par.g * vmag / std::cbrt(h);
    S[IHU] -= cf * u;
    S[IHV] -= cf * v;
  }
}

// sgn_flux2d.h:150-155 (ncpX) and :156-161 (ncpY)
static void ncpX(const double* U, double hx, const Par& par, double* B) {
  const double u = U[IHU] / U[IH];
  B[0] = 0.0; B[1] = 0.0; B[2] = 0.0; B[3] = 0.0;
  B[IHP] = -par.c * par.c * u * hx;
}
static void ncpY(const double* U, double hy, const Par& par, double* B) {
  const double v = U[IHV] / U[IH];
  B[0] = 0.0; B[1] = 0.0; B[2] = 0.0; B[3] = 0.0;
  B[IHP] = -par.c * par.c * v * hy;
}

// sgn_flux2d.h:169-179 (rusanov)
static void rusanov(const double* UL, const double* UR, int dir, double Hf, const Par& par,
                    double* Fn) {
  double FL[NV], FR[NV];
  if (dir == 0) { fluxX(UL, Hf, par, FL); fluxX(UR, Hf, par, FR); }
  else          { fluxY(UL, Hf, par, FL); fluxY(UR, Hf, par, FR); }
  const double lam = std::fmax(waveSpeed(UL, dir, par), waveSpeed(UR, dir, par));
  for (int k = 0; k < NV; ++k)
    Fn[k] = 0.5 * (FL[k] + FR[k]) - 0.5 * lam * (UR[k] - UL[k]);
}

// sgn_flux2d.h:185-191 (energyDensity)
static double energyDensity(const double* U, double Hl, const Par& par) {
  const double h = U[IH];
  const double u = U[IHU] / h, v = U[IHV] / h, w = U[IHW] / h;
  return 0.5 * h * (u * u + v * v + (1.0 / par.gamma) * w * w + par.g * (h - 2.0 * Hl));
}

// ---- 1D FR operators: src/core/fr_operators.cpp ------------------------------

// fr_operators.cpp:27-37 (three-term Legendre recurrence)
static double legendre(int n, double x) {
  if (n == 0) return 1.0;
  if (n == 1) return x;
  double Lm1 = 1.0, L = x;
  for (int k = 1; k < n; ++k) {
    const double Lp1 = ((2.0 * k + 1.0) * x * L - k * Lm1) / (k + 1.0);
    Lm1 = L;
    L = Lp1;
  }
  return L;
}

// fr_operators.cpp:40-51 (derivative, closed form at the endpoints)
static double legendreDeriv(int n, double x) {
  if (n == 0) return 0.0;
  const double denom = x * x - 1.0;
  if (std::abs(denom) < 1e-14) {
    const double s = (x > 0.0) ? 1.0 : ((n - 1) % 2 == 0 ? 1.0 : -1.0);
    return s * 0.5 * n * (n + 1.0);
  }
  return n * (x * legendre(n, x) - legendre(n - 1, x)) / denom;
}

// fr_operators.cpp:59-81 (GLL nodes: Newton on L_P' from Chebyshev guesses)
static std::vector<double> gllNodes(int P) {
  std::vector<double> x(P + 1);
  if (P == 0) { x[0] = 0.0; return x; }
  x[0] = -1.0;
  x[P] = 1.0;
  for (int i = 1; i < P; ++i) {
    double xi = -std::cos(kPi * i / static_cast<double>(P));
    for (int it = 0; it < 100; ++it) {
      const double f = legendreDeriv(P, xi);
      const double fp = (2.0 * xi * legendreDeriv(P, xi) - P * (P + 1.0) * legendre(P, xi)) /
                        (1.0 - xi * xi);
      const double dx = f / fp;
      xi -= dx;
      if (std::abs(dx) < 1e-15) break;
    }
    x[i] = xi;
  }
  return x;
}

// fr_operators.cpp:84-99 (Gauss-Legendre nodes: Newton on L_{P+1})
static std::vector<double> gaussNodes(int P) {
  const int n = P + 1;
  std::vector<double> x(n);
  for (int i = 0; i < n; ++i) {
    double xi = -std::cos(kPi * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      const double f = legendre(n, xi);
      const double fp = legendreDeriv(n, xi);
      const double dx = f / fp;
      xi -= dx;
      if (std::abs(dx) < 1e-15) break;
    }
    x[i] = xi;
  }
  return x;
}

// fr_operators.cpp:103-113 (barycentric weights)
static std::vector<double> baryWeights(const std::vector<double>& x) {
  const int n = static_cast<int>(x.size());
  std::vector<double> w(n, 1.0);
  for (int j = 0; j < n; ++j) {
    double prod = 1.0;
    for (int k = 0; k < n; ++k)
      if (k != j) prod *= (x[j] - x[k]);
    w[j] = 1.0 / prod;
  }
  return w;
}

// fr_operators.cpp:118-127 (Lagrange basis value, exact delta on a node)
static double lagrangeAt(const std::vector<double>& x, const std::vector<double>& w, int j,
                         double xq) {
  const int n = static_cast<int>(x.size());
  for (int k = 0; k < n; ++k)
    if (std::abs(xq - x[k]) < 1e-14) return (k == j) ? 1.0 : 0.0;
  const double num = w[j] / (xq - x[j]);
  double den = 0.0;
  for (int k = 0; k < n; ++k) den += w[k] / (xq - x[k]);
  return num / den;
}

struct Ops {
  int P = 0;
  std::vector<double> x, D, lL, lR, gLp, gRp, qw;
};

// fr_operators.cpp:134-200 (constructor: steps 1-5)
static Ops buildOps(int P, bool gauss) {
  if (P < 0) throw std::invalid_argument("FR operators: P must be >= 0");
  Ops op;
  op.P = P;
  op.x = gauss ? gaussNodes(P) : gllNodes(P);
  const std::vector<double> w = baryWeights(op.x);
  const int n = P + 1;
  op.D.assign(static_cast<size_t>(n) * n, 0.0);
  for (int i = 0; i < n; ++i) {
    double diag = 0.0;
    for (int j = 0; j < n; ++j) {
      if (i == j) continue;
      const double dij = (w[j] / w[i]) / (op.x[i] - op.x[j]);
      op.D[static_cast<size_t>(i) * n + j] = dij;
      diag -= dij;
    }
    op.D[static_cast<size_t>(i) * n + i] = diag;
  }
  op.lL.assign(n, 0.0);
  op.lR.assign(n, 0.0);
  for (int j = 0; j < n; ++j) {
    op.lL[j] = lagrangeAt(op.x, w, j, -1.0);
    op.lR[j] = lagrangeAt(op.x, w, j, +1.0);
  }
  op.gLp.assign(n, 0.0);
  op.gRp.assign(n, 0.0);
  const double signP = (P % 2 == 0) ? 1.0 : -1.0;
  for (int i = 0; i < n; ++i) {
    const double dLP = legendreDeriv(P, op.x[i]);
    const double dLP1 = legendreDeriv(P + 1, op.x[i]);
    op.gLp[i] = 0.5 * signP * (dLP - dLP1);
    op.gRp[i] = 0.5 * (dLP + dLP1);
  }
  op.qw.assign(n, 0.0);
  if (P == 0) {
    op.qw[0] = 2.0;
  } else if (!gauss) {
    for (int i = 0; i < n; ++i) {
      const double LP = legendre(P, op.x[i]);
      op.qw[i] = 2.0 / (P * (P + 1.0) * LP * LP);
    }
  } else {
    for (int i = 0; i < n; ++i) {
      const double dLP1 = legendreDeriv(P + 1, op.x[i]);
      op.qw[i] = 2.0 / ((1.0 - op.x[i] * op.x[i]) * dLP1 * dLP1);
    }
  }
  return op;
}

// ---- the 2D solver: src/solver/sgn2d.cpp + mesh2d.h --------------------------

struct Solver {
  int P, np, npp, Nx, Ny, ne, Nd;
  double Lx, Ly, hx, hy, Jx, Jy;
  Par par;
  Ops op;
  int bc[4];
  std::vector<double> Hnod, Hx, Hy;   // nodal still depth and its gradients
  std::vector<double> HfX, HfY;       // interface still depths (see sgn_oracle.h)
  std::vector<double> inflow;         // left-side ghost coefficients
  double omega = 0.0;

  // gnode of node (a, b) of element e; sgn2d.h:22-27, mesh2d.h:31
  size_t uidx(int e, int a, int b, int var) const {
    return (size_t)var * Nd + (size_t)e * npp + a + np * b;
  }
};

// sgn2d.cpp:30-71 (mesh and bathymetry precompute; face depths supplied)
static Solver* makeSolver(const orc_desc& d) {
  if (d.Nx < 2 || d.Ny < 2) throw std::invalid_argument("SGN2D: Nx,Ny >= 2");
  for (int s = 0; s < 4; ++s)
    if (d.bc[s] < 0 || d.bc[s] > 2) throw std::invalid_argument("bad side code");
  if ((d.bc[0] == ORC_BC_PERIODIC) != (d.bc[1] == ORC_BC_PERIODIC) ||
      (d.bc[2] == ORC_BC_PERIODIC) != (d.bc[3] == ORC_BC_PERIODIC))
    throw std::invalid_argument("periodic sides must come in pairs");
  if (d.bc[1] == ORC_BC_INFLOW || d.bc[2] == ORC_BC_INFLOW || d.bc[3] == ORC_BC_INFLOW)
    throw std::invalid_argument("inflow is supported on the left side only");
  if (d.bc[0] == ORC_BC_INFLOW && !d.inflow) throw std::invalid_argument("inflow table missing");
  auto* s = new Solver();
  s->P = d.P; s->np = d.P + 1; s->npp = s->np * s->np;
  s->Nx = d.Nx; s->Ny = d.Ny; s->ne = d.Nx * d.Ny; s->Nd = s->ne * s->npp;
  s->Lx = d.Lx; s->Ly = d.Ly;
  s->hx = d.Lx / d.Nx; s->hy = d.Ly / d.Ny;          // mesh2d.h:29-30
  s->Jx = 0.5 * s->hx; s->Jy = 0.5 * s->hy;          // mesh2d.h:49-50
  s->par = Par{d.g, d.c, d.gamma, d.nManning};
  s->op = buildOps(d.P, d.gauss_points != 0);
  for (int k = 0; k < 4; ++k) s->bc[k] = d.bc[k];
  const int np = s->np, npp = s->npp, ne = s->ne, Nd = s->Nd;
  s->Hnod.assign(Nd, 0.0);
  s->Hx.assign(Nd, 0.0);
  s->Hy.assign(Nd, 0.0);
  s->HfX.assign((size_t)(s->Nx + 1) * s->Ny * np, 0.0);
  s->HfY.assign((size_t)s->Nx * (s->Ny + 1) * np, 0.0);
  if (d.bc[0] == ORC_BC_INFLOW) {
    s->inflow.assign(d.inflow, d.inflow + (size_t)s->Ny * np * NV * 3);
    s->omega = d.inflow_omega;
  }
  if (d.Hnod) {
    if (d.gauss_points) {
      delete s;
      throw std::invalid_argument("SGN2D: bathymetry requires GLL points");
    }
    std::copy(d.Hnod, d.Hnod + Nd, s->Hnod.begin());
    if (d.HfaceX) std::copy(d.HfaceX, d.HfaceX + s->HfX.size(), s->HfX.begin());
    if (d.HfaceY) std::copy(d.HfaceY, d.HfaceY + s->HfY.size(), s->HfY.begin());
    // sgn2d.cpp:54-71: element-local tensor differentiation of H
    const auto& D = s->op.D;
    for (int e = 0; e < ne; ++e) {
      const double* He = &s->Hnod[(size_t)e * npp];
      for (int b = 0; b < np; ++b)
        for (int a = 0; a < np; ++a) {
          double dHx = 0.0, dHy = 0.0;
          for (int k = 0; k < np; ++k) {
            dHx += D[(size_t)a * np + k] * He[k + np * b];
            dHy += D[(size_t)b * np + k] * He[a + np * k];
          }
          s->Hx[(size_t)e * npp + a + np * b] = dHx / s->Jx;
          s->Hy[(size_t)e * npp + a + np * b] = dHy / s->Jy;
        }
    }
  }
  return s;
}

// Ghost state across a non-periodic side (extension; see file header).
//   UIn: interior face state, side 0..3, line node t of row/column rc.
static void ghostState(const Solver& s, int side, int rc, int t, const double* UIn, double time,
                       double* G) {
  for (int v = 0; v < NV; ++v) G[v] = UIn[v];
  if (s.bc[side] == ORC_BC_WALL) {
    const int vn = (side < 2) ? IHU : IHV;
    G[vn] = -UIn[vn];
  } else {  // inflow (left side only)
    const double sn = std::sin(s.omega * time), cs = std::cos(s.omega * time);
    const double* c = &s.inflow[((size_t)rc * s.np + t) * NV * 3];
    for (int v = 0; v < NV; ++v) G[v] = (c[3 * v] + c[3 * v + 1] * sn) + c[3 * v + 2] * cs;
  }
}

// sgn2d.cpp:125-290 (rhs, five phases), with the side extension.
static void rhs(const Solver& s, const double* U, double* dUdt, double time) {
  const int np = s.np, npp = s.npp, Nx = s.Nx, Ny = s.Ny, ne = s.ne, Nd = s.Nd;
  const auto& D = s.op.D;
  const auto& lL = s.op.lL;
  const auto& lR = s.op.lR;
  const auto& gLp = s.op.gLp;
  const auto& gRp = s.op.gRp;
  const double Jx = s.Jx, Jy = s.Jy;
  const Par& par = s.par;
  std::fill(dUdt, dUdt + (size_t)NV * Nd, 0.0);

  // 1) face extrapolation (sgn2d.cpp:141-177): [(e*np + t)*NV + var]
  std::vector<double> xFL((size_t)ne * np * NV), xFR((size_t)ne * np * NV);
  std::vector<double> yFB((size_t)ne * np * NV), yFT((size_t)ne * np * NV);
  for (int e = 0; e < ne; ++e)
    for (int var = 0; var < NV; ++var) {
      for (int b = 0; b < np; ++b) {
        double sL = 0.0, sR = 0.0;
        for (int a = 0; a < np; ++a) {
          const double val = U[s.uidx(e, a, b, var)];
          sL += lL[a] * val; sR += lR[a] * val;
        }
        xFL[((size_t)e * np + b) * NV + var] = sL;
        xFR[((size_t)e * np + b) * NV + var] = sR;
      }
      for (int a = 0; a < np; ++a) {
        double sB = 0.0, sT = 0.0;
        for (int b = 0; b < np; ++b) {
          const double val = U[s.uidx(e, a, b, var)];
          sB += lL[b] * val; sT += lR[b] * val;
        }
        yFB[((size_t)e * np + a) * NV + var] = sB;
        yFT[((size_t)e * np + a) * NV + var] = sT;
      }
    }

  std::vector<double> Fnod((size_t)npp * NV), Gnod((size_t)npp * NV);
  std::vector<double> FxL(np * NV), FxR(np * NV), GyB(np * NV), GyT(np * NV);
  std::vector<double> hSxL(np), hSxR(np), hSyB(np), hSyT(np);
  std::vector<double> FdXL(np * NV), FdXR(np * NV), GdYB(np * NV), GdYT(np * NV);
  for (int e = 0; e < ne; ++e) {
    const int ex = e % Nx, ey = e / Nx;                  // mesh2d.h:32-33
    // periodic neighbours, mesh2d.h:36-45
    const int eL = ((ex - 1 + Nx) % Nx) + Nx * ey, eR = ((ex + 1) % Nx) + Nx * ey;
    const int eB = ex + Nx * ((ey - 1 + Ny) % Ny), eT = ex + Nx * ((ey + 1) % Ny);
    const bool leftSide = (ex == 0) && s.bc[0] != ORC_BC_PERIODIC;
    const bool rightSide = (ex == Nx - 1) && s.bc[1] != ORC_BC_PERIODIC;
    const bool botSide = (ey == 0) && s.bc[2] != ORC_BC_PERIODIC;
    const bool topSide = (ey == Ny - 1) && s.bc[3] != ORC_BC_PERIODIC;

    // 2) nodal fluxes (sgn2d.cpp:188-196)
    for (int b = 0; b < np; ++b)
      for (int a = 0; a < np; ++a) {
        double Un[NV];
        for (int var = 0; var < NV; ++var) Un[var] = U[s.uidx(e, a, b, var)];
        const double Hl = s.Hnod[(size_t)e * npp + a + np * b];
        fluxX(Un, Hl, par, &Fnod[(a + np * b) * NV]);
        fluxY(Un, Hl, par, &Gnod[(a + np * b) * NV]);
      }

    // 3) Rusanov fluxes and central depths on the four face lines (sgn2d.cpp:198-227)
    for (int b = 0; b < np; ++b) {
      const double* uMe = &xFL[((size_t)e * np + b) * NV];
      const double* uMeR = &xFR[((size_t)e * np + b) * NV];
      const double* uL = &xFR[((size_t)eL * np + b) * NV];
      const double* uR = &xFL[((size_t)eR * np + b) * NV];
      double gL[NV], gR[NV];
      if (leftSide) { ghostState(s, 0, ey, b, uMe, time, gL); uL = gL; }
      if (rightSide) { ghostState(s, 1, ey, b, uMeR, time, gR); uR = gR; }
      const double HxL = s.HfX[((size_t)ey * (Nx + 1) + ex) * np + b];
      const double HxR = s.HfX[((size_t)ey * (Nx + 1) + ex + 1) * np + b];
      rusanov(uL, uMe, 0, HxL, par, &FxL[b * NV]);
      rusanov(uMeR, uR, 0, HxR, par, &FxR[b * NV]);
      hSxL[b] = 0.5 * (uL[IH] + uMe[IH]);
      hSxR[b] = 0.5 * (uMeR[IH] + uR[IH]);
    }
    for (int a = 0; a < np; ++a) {
      const double* uMe = &yFB[((size_t)e * np + a) * NV];
      const double* uMeT = &yFT[((size_t)e * np + a) * NV];
      const double* uB = &yFT[((size_t)eB * np + a) * NV];
      const double* uT = &yFB[((size_t)eT * np + a) * NV];
      double gB[NV], gT[NV];
      if (botSide) { ghostState(s, 2, ex, a, uMe, time, gB); uB = gB; }
      if (topSide) { ghostState(s, 3, ex, a, uMeT, time, gT); uT = gT; }
      const double HyB = s.HfY[((size_t)ey * Nx + ex) * np + a];
      const double HyT = s.HfY[((size_t)(ey + 1) * Nx + ex) * np + a];
      rusanov(uB, uMe, 1, HyB, par, &GyB[a * NV]);
      rusanov(uMeT, uT, 1, HyT, par, &GyT[a * NV]);
      hSyB[a] = 0.5 * (uB[IH] + uMe[IH]);
      hSyT[a] = 0.5 * (uMeT[IH] + uT[IH]);
    }

    // 4) the element's own face fluxes (sgn2d.cpp:229-239)
    for (int b = 0; b < np; ++b) {
      fluxX(&xFL[((size_t)e * np + b) * NV], s.HfX[((size_t)ey * (Nx + 1) + ex) * np + b], par,
            &FdXL[b * NV]);
      fluxX(&xFR[((size_t)e * np + b) * NV], s.HfX[((size_t)ey * (Nx + 1) + ex + 1) * np + b],
            par, &FdXR[b * NV]);
    }
    for (int a = 0; a < np; ++a) {
      fluxY(&yFB[((size_t)e * np + a) * NV], s.HfY[((size_t)ey * Nx + ex) * np + a], par,
            &GdYB[a * NV]);
      fluxY(&yFT[((size_t)e * np + a) * NV], s.HfY[((size_t)(ey + 1) * Nx + ex) * np + a], par,
            &GdYT[a * NV]);
    }

    // 5) per-node assembly (sgn2d.cpp:241-288)
    for (int b = 0; b < np; ++b)
      for (int a = 0; a < np; ++a) {
        const int loc = a + np * b;
        double Un[NV];
        for (int var = 0; var < NV; ++var) Un[var] = U[s.uidx(e, a, b, var)];
        double dF[NV], dG[NV];
        for (int var = 0; var < NV; ++var) {
          double dfx = 0.0, dgy = 0.0;
          for (int k = 0; k < np; ++k) {
            dfx += D[(size_t)a * np + k] * Fnod[(k + np * b) * NV + var];
            dgy += D[(size_t)b * np + k] * Gnod[(a + np * k) * NV + var];
          }
          const double cx = (FxL[b * NV + var] - FdXL[b * NV + var]) * gLp[a] +
                            (FxR[b * NV + var] - FdXR[b * NV + var]) * gRp[a];
          const double cy = (GyB[a * NV + var] - GdYB[a * NV + var]) * gLp[b] +
                            (GyT[a * NV + var] - GdYT[a * NV + var]) * gRp[b];
          dF[var] = (dfx + cx) / Jx;
          dG[var] = (dgy + cy) / Jy;
        }
        double dhx = 0.0, dhy = 0.0;
        for (int k = 0; k < np; ++k) {
          dhx += D[(size_t)a * np + k] * U[s.uidx(e, k, b, IH)];
          dhy += D[(size_t)b * np + k] * U[s.uidx(e, a, k, IH)];
        }
        dhx += (hSxL[b] - xFL[((size_t)e * np + b) * NV + IH]) * gLp[a] +
               (hSxR[b] - xFR[((size_t)e * np + b) * NV + IH]) * gRp[a];
        dhy += (hSyB[a] - yFB[((size_t)e * np + a) * NV + IH]) * gLp[b] +
               (hSyT[a] - yFT[((size_t)e * np + a) * NV + IH]) * gRp[b];
        const double hx = dhx / Jx, hy = dhy / Jy;
        double Bx[NV], By[NV], S[NV];
        ncpX(Un, hx, par, Bx);
        ncpY(Un, hy, par, By);
        source(Un, s.Hnod[(size_t)e * npp + loc], s.Hx[(size_t)e * npp + loc],
               s.Hy[(size_t)e * npp + loc], par, S);
        for (int var = 0; var < NV; ++var)
          dUdt[s.uidx(e, a, b, var)] = -dF[var] - dG[var] - Bx[var] - By[var] + S[var];
      }
  }
}

// sgn2d.cpp:295-304 (maxWaveSpeed)
static double maxWaveSpeedState(const Solver& s, const double* U) {
  double lam = 0.0;
  for (int i = 0; i < s.Nd; ++i) {
    double Un[NV];
    for (int v = 0; v < NV; ++v) Un[v] = U[(size_t)v * s.Nd + i];
    lam = std::fmax(lam, std::fmax(waveSpeed(Un, 0, s.par), waveSpeed(Un, 1, s.par)));
  }
  return lam;
}

// sgn2d.cpp:310-317 (cflTimeStep)
static double cflTimeStep(const Solver& s, const double* U, double cfl) {
  const double lam = maxWaveSpeedState(s, U);
  const double dxEff = std::fmin(s.hx, s.hy) / (2.0 * s.P + 1.0);
  return cfl * dxEff / (lam > 0.0 ? lam : 1e-300);
}

// sgn2d.cpp:319-332 (mass)
static double mass(const Solver& s, const double* U) {
  const auto& qw = s.op.qw;
  const double JJ = s.Jx * s.Jy;
  double m = 0.0;
  for (int e = 0; e < s.ne; ++e)
    for (int b = 0; b < s.np; ++b)
      for (int a = 0; a < s.np; ++a)
        m += qw[a] * qw[b] * U[s.uidx(e, a, b, IH)];
  return m * JJ;
}

// sgn2d.cpp:334-350 (energy)
static double energy(const Solver& s, const double* U) {
  const auto& qw = s.op.qw;
  const double JJ = s.Jx * s.Jy;
  double E = 0.0;
  for (int e = 0; e < s.ne; ++e)
    for (int b = 0; b < s.np; ++b)
      for (int a = 0; a < s.np; ++a) {
        const int loc = a + s.np * b;
        double Un[NV];
        for (int v = 0; v < NV; ++v) Un[v] = U[s.uidx(e, a, b, v)];
        E += qw[a] * qw[b] * energyDensity(Un, s.Hnod[(size_t)e * s.npp + loc], s.par);
      }
  return E * JJ;
}

// ---- integrators: src/solver/rk.cpp -------------------------------------------

using State = std::vector<double>;
using RhsFn = std::function<void(double, const State&, State&)>;

// rk.cpp:23-38 (rk4Step)
static void rk4Step(const RhsFn& f, double t, double dt, State& u) {
  const size_t n = u.size();
  State k1(n), k2(n), k3(n), k4(n), tmp(n);
  f(t, u, k1);
  for (size_t i = 0; i < n; ++i) tmp[i] = u[i] + 0.5 * dt * k1[i];
  f(t + 0.5 * dt, tmp, k2);
  for (size_t i = 0; i < n; ++i) tmp[i] = u[i] + 0.5 * dt * k2[i];
  f(t + 0.5 * dt, tmp, k3);
  for (size_t i = 0; i < n; ++i) tmp[i] = u[i] + dt * k3[i];
  f(t + dt, tmp, k4);
  for (size_t i = 0; i < n; ++i)
    u[i] += (dt / 6.0) * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
}

// rk.cpp:41-54 (bs3Step)
static void bs3Step(const RhsFn& f, double t, double dt, State& u) {
  const size_t n = u.size();
  State k1(n), k2(n), k3(n), tmp(n);
  f(t, u, k1);
  for (size_t i = 0; i < n; ++i) tmp[i] = u[i] + 0.5 * dt * k1[i];
  f(t + 0.5 * dt, tmp, k2);
  for (size_t i = 0; i < n; ++i) tmp[i] = u[i] + 0.75 * dt * k2[i];
  f(t + 0.75 * dt, tmp, k3);
  for (size_t i = 0; i < n; ++i)
    u[i] += dt * (2.0 / 9.0 * k1[i] + 1.0 / 3.0 * k2[i] + 4.0 / 9.0 * k3[i]);
}

// FNV-1a 64 of the bytes of a double array (little-endian doubles).
static unsigned long long fnv64(const double* v, size_t n) {
  unsigned long long h = 1469598103934665603ull;
  const unsigned char* p = reinterpret_cast<const unsigned char*>(v);
  for (size_t i = 0; i < n * sizeof(double); ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

struct CallLog {
  double* t = nullptr;
  unsigned long long* h = nullptr;
  long cap = 0, n = 0;
  void add(double tt, const State& dx) {
    if (n < cap) { t[n] = tt; h[n] = fnv64(dx.data(), dx.size()); }
    ++n;
  }
};

// rk.cpp:67-164 (integrateAdaptive), plus the max_accepted stop and a log of
// (dt, errNorm) per attempted step for the tests.
static orc_rk_stats integrateAdaptive(const RhsFn& f0, State& u, double t0, double tEnd,
                                      double dt0, const orc_rk_opts& opt, double* dtLog,
                                      long dtCap, CallLog* log = nullptr) {
  // every RHS call optionally logs (t, FNV-1a digest of its output)
  const RhsFn f = [&](double t, const State& x, State& dx) {
    f0(t, x, dx);
    if (log) log->add(t, dx);
  };
  const int orderLow = 2;
  const double expo = 1.0 / (orderLow + 1.0);
  const double alpha = 0.7 * expo;
  const double beta = 0.4 * expo;
  orc_rk_stats st{};
  st.tReached = t0;
  const size_t n = u.size();
  if (n == 0) return st;
  State k1(n), k2(n), k3(n), k4(n), tmp(n), uNew(n);
  double t = t0;
  double dt = std::min(dt0, opt.dtMax);
  double errPrev = 1.0;
  f(t, u, k1);
  ++st.rhsEvals;
  long attempt = 0;
  while (t < tEnd - 1e-15 * std::max(1.0, std::abs(tEnd))) {
    if (opt.max_accepted > 0 && st.accepted >= opt.max_accepted) break;
    if (t + dt > tEnd) dt = tEnd - t;
    for (size_t i = 0; i < n; ++i) tmp[i] = u[i] + 0.5 * dt * k1[i];
    f(t + 0.5 * dt, tmp, k2);
    for (size_t i = 0; i < n; ++i) tmp[i] = u[i] + 0.75 * dt * k2[i];
    f(t + 0.75 * dt, tmp, k3);
    for (size_t i = 0; i < n; ++i)
      uNew[i] = u[i] + dt * (2.0 / 9.0 * k1[i] + 1.0 / 3.0 * k2[i] + 4.0 / 9.0 * k3[i]);
    f(t + dt, uNew, k4);
    st.rhsEvals += 3;
    double errL2 = 0.0, errMax = 0.0;
    for (size_t i = 0; i < n; ++i) {
      const double uEmb =
          u[i] + dt * (7.0 / 24.0 * k1[i] + 0.25 * k2[i] + 1.0 / 3.0 * k3[i] + 0.125 * k4[i]);
      const double e = uNew[i] - uEmb;
      const double sc = opt.atol + opt.rtol * std::max(std::abs(u[i]), std::abs(uNew[i]));
      const double r = e / sc;
      errL2 += r * r;
      errMax = std::max(errMax, std::abs(r));
    }
    errL2 = std::sqrt(errL2 / static_cast<double>(n));
    const double errNorm = std::max(errL2, 1e-300);
    if (dtLog && attempt < dtCap) { dtLog[2 * attempt] = dt; dtLog[2 * attempt + 1] = errNorm; }
    ++attempt;
    if (errNorm <= 1.0) {
      t += dt;
      u.swap(uNew);
      std::swap(k1, k4);
      ++st.accepted;
      st.tReached = t;
      st.dtLast = dt;
      st.maxNormLast = errMax;
      errPrev = errNorm;
    } else {
      ++st.rejected;
    }
    double fac = opt.safety * std::pow(errNorm, -alpha) * std::pow(errPrev, beta);
    fac = std::clamp(fac, opt.facMin, opt.facMax);
    double dtNext = dt * fac;
    dtNext = std::min(dtNext, opt.dtMax);
    if (dtNext < opt.dtMin) {
      st.aborted = 1;
      st.dtLast = dtNext;
      break;
    }
    dt = dtNext;
  }
  return st;
}

// Test ODEs of the reference RK tests (test/test_rk.cpp:39-78).
static RhsFn odeFn(int kind, double param) {
  if (kind == 0) return [](double, const State& u, State& du) { du[0] = -u[0]; };
  if (kind == 1) return [](double, const State& u, State& du) { du[0] = u[1]; du[1] = -u[0]; };
  return [param](double, const State& u, State& du) { du[0] = param * u[0]; };
}

}  // namespace orc

// ---- C-ABI ------------------------------------------------------------------

struct orc_solver { orc::Solver* s; };

extern "C" {

int orc_fr_operators(int P, int gauss_points, double* x, double* D, double* lL, double* lR,
                     double* gLp, double* gRp, double* qw) {
  try {
    const orc::Ops op = orc::buildOps(P, gauss_points != 0);
    const int n = P + 1;
    std::copy(op.x.begin(), op.x.end(), x);
    std::copy(op.D.begin(), op.D.end(), D);
    std::copy(op.lL.begin(), op.lL.end(), lL);
    std::copy(op.lR.begin(), op.lR.end(), lR);
    std::copy(op.gLp.begin(), op.gLp.end(), gLp);
    std::copy(op.gRp.begin(), op.gRp.end(), gRp);
    std::copy(op.qw.begin(), op.qw.begin() + n, qw);
    return 0;
  } catch (...) {
    return 1;
  }
}
double orc_legendre(int n, double x) { return orc::legendre(n, x); }
double orc_legendre_deriv(int n, double x) { return orc::legendreDeriv(n, x); }

void orc_flux_x(const double* U, double H, const double* p4, double* F) {
  orc::fluxX(U, H, orc::parFrom(p4), F);
}
void orc_flux_y(const double* U, double H, const double* p4, double* G) {
  orc::fluxY(U, H, orc::parFrom(p4), G);
}
void orc_source(const double* U, double H, double Hx, double Hy, const double* p4, double* S) {
  orc::source(U, H, Hx, Hy, orc::parFrom(p4), S);
}
void orc_ncp_x(const double* U, double hx, const double* p4, double* B) {
  orc::ncpX(U, hx, orc::parFrom(p4), B);
}
void orc_ncp_y(const double* U, double hy, const double* p4, double* B) {
  orc::ncpY(U, hy, orc::parFrom(p4), B);
}
double orc_signal_speed(const double* U, const double* p4) {
  return orc::signalSpeed(U, orc::parFrom(p4));
}
double orc_max_wave_speed(const double* U, int dir, const double* p4) {
  return orc::waveSpeed(U, dir, orc::parFrom(p4));
}
void orc_rusanov(const double* UL, const double* UR, int dir, double Hf, const double* p4,
                 double* Fn) {
  orc::rusanov(UL, UR, dir, Hf, orc::parFrom(p4), Fn);
}
double orc_energy_density(const double* U, double H, const double* p4) {
  return orc::energyDensity(U, H, orc::parFrom(p4));
}

// sgn2d.cpp:89-102 (nodeCoords)
void orc_node_coords(int P, int Nx, int Ny, double Lx, double Ly, double* x, double* y) {
  const orc::Ops op = orc::buildOps(P, false);
  const int np = P + 1, npp = np * np;
  const double hx = Lx / Nx, hy = Ly / Ny;
  for (int e = 0; e < Nx * Ny; ++e) {
    const int ex = e % Nx, ey = e / Nx;
    for (int b = 0; b < np; ++b)
      for (int a = 0; a < np; ++a) {
        const size_t g = (size_t)e * npp + a + np * b;
        x[g] = ex * hx + 0.5 * hx * (op.x[a] + 1.0);
        y[g] = ey * hy + 0.5 * hy * (op.x[b] + 1.0);
      }
  }
}

// sgn2d.cpp:75-83 (face still-depth sample points), on every interface line.
void orc_face_coords(int P, int Nx, int Ny, double Lx, double Ly, int wrapx, int wrapy,
                     double* fxx, double* fxy, double* fyx, double* fyy) {
  const orc::Ops op = orc::buildOps(P, false);
  const int np = P + 1;
  const double hx = Lx / Nx, hy = Ly / Ny;
  for (int ey = 0; ey < Ny; ++ey)
    for (int i = 0; i <= Nx; ++i) {
      const int ic = (wrapx && i == Nx) ? 0 : i;
      for (int t = 0; t < np; ++t) {
        const size_t k = ((size_t)ey * (Nx + 1) + i) * np + t;
        fxx[k] = ic * hx;
        fxy[k] = ey * hy + 0.5 * hy * (op.x[t] + 1.0);
      }
    }
  for (int j = 0; j <= Ny; ++j) {
    const int jc = (wrapy && j == Ny) ? 0 : j;
    for (int ex = 0; ex < Nx; ++ex)
      for (int t = 0; t < np; ++t) {
        const size_t k = ((size_t)j * Nx + ex) * np + t;
        fyx[k] = ex * hx + 0.5 * hx * (op.x[t] + 1.0);
        fyy[k] = jc * hy;
      }
  }
}

orc_solver* orc_create(const orc_desc* d, char* err, size_t errlen) {
  try {
    auto* h = new orc_solver{orc::makeSolver(*d)};
    return h;
  } catch (const std::exception& ex) {
    if (err && errlen) std::snprintf(err, errlen, "%s", ex.what());
    return nullptr;
  }
}
void orc_destroy(orc_solver* s) {
  if (s) { delete s->s; delete s; }
}
long orc_ndof(const orc_solver* s) { return s->s->Nd; }
void orc_bathymetry(const orc_solver* s, double* Hnod, double* Hx, double* Hy) {
  const auto& S = *s->s;
  std::copy(S.Hnod.begin(), S.Hnod.end(), Hnod);
  std::copy(S.Hx.begin(), S.Hx.end(), Hx);
  std::copy(S.Hy.begin(), S.Hy.end(), Hy);
}
int orc_rhs(const orc_solver* s, const double* U, double* dU, double t) {
  orc::rhs(*s->s, U, dU, t);
  return 0;
}
double orc_max_wave_speed_state(const orc_solver* s, const double* U) {
  return orc::maxWaveSpeedState(*s->s, U);
}
double orc_cfl_time_step(const orc_solver* s, const double* U, double cfl) {
  return orc::cflTimeStep(*s->s, U, cfl);
}
double orc_mass(const orc_solver* s, const double* U) { return orc::mass(*s->s, U); }
double orc_energy(const orc_solver* s, const double* U) { return orc::energy(*s->s, U); }

static orc::RhsFn solverFn(const orc_solver* s) {
  const orc::Solver* S = s->s;
  return [S](double t, const orc::State& u, orc::State& du) {
    du.resize(u.size());
    orc::rhs(*S, u.data(), du.data(), t);
  };
}

void orc_rk4_step(const orc_solver* s, double t, double dt, double* U) {
  const size_t n = (size_t)orc::NV * s->s->Nd;
  orc::State u(U, U + n);
  orc::rk4Step(solverFn(s), t, dt, u);
  std::copy(u.begin(), u.end(), U);
}
void orc_bs3_step(const orc_solver* s, double t, double dt, double* U) {
  const size_t n = (size_t)orc::NV * s->s->Nd;
  orc::State u(U, U + n);
  orc::bs3Step(solverFn(s), t, dt, u);
  std::copy(u.begin(), u.end(), U);
}
int orc_integrate(const orc_solver* s, double* U, double t0, double tEnd, double dt0,
                  const orc_rk_opts* o, orc_rk_stats* st, double* dt_log, long dt_cap,
                  double* call_t, unsigned long long* call_h, long call_cap) {
  const size_t n = (size_t)orc::NV * s->s->Nd;
  orc::State u(U, U + n);
  orc::CallLog log;
  log.t = call_t; log.h = call_h; log.cap = call_t && call_h ? call_cap : 0;
  *st = orc::integrateAdaptive(solverFn(s), u, t0, tEnd, dt0, *o, dt_log, dt_cap, &log);
  std::copy(u.begin(), u.end(), U);
  return 0;
}
int orc_integrate_ode(int kind, double param, double* u, int n, double t0, double tEnd,
                      double dt0, const orc_rk_opts* o, orc_rk_stats* st) {
  orc::State v(u, u + n);
  *st = orc::integrateAdaptive(orc::odeFn(kind, param), v, t0, tEnd, dt0, *o, nullptr, 0);
  std::copy(v.begin(), v.end(), u);
  return 0;
}
void orc_rk4_step_ode(int kind, double param, double* u, int n, double t, double dt) {
  orc::State v(u, u + n);
  orc::rk4Step(orc::odeFn(kind, param), t, dt, v);
  std::copy(v.begin(), v.end(), u);
}
unsigned long long orc_fnv64(const double* v, size_t n) { return orc::fnv64(v, n); }

void orc_bs3_step_ode(int kind, double param, double* u, int n, double t, double dt) {
  orc::State v(u, u + n);
  orc::bs3Step(orc::odeFn(kind, param), t, dt, v);
  std::copy(v.begin(), v.end(), u);
}

}  // extern "C"
