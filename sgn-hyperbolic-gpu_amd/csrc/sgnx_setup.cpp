// =============================================================================
// sgnx_setup.cpp -- host-side setup: 1D FR operator tables on Gauss-Lobatto
// points and the node / face sample coordinates of the 2D mesh.
//
// The kernels are bitwise against the reference only when they are fed the
// reference's exact table bits, so the tables are built by the same
// mathematical steps as src/core/fr_operators.cpp:134-200 (Newton iteration
// for the interior GLL nodes from Chebyshev-Lobatto guesses, barycentric
// weights, the Berrut-Trefethen differentiation matrix with the negative-sum
// diagonal, Huynh's g_DG correction derivatives, GLL weights). Coordinates
// follow src/solver/sgn2d.cpp:75-102 term by term.
// =============================================================================

#include <cmath>
#include <vector>

#include "../../include/sgnx.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

// Legendre L_n(x) by the three-term recurrence (n+1) L_{n+1} = (2n+1) x L_n - n L_{n-1}.
double leg(int n, double x) {
  if (n == 0) return 1.0;
  if (n == 1) return x;
  double prev = 1.0, cur = x;
  for (int k = 1; k < n; ++k) {
    const double next = ((2.0 * k + 1.0) * x * cur - k * prev) / (k + 1.0);
    prev = cur;
    cur = next;
  }
  return cur;
}

// L_n'(x) = n (x L_n - L_{n-1}) / (x^2 - 1), with the closed form
// L_n'(+-1) = (+-1)^(n-1) n (n+1) / 2 where the quotient degenerates.
double legd(int n, double x) {
  if (n == 0) return 0.0;
  const double den = x * x - 1.0;
  if (std::abs(den) < 1e-14) {
    const double sgn = (x > 0.0) ? 1.0 : ((n - 1) % 2 == 0 ? 1.0 : -1.0);
    return sgn * 0.5 * n * (n + 1.0);
  }
  return n * (x * leg(n, x) - leg(n - 1, x)) / den;
}

struct GllTables {
  std::vector<double> x, D, lL, lR, gLp, gRp, qw;
};

GllTables gll(int P) {
  GllTables T;
  const int n = P + 1;
  T.x.assign(n, 0.0);
  if (P > 0) {
    T.x[0] = -1.0;
    T.x[P] = 1.0;
    for (int i = 1; i < P; ++i) {
      // interior nodes: roots of L_P', Newton with L_P'' from the Legendre ODE
      double r = -std::cos(kPi * i / static_cast<double>(P));
      for (int it = 0; it < 100; ++it) {
        const double f = legd(P, r);
        const double fp = (2.0 * r * legd(P, r) - P * (P + 1.0) * leg(P, r)) / (1.0 - r * r);
        const double step = f / fp;
        r -= step;
        if (std::abs(step) < 1e-15) break;
      }
      T.x[i] = r;
    }
  }
  // barycentric weights 1 / prod_{k != j} (x_j - x_k)
  std::vector<double> w(n, 1.0);
  for (int j = 0; j < n; ++j) {
    double prod = 1.0;
    for (int k = 0; k < n; ++k)
      if (k != j) prod *= (T.x[j] - T.x[k]);
    w[j] = 1.0 / prod;
  }
  // D_ij = (w_j / w_i) / (x_i - x_j), D_ii = -sum_{j != i} D_ij
  T.D.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) {
    double diag = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      const double d = (w[j] / w[i]) / (T.x[i] - T.x[j]);
      T.D[(size_t)i * n + j] = d;
      diag -= d;
    }
    T.D[(size_t)i * n + i] = diag;
  }
  // value of the Lagrange basis at the faces: a Kronecker delta on GLL points
  // (the barycentric quotient is never needed because -1 and +1 are nodes)
  T.lL.assign(n, 0.0);
  T.lR.assign(n, 0.0);
  auto lag = [&](int j, double xq) {
    for (int k = 0; k < n; ++k)
      if (std::abs(xq - T.x[k]) < 1e-14) return (k == j) ? 1.0 : 0.0;
    double den = 0.0;
    for (int k = 0; k < n; ++k) den += w[k] / (xq - T.x[k]);
    return (w[j] / (xq - T.x[j])) / den;
  };
  for (int j = 0; j < n; ++j) {
    T.lL[j] = lag(j, -1.0);
    T.lR[j] = lag(j, +1.0);
  }
  // Huynh g_DG: gL = (-1)^P/2 (L_P - L_{P+1}), gR = 1/2 (L_P + L_{P+1})
  T.gLp.assign(n, 0.0);
  T.gRp.assign(n, 0.0);
  const double sP = (P % 2 == 0) ? 1.0 : -1.0;
  for (int i = 0; i < n; ++i) {
    const double a = legd(P, T.x[i]), b = legd(P + 1, T.x[i]);
    T.gLp[i] = 0.5 * sP * (a - b);
    T.gRp[i] = 0.5 * (a + b);
  }
  // GLL weights 2 / (P (P+1) L_P(x_i)^2)
  T.qw.assign(n, 0.0);
  if (P == 0) {
    T.qw[0] = 2.0;
  } else {
    for (int i = 0; i < n; ++i) {
      const double L = leg(P, T.x[i]);
      T.qw[i] = 2.0 / (P * (P + 1.0) * L * L);
    }
  }
  return T;
}

template <class V>
void put(const V& v, double* out) {
  if (!out) return;
  for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
}

}  // namespace

extern "C" {

int sgnx_fr_operators(int P, double* x, double* D, double* lL, double* lR, double* gLp,
                      double* gRp, double* qw) {
  if (P < 0 || P > 16) return SGNX_EINVAL;
  const GllTables T = gll(P);
  put(T.x, x);
  put(T.D, D);
  put(T.lL, lL);
  put(T.lR, lR);
  put(T.gLp, gLp);
  put(T.gRp, gRp);
  put(T.qw, qw);
  return SGNX_OK;
}

int sgnx_node_coords(int P, int Nx, int Ny, double Lx, double Ly, double* x, double* y) {
  if (P < 0 || P > 16 || Nx < 1 || Ny < 1 || !x || !y) return SGNX_EINVAL;
  const GllTables T = gll(P);
  const int np = P + 1, npp = np * np;
  const double hx = Lx / Nx, hy = Ly / Ny;
  for (long long e = 0; e < (long long)Nx * Ny; ++e) {
    const int ex = (int)(e % Nx), ey = (int)(e / Nx);
    for (int b = 0; b < np; ++b)
      for (int a = 0; a < np; ++a) {
        const long long k = e * npp + a + np * b;
        x[k] = ex * hx + 0.5 * hx * (T.x[a] + 1.0);
        y[k] = ey * hy + 0.5 * hy * (T.x[b] + 1.0);
      }
  }
  return SGNX_OK;
}

int sgnx_face_coords(int P, int Nx, int Ny, double Lx, double Ly, int wrapx, int wrapy,
                     double* fxx, double* fxy, double* fyx, double* fyy) {
  if (P < 0 || P > 16 || Nx < 1 || Ny < 1 || !fxx || !fxy || !fyx || !fyy) return SGNX_EINVAL;
  const GllTables T = gll(P);
  const int np = P + 1;
  const double hx = Lx / Nx, hy = Ly / Ny;
  for (int ey = 0; ey < Ny; ++ey)
    for (int i = 0; i <= Nx; ++i) {
      const int col = (wrapx && i == Nx) ? 0 : i;
      for (int t = 0; t < np; ++t) {
        const long long k = ((long long)ey * (Nx + 1) + i) * np + t;
        fxx[k] = col * hx;
        fxy[k] = ey * hy + 0.5 * hy * (T.x[t] + 1.0);
      }
    }
  for (int j = 0; j <= Ny; ++j) {
    const int row = (wrapy && j == Ny) ? 0 : j;
    for (int ex = 0; ex < Nx; ++ex)
      for (int t = 0; t < np; ++t) {
        const long long k = ((long long)j * Nx + ex) * np + t;
        fyx[k] = ex * hx + 0.5 * hx * (T.x[t] + 1.0);
        fyy[k] = row * hy;
      }
  }
  return SGNX_OK;
}

}  // extern "C"
