/*
 * sgn_oracle.h -- C-ABI of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * The oracle is a plain C++17 restatement of the reference CPU solver's 2D
 * flux-reconstruction right-hand side and its Runge-Kutta integrators,
 * operation for operation, built with -ffp-contract=off so that its bits are
 * those of the reference CPU build. Only the tests, the smoke check and the
 * CPU-baseline leg of the benchmark may load it; the product never does.
 *
 * Reference files restated (paths relative to the reference checkout):
 *   src/physics/sgn_flux2d.h      pointwise physics
 *   src/core/fr_operators.cpp     GLL / Gauss nodes, D, lL/lR, gL'/gR', weights
 *   src/core/mesh2d.h             element numbering, neighbours, Jacobians
 *   src/solver/sgn2d.cpp          bathymetry precompute, rhs(), CFL, mass, energy
 *   src/solver/rk.cpp             rk4Step, bs3Step, integrateAdaptive
 * plus one extension the 2D reference lacks: wall and inflow sides (the
 * reference 2D mesh is periodic on all four sides).
 */
#ifndef SGN_ORACLE_H
#define SGN_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Side codes and side order (left, right, bottom, top). */
#define ORC_BC_PERIODIC 0
#define ORC_BC_WALL 1
#define ORC_BC_INFLOW 2

typedef struct orc_desc {
  int P, Nx, Ny;
  int gauss_points;          /* 0 = Gauss-Lobatto (production), 1 = Gauss-Legendre */
  double Lx, Ly;
  double g, c, gamma, nManning;
  int bc[4];                 /* left, right, bottom, top */
  /* Still depth: nodal (Nd) and on every interface line.
   *   HfaceX[(ey*(Nx+1) + i)*np + t]  x-interface i = 0..Nx (left edge of column i)
   *   HfaceY[(j*Nx + ex)*np + t]      y-interface j = 0..Ny (bottom edge of row j)
   * All three NULL = flat bottom H == 0 (the reference's empty stillDepth). */
  const double* Hnod;
  const double* HfaceX;
  const double* HfaceY;
  /* Inflow ghost on the left side: per (ey, t) node, NVAR x {base, sin, cos}
   * coefficients, ghost_v = (base + s_coef*sin(w t)) + c_coef*cos(w t). */
  const double* inflow;
  double inflow_omega;
} orc_desc;

typedef struct orc_rk_opts {
  double atol, rtol, dtMin, dtMax, safety, facMin, facMax;
  long max_accepted;         /* extension: stop after this many accepted steps (0 = off) */
} orc_rk_opts;

typedef struct orc_rk_stats {
  long accepted, rejected, rhsEvals;
  double tReached, dtLast, maxNormLast;
  int aborted;
} orc_rk_stats;

typedef struct orc_solver orc_solver;

/* FR tables of degree P: x, w_bary (P+1), D ((P+1)^2 row-major), lL, lR, gLp, gRp, qw. */
int orc_fr_operators(int P, int gauss_points, double* x, double* D, double* lL,
                     double* lR, double* gLp, double* gRp, double* qw);
double orc_legendre(int n, double x);
double orc_legendre_deriv(int n, double x);

/* Pointwise physics (sgn_flux2d.h). Arrays have NVAR = 5 entries. */
void orc_flux_x(const double* U, double H, const double* par4, double* F);
void orc_flux_y(const double* U, double H, const double* par4, double* G);
void orc_source(const double* U, double H, double Hx, double Hy, const double* par4, double* S);
void orc_ncp_x(const double* U, double hx, const double* par4, double* B);
void orc_ncp_y(const double* U, double hy, const double* par4, double* B);
double orc_signal_speed(const double* U, const double* par4);
double orc_max_wave_speed(const double* U, int dir, const double* par4);
void orc_rusanov(const double* UL, const double* UR, int dir, double Hf, const double* par4, double* Fn);
double orc_energy_density(const double* U, double H, const double* par4);

/* Geometry helpers with the reference's coordinate formulas. */
void orc_node_coords(int P, int Nx, int Ny, double Lx, double Ly, double* x, double* y);
/* Coordinates of the interface nodes of HfaceX / HfaceY (layouts above).  With
 * wrap != 0 the last interface line uses the periodic (wrapped) coordinate 0,
 * as the reference does for its right / top edges. */
void orc_face_coords(int P, int Nx, int Ny, double Lx, double Ly, int wrapx, int wrapy,
                     double* fxx, double* fxy, double* fyx, double* fyy);

orc_solver* orc_create(const orc_desc* d, char* err, size_t errlen);
void orc_destroy(orc_solver* s);
long orc_ndof(const orc_solver* s);
void orc_bathymetry(const orc_solver* s, double* Hnod, double* Hx, double* Hy);
int orc_rhs(const orc_solver* s, const double* U, double* dU, double t);
double orc_max_wave_speed_state(const orc_solver* s, const double* U);
double orc_cfl_time_step(const orc_solver* s, const double* U, double cfl);
double orc_mass(const orc_solver* s, const double* U);
double orc_energy(const orc_solver* s, const double* U);
void orc_rk4_step(const orc_solver* s, double t, double dt, double* U);
void orc_bs3_step(const orc_solver* s, double t, double dt, double* U);
/* Adaptive BS3(2) on the solver's RHS. dt_log (optional, capacity dt_cap)
 * receives every attempted dt followed by its error norm (2 doubles per step);
 * call_t / call_h (optional, capacity call_cap) receive the time argument and
 * the FNV-1a 64 digest of the output of every RHS evaluation. */
int orc_integrate(const orc_solver* s, double* U, double t0, double tEnd, double dt0,
                  const orc_rk_opts* o, orc_rk_stats* st, double* dt_log, long dt_cap,
                  double* call_t, unsigned long long* call_h, long call_cap);
/* FNV-1a 64 over the bytes of n little-endian doubles. */
unsigned long long orc_fnv64(const double* v, size_t n);
/* The same integrator on the scalar test ODEs of the reference RK tests:
 * kind 0: u' = -u (n = 1);  kind 1: harmonic oscillator (n = 2);
 * kind 2: u' = lam*u (n = 1, lam = param). */
int orc_integrate_ode(int kind, double param, double* u, int n, double t0, double tEnd,
                      double dt0, const orc_rk_opts* o, orc_rk_stats* st);
void orc_rk4_step_ode(int kind, double param, double* u, int n, double t, double dt);
void orc_bs3_step_ode(int kind, double param, double* u, int n, double t, double dt);

#ifdef __cplusplus
}
#endif

#endif /* SGN_ORACLE_H */
