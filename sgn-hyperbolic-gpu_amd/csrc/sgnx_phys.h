// =============================================================================
// sgnx_phys.h -- pointwise hyperbolic-SGN physics for the gfx950 kernels.
//
// The expression trees are those of src/physics/sgn_flux2d.h (the reference's
// pointwise physics, compiled there for host and device from one text); they
// are the bitwise contract, so every product, sum and parenthesisation below is
// kept. What differs is structure: the four primitive quotients u, v, w, p are
// formed once per state (Prim) and shared by both fluxes, the wave speed, the
// source and the non-conservative products. Every reference function divides
// the same operands, so sharing the quotient changes no bit.
//
// Compiled with -ffp-contract=off: no multiply-add is fused anywhere.
// =============================================================================
#ifndef SGNX_PHYS_H
#define SGNX_PHYS_H

#include <hip/hip_runtime.h>

namespace sgnx {

constexpr int NV = 5;                          // h, hu, hv, hw, hp
constexpr int IH = 0, IHU = 1, IHV = 2, IHW = 3, IHP = 4;

// Run constants; the derived ones are the reference's sub-expressions
// (0.5*par.g, par.c*par.c, -par.c*par.c), which evaluate to the same double
// wherever they are computed.
struct Phys {
  double g, c, gamma, nManning;
  double halfg;   // 0.5 * g           (sgn_flux2d.h:97, :112)
  double c2;      // c * c             (sgn_flux2d.h:95, :110, :127)
  double mc2;     // (-c) * c          (sgn_flux2d.h:154, :160)
  double m2c2;    // -2.0 * c2         (sgn_flux2d.h:134)
};

struct Prim { double u, v, w, p; };

__device__ __forceinline__ Prim prims(const double* U) {
  Prim q;
  q.u = U[IHU] / U[IH];
  q.v = U[IHV] / U[IH];
  q.w = U[IHW] / U[IH];
  q.p = U[IHP] / U[IH];
  return q;
}

// sgn_flux2d.h:73-78: sqrt(max(g h + p + c^2, 0))
__device__ __forceinline__ double signal_speed(double h, const Prim& q, const Phys& ph) {
  const double arg = ph.g * h + q.p + ph.c2;
  return sqrt(arg > 0.0 ? arg : 0.0);
}

// sgn_flux2d.h:89-101 (x) and :104-116 (y)
__device__ __forceinline__ void flux_x(const double* U, const Prim& q, double Hl,
                                       const Phys& ph, double* F) {
  const double h = U[IH], hu = U[IHU];
  F[IH] = hu;
  F[IHU] = hu * q.u + ph.halfg * (h * h - Hl * Hl) + h * q.p;
  F[IHV] = hu * q.v;
  F[IHW] = hu * q.w;
  F[IHP] = hu * (q.p + ph.c2);
}
__device__ __forceinline__ void flux_y(const double* U, const Prim& q, double Hl,
                                       const Phys& ph, double* G) {
  const double h = U[IH], hv = U[IHV];
  G[IH] = hv;
  G[IHU] = hv * q.u;
  G[IHV] = hv * q.v + ph.halfg * (h * h - Hl * Hl) + h * q.p;
  G[IHW] = hv * q.w;
  G[IHP] = hv * (q.p + ph.c2);
}

// One interface node: Rusanov flux (sgn_flux2d.h:169-179) plus the two
// one-sided fluxes it is built from, which are also the two elements' own
// discontinuous face fluxes (sgn2d.cpp:229-239 evaluates them with the same
// state and face depth).
template <int DIR>
__device__ __forceinline__ void interface_flux(const double* UL, const double* UR, double Hf,
                                               const Phys& ph, double* Fs, double* FL,
                                               double* FR) {
  const Prim qL = prims(UL), qR = prims(UR);
  if (DIR == 0) { flux_x(UL, qL, Hf, ph, FL); flux_x(UR, qR, Hf, ph, FR); }
  else          { flux_y(UL, qL, Hf, ph, FL); flux_y(UR, qR, Hf, ph, FR); }
  // sgn_flux2d.h:82-86: |normal velocity| + Ce, the normal velocity being the
  // same quotient hu/h (hv/h) as the primitive
  const double sL = fabs(DIR == 0 ? qL.u : qL.v) + signal_speed(UL[IH], qL, ph);
  const double sR = fabs(DIR == 0 ? qR.u : qR.v) + signal_speed(UR[IH], qR, ph);
  const double lam = fmax(sL, sR);
  const double hl = 0.5 * lam;
#pragma unroll
  for (int k = 0; k < NV; ++k) Fs[k] = 0.5 * (FL[k] + FR[k]) - hl * (UR[k] - UL[k]);
}

// sgn_flux2d.h:122-143 (source; the Manning branch uses the device cbrt, so
// it is the one place that is not bitwise against the host libm). Split where
// the still-depth gradients Hx, Hy enter: source_part needs only the state
// and the nodal depth, source_rest adds the gradient terms, so a kernel can
// form Hx, Hy after the state has been consumed. source() is the two in turn.
struct SrcPart {
  double bath;   // g (h - H) + gamma p
  double pw;     // gamma p
  double cf;     // Manning: n^2 g |u| / cbrt(h)
};

__device__ __forceinline__ SrcPart source_part(const double* U, const Prim& q, double Hl,
                                               const Phys& ph) {
  const double h = U[IH];
  SrcPart sp;
  sp.bath = ph.g * (h - Hl) + ph.gamma * q.p;
  sp.pw = ph.gamma * q.p;
  sp.cf = 0.0;
  if (ph.nManning != 0.0) {
    const double n2 = ph.nManning * ph.nManning;
    const double vmag = sqrt(q.u * q.u + q.v * q.v);
    sp.cf = n2 * ph.g * vmag / cbrt(h);
  }
  return sp;
}

__device__ __forceinline__ void source_rest(const SrcPart& sp, double qu, double qv, double qw,
                                            double Hx, double Hy, const Phys& ph, double* S) {
  S[IH] = 0.0;
  S[IHU] = sp.bath * Hx;
  S[IHV] = sp.bath * Hy;
  S[IHW] = sp.pw;
  S[IHP] = ph.m2c2 * (qw + qu * Hx + qv * Hy);
  if (ph.nManning != 0.0) {
    S[IHU] -= sp.cf * qu;
    S[IHV] -= sp.cf * qv;
  }
}

__device__ __forceinline__ void source(const double* U, const Prim& q, double Hl, double Hx,
                                       double Hy, const Phys& ph, double* S) {
  source_rest(source_part(U, q, Hl, ph), q.u, q.v, q.w, Hx, Hy, ph, S);
}

}  // namespace sgnx

#endif  // SGNX_PHYS_H
