// =============================================================================
// sgnx_adjoint.h -- device side of the discrete adjoint: the arguments of the
// RHS vector-Jacobian product and the launchers of sgnx_adjoint.hip.
// =============================================================================
#ifndef SGNX_ADJOINT_H
#define SGNX_ADJOINT_H

#include <hip/hip_runtime.h>

#include "../../include/sgnx.h"
#include "sgnx_internal.h"

namespace sgnx {

// out = acc0 + acc1 + acc2 + (dR/dU (U))^T W,  W = ca*A + cb*B,
// thetaBar += (dR/dtheta)^T W  (when not null).
// A, B, acc* and out are 5*Nd arrays; any acc* or B may be null; out may be
// the same array as A or an acc* (every node is read before it is written,
// by the thread that writes it). coef: 120*Nx*Ny doubles of scratch.
struct VjpArgs {
  const double* U;
  const double* A;
  double ca;
  const double* B;
  double cb;
  const double* acc0;
  const double* acc1;
  const double* acc2;
  double* out;
  double* thetaBar;
  double* coef;
  // Optional (coefNext not null): the launch also writes the face
  // coefficients of the step's next product, whose cotangent is
  // can*An + cbn*out (out: this product's result) at the state Un, into
  // coefNext -- what the coefficient kernel would write for it.
  const double* Un;
  const double* An;
  double can, cbn;
  double* coefNext;
  bool coefReady;   // coef already holds this product's coefficients
};

inline long long vjp_coef_doubles(const Geometry& g) { return 120LL * g.Nx * g.Ny; }

// Batched forms (one launch over every member, tiles packed behind a
// per-member offset table as in the batched RHS): a product per member ...
struct VjpMember {
  VjpArgs va;
  Geometry g;
  Phys ph;
  GhostTime gt;
};
// ... and one gauge log row's cotangents per member (Gb null: none)
struct GaugeMember {
  const double* Gb;
  const long long* nodes;
  int n;
  double* lam;
  double* thetaBar;
};
long long vjp_tiles(const Geometry& g);       // workgroups of one product launch
hipError_t launch_rhs_vjp_batch(const VjpMember* mem, const long long* tileStart, int nMem,
                                long long totalTiles, bool coefReady, hipStream_t st);
hipError_t launch_gauge_bar_batch(const GaugeMember* gm, int nMem, hipStream_t st);

hipError_t launch_rhs_vjp(const Geometry& g, const Phys& ph, const VjpArgs& va, GhostTime gt,
                          hipStream_t st);
// HfX, HfY from the nodal still depth theta: the mean of the two edge values
// meeting at a line node (the wrapped pair on a periodic axis), the one
// interior edge value on a non-periodic side.
hipError_t launch_face_depth(const Geometry& g, const double* theta, double* HfX, double* HfY,
                             hipStream_t st);
// lam[nodes[k]] += Gb[k]; thetaBar[nodes[k]] -= Gb[k] (thetaBar may be null)
hipError_t launch_gauge_bar(const double* Gb, const long long* nodes, int n, double* lam,
                            double* thetaBar, hipStream_t st);

}  // namespace sgnx

#endif  // SGNX_ADJOINT_H
