// =============================================================================
// sgnx_adjoint.hip -- kernels of the discrete adjoint: the vector-Jacobian
// product of the FR/SGN right-hand side, the still depth set from one nodal
// field theta, and the gauge cotangent scatter. No counterpart in the
// reference, which has no adjoint.
//
// R is the composition of the five phases of src/solver/sgn2d.cpp:125-290
// (with the wall / inflow sides): nodal fluxes, sources and non-conservative
// products, a Rusanov flux on every interface line, the FR lift of the jumps.
// For a cotangent W of dU/dt the product returns
//   Ubar      = (dR/dU)^T W
//   thetaBar += (dR/dtheta)^T W,  theta the nodal still depth, with
//               Hx, Hy = (D theta)/J per element line (sgn2d.cpp:54-71) and the
//               face depth of an interface line node the mean of the two edge
//               values meeting there (the one interior value on a non-periodic
//               side); sgnx_set_depth builds that still depth.
// Two launches:
//   k_vjp_coef  per element line: the cotangents of the FR lift's face jumps
//               (the transposed lift, Lambda = -(g' . W)/J) and of the
//               central face depths of the non-conservative products
//               (lambda = (g' . W_hp c^2 u)/J), 6 numbers per face node;
//   k_vjp_elem  per element: the volume adjoint (source, non-conservative
//               products, nodal fluxes through D^T), and for each face node the
//               Rusanov adjoint of every interface line its edge state feeds,
//               kept for the own side only (each interior line is thus
//               evaluated by both of its elements), the derivative of
//               max(sL, sR) taken on the side that attains it, walls mirrored
//               back onto the interior state, the inflow ghost constant.
// Every sum has one writer and a fixed order, so the results are
// deterministic. The products are the derivatives of R away from its
// non-smooth points (a zero normal velocity, sL = sR off a wall, a clamped
// wave-speed argument). The Manning term is not differentiated (the entry
// points refuse nManning != 0).
// =============================================================================

#include <hip/hip_runtime.h>

#include "sgnx_adjoint.h"

namespace sgnx {
namespace {

constexpr int kT = 256;    // threads per block of the elementwise kernels
constexpr int kEW = 5;     // elements per tile of k_vjp_coef / k_vjp_elem (one element row)
constexpr int kVT = ((kEW * NPP + 63) / 64) * 64;   // their threads per block (whole waves)
static_assert(20 * kEW <= kVT, "one thread per element face node");

// --- pointwise adjoints of sgnx_phys.h (nManning = 0) --------------------------
// Every term is differentiated with the primitives u = hu/h, v, w, p as
// intermediate variables: a term adds its explicit U-derivatives to Ub and its
// primitive cotangents to a PrimBar, and prims_fin maps the summed PrimBar
// onto U once per state.

// a state with its reciprocal depth and primitives, formed once
struct St {
  double h, hu, hv, hw, hp;
  double ih, u, v, w, p;
};
__device__ __forceinline__ St make_st(const double* U) {
  St s;
  s.h = U[IH];
  s.hu = U[IHU];
  s.hv = U[IHV];
  s.hw = U[IHW];
  s.hp = U[IHP];
  s.ih = 1.0 / s.h;
  s.u = s.hu * s.ih;
  s.v = s.hv * s.ih;
  s.w = s.hw * s.ih;
  s.p = s.hp * s.ih;
  return s;
}

struct PrimBar { double u, v, w, p; };

// Ub += (d(u, v, w, p)/dU)^T q
__device__ __forceinline__ void prims_fin(const St& s, const PrimBar& q, double* Ub) {
  Ub[IHU] += q.u * s.ih;
  Ub[IHV] += q.v * s.ih;
  Ub[IHW] += q.w * s.ih;
  Ub[IHP] += q.p * s.ih;
  Ub[IH] -= (((q.u * s.u + q.v * s.v) + q.w * s.w) + q.p * s.p) * s.ih;
}

// flux along x (DIR 0, normal momentum hu) or y (DIR 1, hv) at still depth Hl,
// F = (m, m u, m v, m w, m (p + c^2)) + P e_N,  P = g/2 (h^2 - Hl^2) + h p:
// cotangent Fb; returns (dF/dHl) . Fb
template <int DIR>
__device__ __forceinline__ double flux_acc(const St& s, double Hl, const Phys& ph,
                                           const double* Fb, double* Ub, PrimBar& q) {
  constexpr int N = DIR == 0 ? IHU : IHV;
  const double m = DIR == 0 ? s.hu : s.hv;
  Ub[N] += Fb[IH] + (((Fb[IHU] * s.u + Fb[IHV] * s.v) + Fb[IHW] * s.w) + Fb[IHP] * (s.p + ph.c2));
  const double Pb = Fb[N];
  Ub[IH] += Pb * (ph.g * s.h + s.p);
  q.u += Fb[IHU] * m;
  q.v += Fb[IHV] * m;
  q.w += Fb[IHW] * m;
  q.p += Fb[IHP] * m + Pb * s.h;
  return -Pb * (ph.g * Hl);
}

// wave speed |vn| + sig, sig = sqrt(max(g h + p + c^2, 0)) given: cotangent sb
template <int DIR>
__device__ __forceinline__ void speed_acc(const St& s, double sig, const Phys& ph, double sb,
                                          double* Ub, PrimBar& q) {
  const double vn = DIR == 0 ? s.u : s.v;
  const double sg = vn > 0.0 ? sb : (vn < 0.0 ? -sb : 0.0);
  if (DIR == 0) q.u += sg;
  else q.v += sg;
  if (sig > 0.0) {
    const double ab = sb * (0.5 / sig);
    Ub[IH] += ab * ph.g;
    q.p += ab;
  }
}

// the node's own state terms: source S and the non-conservative products
// -Bx - By = c^2 (u hx + v hy) e_hp, with cotangent W; sets the still-depth
// cotangents Hlb, Hxb, Hyb
__device__ __forceinline__ void node_terms_acc(const St& s, double Hl, double Hx, double Hy,
                                               double hx, double hy, const Phys& ph,
                                               const double* W, double* Ub, PrimBar& q,
                                               double& Hlb, double& Hxb, double& Hyb) {
  const double bath = ph.g * (s.h - Hl) + ph.gamma * s.p;
  const double bathb = W[IHU] * Hx + W[IHV] * Hy;
  const double k = ph.m2c2 * W[IHP];              // S_hp = -2 c^2 (w + u Hx + v Hy)
  const double c = W[IHP] * ph.c2;                // ncp: c^2 (u hx + v hy)
  Hxb = W[IHU] * bath + k * s.u;
  Hyb = W[IHV] * bath + k * s.v;
  Hlb = -ph.g * bathb;
  Ub[IH] += ph.g * bathb;
  q.u += k * Hx + c * hx;
  q.v += k * Hy + c * hy;
  q.w += k;
  q.p += ph.gamma * W[IHW] + ph.gamma * bathb;
}

// One interface line node: Fs = 0.5 (FL + FR) - 0.5 lam (UR - UL) with its
// cotangent LamL + LamR, the two sides' own face fluxes FL, FR (same state,
// same face depth) with cotangents -LamL, -LamR, and the central depth
// 0.5 (hL + hR) with cotangent lh. UbL, UbR = the state cotangents; returns the
// face-depth cotangent.
template <int DIR>
__device__ __forceinline__ double line_bar(const double* UL, const double* UR, double Hf,
                                           const double* LamL, const double* LamR, double lh,
                                           const Phys& ph, double* UbL, double* UbR) {
  const St L = make_st(UL), R = make_st(UR);
  const double aL = ph.g * L.h + L.p + ph.c2, aR = ph.g * R.h + R.p + ph.c2;
  const double gL = sqrt(aL > 0.0 ? aL : 0.0), gR = sqrt(aR > 0.0 ? aR : 0.0);
  const double sL = fabs(DIR == 0 ? L.u : L.v) + gL;
  const double sR = fabs(DIR == 0 ? R.u : R.v) + gR;
  const double hl = 0.5 * fmax(sL, sR);
  PrimBar qL{0.0, 0.0, 0.0, 0.0}, qR{0.0, 0.0, 0.0, 0.0};
  double FbL[NV], FbR[NV], hlb = 0.0;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double Fsb = LamL[k] + LamR[k];
    FbL[k] = 0.5 * Fsb - LamL[k];
    FbR[k] = 0.5 * Fsb - LamR[k];
    UbL[k] = hl * Fsb;
    UbR[k] = -hl * Fsb;
    hlb -= Fsb * (UR[k] - UL[k]);
  }
  UbL[IH] += 0.5 * lh;
  UbR[IH] += 0.5 * lh;
  if (sL > sR) speed_acc<DIR>(L, gL, ph, 0.5 * hlb, UbL, qL);
  else speed_acc<DIR>(R, gR, ph, 0.5 * hlb, UbR, qR);
  const double hfb = flux_acc<DIR>(L, Hf, ph, FbL, UbL, qL) + flux_acc<DIR>(R, Hf, ph, FbR, UbR, qR);
  prims_fin(L, qL, UbL);
  prims_fin(R, qR, UbR);
  return hfb;
}

__device__ __forceinline__ double cot_at(const VjpArgs& va, long long j) {
  double w = va.ca * va.A[j];
  if (va.B) w = w + va.cb * va.B[j];
  return w;
}

// local node of face f (0 left, 1 right, 2 bottom, 3 top), line node n
__device__ __forceinline__ int face_loc(int f, int n) {
  return f == 0 ? NP * n : f == 1 ? NP - 1 + NP * n : f == 2 ? n : n + NP * (NP - 1);
}

// ---- element tiles --------------------------------------------------------------
// Both launches work on tiles of kEW elements of one element row, one node per
// thread; the tile's cotangents W and the integrands of the depth-gradient
// cotangents, c^2 W_hp u / Jx and c^2 W_hp v / Jy, are staged in LDS.
struct TileLds {
  double D[NPP], gL[NP], gR[NP];
  double W[NV][kEW * NPP];
  double DX[kEW * NPP], DY[kEW * NPP];
};

// Lambda_c = -(g' . W_c)/J along the face's element line (the cotangent of
// the FR lift's face jump) and lambda = g' . (c^2 W_hp vel / J) (that of the
// central face depth of the non-conservative products), g' = gL' on a low
// face, gR' on a high one; f = 0 left, 1 right, 2 bottom, 3 top, line node n
__device__ __forceinline__ void face_coef(const TileLds& L, int jb, int f, int n, double iJx,
                                          double iJy, double* Lam, double& lam) {
  const int dir = f >> 1;
  const bool hi = (f & 1) != 0;
  double sum[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double q = 0.0;
  // rolled: fewer registers live in the product kernel's face phase (its
  // four-wave build spills 3 registers instead of 5)
#pragma unroll 1
  for (int k = 0; k < NP; ++k) {
    const int node = jb + (dir == 0 ? k + NP * n : n + NP * k);
    const double gk = hi ? L.gR[k] : L.gL[k];
#pragma unroll
    for (int c = 0; c < NV; ++c) sum[c] = sum[c] + gk * L.W[c][node];
    q = q + gk * (dir == 0 ? L.DX[node] : L.DY[node]);
  }
  const double iJ = dir == 0 ? iJx : iJy;
#pragma unroll
  for (int c = 0; c < NV; ++c) Lam[c] = -sum[c] * iJ;
  lam = q;
}

// the tile's operator tables and node data; returns the thread's node index
__device__ __forceinline__ void tile_stage(const VjpArgs& va, const Geometry& g, const Phys& ph,
                                           TileLds& L, int t, bool isNode, long long i,
                                           double* sH) {
  const OpTables* T = g.ops;
  if (t < NPP) L.D[t] = T->D[t];
  if (t < NP) {
    L.gL[t] = T->gLp[t];
    L.gR[t] = T->gRp[t];
  }
  if (!isNode) return;
  const long long Nd = g.Nd;
  double W[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    W[c] = cot_at(va, (long long)c * Nd + i);
    L.W[c][t] = W[c];
  }
  const double h = va.U[i];
  if (sH) sH[t] = h;
  const double q = W[IHP] * ph.c2 / h;
  L.DX[t] = q * va.U[(long long)IHU * Nd + i] * (1.0 / g.Jx);
  L.DY[t] = q * va.U[(long long)IHV * Nd + i] * (1.0 / g.Jy);
}

// ---- launch 1: face-line cotangents -----------------------------------------------
// coef[((e * 4 + f) * NP + n) * 6 + c]: c < 5 Lambda_c, c = 5 lambda
__device__ __forceinline__ void vjp_coef_tile(const VjpArgs& va, const Geometry& g,
                                              const Phys& ph, int tile) {
  __shared__ TileLds L;
  const int t = threadIdx.x;
  const int Nx = g.Nx;
  const int strips = (Nx + kEW - 1) / kEW;
  const int ey = tile / strips;
  const int ex0 = (tile - ey * strips) * kEW;
  const int nW = min(kEW, Nx - ex0);
  const int j = t / NPP;
  const long long i = ((long long)ey * Nx + ex0) * NPP + t;
  tile_stage(va, g, ph, L, t, j < nW, i, nullptr);
  __syncthreads();
  if (t < 20 * nW) {
    const int jf = t / 20, r = t - 20 * jf;
    const int f = r / NP, n = r - NP * f;
    double Lam[NV], lam;
    face_coef(L, NPP * jf, f, n, 1.0 / g.Jx, 1.0 / g.Jy, Lam, lam);
    double* cp = va.coef + (((long long)ey * Nx + ex0) * 20 + t) * 6;
#pragma unroll
    for (int c = 0; c < NV; ++c) cp[c] = Lam[c];
    cp[NV] = lam;
  }
}

// ---- launch 2: one element row strip per block ---------------------------------
__device__ __forceinline__ void vjp_elem_tile(const VjpArgs& va, const Geometry& g,
                                              const Phys& ph, const GhostTime& gt, int tile) {
  __shared__ TileLds L;
  __shared__ double sH[kEW * NPP];
  __shared__ double sHS[kEW * 20], sLam[kEW * 20], sTF[kEW * 20];
  __shared__ double sUF[NV][kEW * 20];
  const int t = threadIdx.x;
  const int Nx = g.Nx, Ny = g.Ny;
  const long long Nd = g.Nd;
  const int strips = (Nx + kEW - 1) / kEW;
  const int ey = tile / strips;
  const int ex0 = (tile - ey * strips) * kEW;
  const int nW = min(kEW, Nx - ex0);
  const double* U = va.U;
  const double* sD = L.D;
  const double* sgL = L.gL;
  const double* sgR = L.gR;
  const double iJx = 1.0 / g.Jx, iJy = 1.0 / g.Jy;

  // phase 1: the cotangent loads, the cotangents of the two D-sums and of the
  // depth gradients of the non-conservative products; the face threads (20
  // per element, see phase 2) issue their global loads before the barrier too
  const int j = t / NPP, loc = t - NPP * (t / NPP);
  const int b = loc / NP, a = loc - NP * b;
  const bool isNode = j < nW;
  const long long i = ((long long)ey * Nx + ex0 + j) * NPP + loc;
  const bool isFace = t < 20 * nW;
  const int jf = t / 20, r = t - 20 * (t / 20);
  const int f = r / NP, n = r - NP * f;
  const int ex = ex0 + jf;
  const long long e = (long long)ey * Nx + ex;
  const int dir = f >> 1;
  const bool ownHigh = (f & 1) == 0;               // own element on the line's high side
  const int pos = dir == 0 ? ex : ey, nL = dir == 0 ? Nx : Ny;
  const int sideLo = dir == 0 ? 0 : 2;              // bc index of the low side
  const bool periodic = g.bc[sideLo] == SGNX_BC_PERIODIC;
  const int posO = ownHigh ? pos - 1 : pos + 1;
  const bool inLine = posO >= 0 && posO < nL;
  const int wrapO = posO < 0 ? nL - 1 : (posO >= nL ? 0 : posO);
  const int bcSide = ownHigh ? sideLo : sideLo + 1;
  const bool ghost = !inLine && !periodic;
  const bool wall = ghost && g.bc[bcSide] == SGNX_BC_WALL;
  // pass 0: the own line; pass 1 (an edge on the wrap of a periodic axis):
  // the line at the far end, where only the far element's face is in line
  const int passes = (!inLine && periodic) ? 2 : 1;
  double Uo[NV], Uw[NV], Lw[NV], lw = 0.0, hfp[2] = {0.0, 0.0};
  if (isFace) {
    const long long io = e * NPP + face_loc(f, n);
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      Uo[c] = U[(long long)c * Nd + io];
      Lw[c] = 0.0;
    }
    if (!ghost) {
      // the element across the line (across the wrap on a periodic axis)
      const long long eo = dir == 0 ? (long long)ey * Nx + wrapO : (long long)wrapO * Nx + ex;
      const long long ioth = eo * NPP + face_loc(f ^ 1, n);
      const double* cw = va.coef + ((eo * 4 + (f ^ 1)) * NP + n) * 6;
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        Uw[c] = U[(long long)c * Nd + ioth];
        Lw[c] = cw[c];
      }
      lw = cw[NV];
    } else if (!wall) {
      const double* cf = g.inflow + ((long long)ey * NP + n) * (NV * 3);
#pragma unroll
      for (int c = 0; c < NV; ++c) Uw[c] = (cf[3 * c] + cf[3 * c + 1] * gt.s) + cf[3 * c + 2] * gt.c;
    }
#pragma unroll
    for (int ps = 0; ps < 2; ++ps) {
      if (ps < passes) {
        const int line = ps == 0 ? pos + (ownHigh ? 0 : 1) : (ownHigh ? nL : 0);
        hfp[ps] = dir == 0 ? g.HfX[((long long)ey * (Nx + 1) + line) * NP + n]
                           : g.HfY[((long long)line * Nx + ex) * NP + n];
      }
    }
  }
  tile_stage(va, g, ph, L, t, isNode, i, sH);
  __syncthreads();

  // phase 2: face nodes (20 per element): the interface lines each edge
  // state feeds, their adjoint kept for the own side. A y line is handled in
  // the x frame: with P the exchange of hu and hv, flux_y(U) = P flux_x(P U)
  // and the y wave speed is the x one of P U, so its adjoint is P applied to
  // the x adjoint of the exchanged states and cotangents.
  if (isFace) {
    double Lo[NV], lo;
    face_coef(L, NPP * jf, f, n, iJx, iJy, Lo, lo);
    if (wall) {
#pragma unroll
      for (int c = 0; c < NV; ++c) Uw[c] = Uo[c];
      Uw[dir == 0 ? IHU : IHV] = -Uo[dir == 0 ? IHU : IHV];
    }
    auto to_x = [&](double* x) {
      if (dir == 1) {
        const double tmp = x[IHU];
        x[IHU] = x[IHV];
        x[IHV] = tmp;
      }
    };
    to_x(Uo);
    to_x(Lo);
    to_x(Uw);
    to_x(Lw);
    double Ub[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double tf = 0.0;
    for (int ps = 0; ps < passes; ++ps) {
      const double hf = ps == 0 ? hfp[0] : hfp[1];
      const bool ownIn = ps == 0, othIn = ps == 1 || inLine;
      double UL[NV], UR[NV], LL[NV], LR[NV], UbL[NV], UbR[NV];
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        const double lo_c = ownIn ? Lo[c] : 0.0, lw_c = othIn ? Lw[c] : 0.0;
        UL[c] = ownHigh ? Uw[c] : Uo[c];
        UR[c] = ownHigh ? Uo[c] : Uw[c];
        LL[c] = ownHigh ? lw_c : lo_c;
        LR[c] = ownHigh ? lo_c : lw_c;
      }
      const double lh = (ownIn ? lo : 0.0) + (othIn ? lw : 0.0);
      const double hfb = line_bar<0>(UL, UR, hf, LL, LR, lh, ph, UbL, UbR);
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        const double own = ownHigh ? UbR[c] : UbL[c], oth = ownHigh ? UbL[c] : UbR[c];
        // a wall ghost is the own state with the normal momentum negated
        Ub[c] += wall ? (c == IHU ? own - oth : own + oth) : own;
      }
      // face depth: the mean of two edge values, or the one interior value
      // on a non-periodic side
      tf += ghost ? hfb : 0.5 * hfb;
      if (ps == 0) sHS[t] = 0.5 * (UL[IH] + UR[IH]);
    }
    to_x(Ub);
#pragma unroll
    for (int c = 0; c < NV; ++c) sUF[c][t] = Ub[c];
    sLam[t] = lo;
    sTF[t] = tf;
  }
  __syncthreads();

  // the node's own operands of phase 3, loaded after the face phase (h, hu,
  // hv are lines tile_stage has just read), which keeps them out of the
  // registers live across it: 0.7 % faster than loading them in phase 1
  double Un[NV], Hl = 0.0, Hxl = 0.0, Hyl = 0.0;
  if (isNode) {
#pragma unroll
    for (int c = 0; c < NV; ++c) Un[c] = U[(long long)c * Nd + i];
    Hl = g.Hnod[i];
    Hxl = g.Hx[i];
    Hyl = g.Hy[i];
  }
  // phase 3: node adjoint
  const int jb = NPP * j, fb = 20 * j;
  double Hxb = 0.0, Hyb = 0.0, th = 0.0;
  double outv[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};   // the node's result, for the next product
  if (isNode) {
    double dhx = 0.0, dhy = 0.0;
    // rolled, like the D^T loop below: unrolled, the compiler hoists its LDS
    // operands into 20 more VGPRs, which keeps the kernel off four waves
#pragma unroll 1
    for (int k = 0; k < NP; ++k) {
      dhx = dhx + sD[a * NP + k] * sH[jb + k + NP * b];
      dhy = dhy + sD[b * NP + k] * sH[jb + a + NP * k];
    }
    dhx = dhx + ((sHS[fb + b] - sH[jb + NP * b]) * sgL[a] +
                 (sHS[fb + NP + b] - sH[jb + NP - 1 + NP * b]) * sgR[a]);
    dhy = dhy + ((sHS[fb + 2 * NP + a] - sH[jb + a]) * sgL[b] +
                 (sHS[fb + 3 * NP + a] - sH[jb + a + NP * (NP - 1)]) * sgR[b]);
    const double hx = dhx * iJx, hy = dhy * iJy;
    double Ub[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double Hlb, Wn[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) Wn[c] = L.W[c][t];
    const St sn = make_st(Un);
    PrimBar qn{0.0, 0.0, 0.0, 0.0};
    node_terms_acc(sn, Hl, Hxl, Hyl, hx, hy, ph, Wn, Ub, qn, Hlb, Hxb, Hyb);
    // the nodal flux cotangents: D^T applied along the row (x) and the column
    // (y) to the cotangents -W/Jx, -W/Jy of the two D-sums. The k loop stays
    // rolled so that only one k's ten LDS operands are live at a time.
    double Fb[NV] = {0.0, 0.0, 0.0, 0.0, 0.0}, Gb[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
    for (int k = 0; k < NP; ++k) {
      const double dxk = sD[k * NP + a], dyk = sD[k * NP + b];
      const int nx = jb + k + NP * b, ny = jb + a + NP * k;
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        Fb[c] = Fb[c] + dxk * L.W[c][nx];
        Gb[c] = Gb[c] + dyk * L.W[c][ny];
      }
    }
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      Fb[c] = -Fb[c] * iJx;
      Gb[c] = -Gb[c] * iJy;
    }
    Hlb += flux_acc<0>(sn, Hl, ph, Fb, Ub, qn);
    Hlb += flux_acc<1>(sn, Hl, ph, Gb, Ub, qn);
    prims_fin(sn, qn, Ub);
    double hb = 0.0;
#pragma unroll
    for (int k = 0; k < NP; ++k)
      hb = hb + (sD[k * NP + a] * L.DX[jb + k + NP * b] + sD[k * NP + b] * L.DY[jb + a + NP * k]);
    if (a == 0) hb -= sLam[fb + b];
    if (a == NP - 1) hb -= sLam[fb + NP + b];
    if (b == 0) hb -= sLam[fb + 2 * NP + a];
    if (b == NP - 1) hb -= sLam[fb + 3 * NP + a];
    Ub[IH] += hb;
    double tb = 0.0;
    auto face_add = [&](int slot) {
#pragma unroll
      for (int c = 0; c < NV; ++c) Ub[c] += sUF[c][slot];
      tb += sTF[slot];
    };
    if (a == 0) face_add(fb + b);
    if (a == NP - 1) face_add(fb + NP + b);
    if (b == 0) face_add(fb + 2 * NP + a);
    if (b == NP - 1) face_add(fb + 3 * NP + a);
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      const long long o = (long long)c * Nd + i;
      double base = 0.0;
      if (va.acc0) base = va.acc0[o];
      if (va.acc1) base = base + va.acc1[o];
      if (va.acc2) base = base + va.acc2[o];
      outv[c] = base + Ub[c];
      va.out[o] = outv[c];
    }
    th = Hlb + tb;
  }
  if (va.thetaBar) {   // uniform over the block
    // still depth: Hnod = theta, Hx = (D theta)/Jx, Hy = (D theta)/Jy per line
    __syncthreads();
    if (isNode) {
      L.DX[t] = Hxb * iJx;
      L.DY[t] = Hyb * iJy;
    }
    __syncthreads();
    if (isNode) {
#pragma unroll
      for (int k = 0; k < NP; ++k)
        th = th + (sD[k * NP + a] * L.DX[jb + k + NP * b] + sD[k * NP + b] * L.DY[jb + a + NP * k]);
      va.thetaBar[i] += th;
    }
  }
  if (va.coefNext) {   // uniform over the block
    // the next product's cotangent can*An + cbn*out on this tile, staged as
    // tile_stage stages a product's own, then its face coefficients as
    // vjp_coef_tile forms them
    __syncthreads();
    if (isNode) {
      double Wn[NV];
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        Wn[c] = va.can * va.An[(long long)c * Nd + i];
        Wn[c] = Wn[c] + va.cbn * outv[c];
        L.W[c][t] = Wn[c];
      }
      const double q = Wn[IHP] * ph.c2 / va.Un[i];
      L.DX[t] = q * va.Un[(long long)IHU * Nd + i] * (1.0 / g.Jx);
      L.DY[t] = q * va.Un[(long long)IHV * Nd + i] * (1.0 / g.Jy);
    }
    __syncthreads();
    if (isFace) {
      double Ln[NV], ln;
      face_coef(L, NPP * jf, f, n, iJx, iJy, Ln, ln);
      double* cp = va.coefNext + (((long long)ey * Nx + ex0) * 20 + t) * 6;
#pragma unroll
      for (int c = 0; c < NV; ++c) cp[c] = Ln[c];
      cp[NV] = ln;
    }
  }
}

__global__ void __launch_bounds__(kVT)
k_vjp_coef(VjpArgs va, Geometry g, Phys ph) {
  vjp_coef_tile(va, g, ph, blockIdx.x);
}

// four waves per SIMD (at most 128 VGPRs; 3 spilled, 12 B of scratch per
// lane): 5.6 % faster per product and 8.7 % per reversed step than three
// waves (134 VGPRs), same bits
__global__ void __launch_bounds__(kVT, 4)
k_vjp_elem(VjpArgs va, Geometry g, Phys ph, GhostTime gt) {
  vjp_elem_tile(va, g, ph, gt, blockIdx.x);
}

// ---- batches: every member's tiles packed into one launch -------------------------
// the member owning packed tile w: tileStart[m] <= w < tileStart[m + 1]
__device__ __forceinline__ int member_of(const long long* tileStart, int nMem, long long w) {
  int lo = 0, hi = nMem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tileStart[mid] <= w) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(kVT)
k_vjp_coef_batch(const VjpMember* __restrict__ mem, const long long* __restrict__ tileStart,
                 int nMem) {
  const long long w = blockIdx.x;
  const int m = member_of(tileStart, nMem, w);
  vjp_coef_tile(mem[m].va, mem[m].g, mem[m].ph, (int)(w - tileStart[m]));
}

__global__ void __launch_bounds__(kVT, 4)
k_vjp_elem_batch(const VjpMember* __restrict__ mem, const long long* __restrict__ tileStart,
                 int nMem) {
  const long long w = blockIdx.x;
  const int m = member_of(tileStart, nMem, w);
  vjp_elem_tile(mem[m].va, mem[m].g, mem[m].ph, mem[m].gt, (int)(w - tileStart[m]));
}

// lam[node] += Gb[k], thetaBar[node] -= Gb[k] for the n gauges of one log row,
// by one block: the thread of a node's first gauge adds that node's gauges in
// gauge order, so every node receives the sums a sequential scan over the
// gauges forms (a node may carry several gauges), bit for bit, and distinct
// nodes proceed in parallel instead of one dependent load after another.
__device__ __forceinline__ void gauge_bar_block(const double* __restrict__ Gb,
                                                const long long* __restrict__ nodes, int n,
                                                double* __restrict__ lam,
                                                double* __restrict__ thetaBar) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const long long i = nodes[k];
    bool first = true;
    for (int j = 0; j < k && first; ++j) first = nodes[j] != i;
    if (!first) continue;
    double l = lam[i];
    double th = thetaBar ? thetaBar[i] : 0.0;
    for (int j = k; j < n; ++j) {
      if (nodes[j] != i) continue;
      l += Gb[j];
      th -= Gb[j];
    }
    lam[i] = l;
    if (thetaBar) thetaBar[i] = th;
  }
}

// the gauge cotangents of one log row per member (one block per member)
__global__ void k_gauge_bar_batch(const GaugeMember* __restrict__ gm, int nMem) {
  if ((int)blockIdx.x >= nMem) return;
  const GaugeMember q = gm[blockIdx.x];
  if (!q.Gb) return;
  gauge_bar_block(q.Gb, q.nodes, q.n, q.lam, q.thetaBar);
}

// ---- still depth from theta -------------------------------------------------------
__global__ void __launch_bounds__(kT)
k_face_depth(const double* __restrict__ th, int Nx, int Ny, int perX, int perY,
             double* __restrict__ HfX, double* __restrict__ HfY) {
  const long long nfx = (long long)(Nx + 1) * Ny * NP, nfy = (long long)Nx * (Ny + 1) * NP;
  const long long gid = (long long)blockIdx.x * kT + threadIdx.x;
  if (gid < nfx) {
    const long long q = gid / NP;
    const int n = (int)(gid - q * NP);
    const int ey = (int)(q / (Nx + 1)), l = (int)(q - (long long)ey * (Nx + 1));
    const long long row = (long long)ey * Nx;
    const double lo = th[(row + (l > 0 ? l - 1 : Nx - 1)) * NPP + NP - 1 + NP * n];
    const double hi = th[(row + (l < Nx ? l : 0)) * NPP + NP * n];
    double v;
    if ((l > 0 && l < Nx) || perX) v = 0.5 * (lo + hi);
    else v = l == 0 ? hi : lo;
    HfX[gid] = v;
  } else if (gid < nfx + nfy) {
    const long long k = gid - nfx;
    const long long q = k / NP;
    const int n = (int)(k - q * NP);
    const int l = (int)(q / Nx), ex = (int)(q - (long long)l * Nx);
    const double lo = th[((long long)(l > 0 ? l - 1 : Ny - 1) * Nx + ex) * NPP + n + NP * (NP - 1)];
    const double hi = th[((long long)(l < Ny ? l : 0) * Nx + ex) * NPP + n];
    double v;
    if ((l > 0 && l < Ny) || perY) v = 0.5 * (lo + hi);
    else v = l == 0 ? hi : lo;
    HfY[k] = v;
  }
}

// the gauge cotangents of one log row (one block)
__global__ void k_gauge_bar(const double* __restrict__ Gb, const long long* __restrict__ nodes,
                            int n, double* __restrict__ lam, double* __restrict__ thetaBar) {
  if (blockIdx.x != 0) return;
  gauge_bar_block(Gb, nodes, n, lam, thetaBar);
}

}  // namespace

hipError_t launch_rhs_vjp(const Geometry& g, const Phys& ph, const VjpArgs& va, GhostTime gt,
                          hipStream_t st) {
  const long long nb = (long long)g.Ny * ((g.Nx + kEW - 1) / kEW);
  if (!va.coefReady) {
    hipLaunchKernelGGL(k_vjp_coef, dim3((unsigned)nb), dim3(kVT), 0, st, va, g, ph);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_vjp_elem, dim3((unsigned)nb), dim3(kVT), 0, st, va, g, ph, gt);
  return hipGetLastError();
}

long long vjp_tiles(const Geometry& g) { return (long long)g.Ny * ((g.Nx + kEW - 1) / kEW); }

hipError_t launch_rhs_vjp_batch(const VjpMember* mem, const long long* tileStart, int nMem,
                                long long totalTiles, bool coefReady, hipStream_t st) {
  if (totalTiles <= 0) return hipSuccess;
  if (!coefReady) {
    hipLaunchKernelGGL(k_vjp_coef_batch, dim3((unsigned)totalTiles), dim3(kVT), 0, st, mem,
                       tileStart, nMem);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_vjp_elem_batch, dim3((unsigned)totalTiles), dim3(kVT), 0, st, mem,
                     tileStart, nMem);
  return hipGetLastError();
}

hipError_t launch_gauge_bar_batch(const GaugeMember* gm, int nMem, hipStream_t st) {
  if (nMem <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gauge_bar_batch, dim3((unsigned)nMem), dim3(64), 0, st, gm, nMem);
  return hipGetLastError();
}

hipError_t launch_face_depth(const Geometry& g, const double* theta, double* HfX, double* HfY,
                             hipStream_t st) {
  const long long n = (long long)(g.Nx + 1) * g.Ny * NP + (long long)g.Nx * (g.Ny + 1) * NP;
  hipLaunchKernelGGL(k_face_depth, dim3((unsigned)((n + kT - 1) / kT)), dim3(kT), 0, st, theta,
                     g.Nx, g.Ny, (int)(g.bc[0] == SGNX_BC_PERIODIC),
                     (int)(g.bc[2] == SGNX_BC_PERIODIC), HfX, HfY);
  return hipGetLastError();
}

hipError_t launch_gauge_bar(const double* Gb, const long long* nodes, int n, double* lam,
                            double* thetaBar, hipStream_t st) {
  hipLaunchKernelGGL(k_gauge_bar, dim3(1), dim3(64), 0, st, Gb, nodes, n, lam, thetaBar);
  return hipGetLastError();
}

}  // namespace sgnx
