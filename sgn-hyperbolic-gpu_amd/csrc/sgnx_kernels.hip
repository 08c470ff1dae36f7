// =============================================================================
// sgnx_kernels.hip -- gfx950 kernels of the FR/SGN right-hand side and of the
// Bogacki-Shampine 3(2) stage, update and error-norm passes.
//
// RHS (one launch, replaces the five phases of src/solver/sgn2d.cpp:125-290):
//   A workgroup owns a tile of TW consecutive elements of one element row
//   (25*TW nodes, one work-item per node; the state layout keeps each
//   variable's element blocks contiguous, so the tile is one coalesced span
//   per variable). Three phases separated by workgroup barriers:
//     0  load U, H at the node; primitives; nodal fluxes F, G -> LDS
//     1  one work-item per interface node of the tile (TW+1 x-lines, TW
//        bottom and TW top y-lines): Rusanov flux, central depth h*, and the
//        two one-sided fluxes, which are the elements' own face fluxes
//     2  one work-item per node: D.F and D.G from LDS, FR lift of the face
//        jumps, FR depth gradients, non-conservative products, source, dU
//   Face extrapolation on GLL points is the edge node itself: the reference's
//   sum 0 + 1*v0 + 0*v1 + ... equals v0 + 0.0 for finite values (the +0.0
//   turns -0 into +0 as that sum does), so it is never materialised.
//   Neighbour face nodes outside the tile are read from global memory; the
//   domain sides are periodic (wrapped index), walls (mirror ghost) or the
//   left inflow ghost.
//   Tiles are handed out XCD-contiguously: consecutive workgroup ids land on
//   different XCDs, so ids are remapped such that each XCD sweeps its own
//   contiguous band of tiles and finds the row below still in its L2.
//
// Every floating-point expression keeps the reference's evaluation order; the
// file is compiled with -ffp-contract=off.
// =============================================================================

#include "sgnx_internal.h"

namespace sgnx {

namespace {

constexpr int kThreads = 256;     // elementwise / reduction kernels

template <int TW>
struct Tile {
  static constexpr int NT = ((NPP * TW + 63) / 64) * 64;  // work-items (whole waves)
  static constexpr int NXI = NP * (TW + 1);               // x-interface nodes
  static constexpr int NYI = NP * TW;                     // y-interface nodes per side
};

// Node (a, b) of element (ex, ey), variable v (sgn2d.h:22-27, mesh2d.h:31).
__device__ __forceinline__ long long nidx(const Geometry& g, int ex, int ey, int a, int b) {
  return ((long long)ey * g.Nx + ex) * NPP + a + NP * b;
}

// Face state of a node: the GLL extrapolation value v + 0.0 (see header).
__device__ __forceinline__ void load_face(const double* __restrict__ U, long long Nd,
                                          long long node, double* W) {
#pragma unroll
  for (int v = 0; v < NV; ++v) W[v] = U[(long long)v * Nd + node] + 0.0;
}

// Compensated accumulation (Knuth's TwoSum): s + c carries the running sum
// with the rounding error of every addition kept in c.
__device__ __forceinline__ void acc_add(double& s, double& c, double x) {
  const double t = s + x;
  const double bp = t - s;
  const double e = (s - (t - bp)) + (x - bp);
  s = t;
  c = c + e;
}

// Fixed-order reductions of (s, c, m) triples: a compensated sum s + c and a
// max m. Within a wavefront, five butterfly levels of swizzles (lane ^ 1, 2,
// 4, 8, 16 inside each half) and lane 32's triple read into lane 0; then the
// waves' triples in wave order. Each level merges the partner's triple into
// the lane's own in the same way, so lane 0's result is one fixed tree and
// the same bits on every run.
template <int XOR>
__device__ __forceinline__ double swz_xor(double x) {
  constexpr int kPat = (XOR << 10) | 0x1F;   // bit mode: and 0x1f, or 0, xor XOR
  const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(x), kPat);
  const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(x), kPat);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ void merge3(double& s, double& c, double& m, double s2, double c2,
                                       double m2) {
  acc_add(s, c, s2);
  c = c + c2;
  m = fmax(m, m2);
}

template <int XOR>
__device__ __forceinline__ void wave_level(double& s, double& c, double& m) {
  const double s2 = swz_xor<XOR>(s), c2 = swz_xor<XOR>(c), m2 = swz_xor<XOR>(m);
  merge3(s, c, m, s2, c2, m2);
}

__device__ __forceinline__ double lane32(double x) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), 32);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), 32);
  return __hiloint2double(hi, lo);
}

// Every thread of the block calls it (no divergence); the result is in
// thread 0's s, c, m. ws: 3 doubles of LDS per wave.
__device__ __forceinline__ void block_reduce(double& s, double& c, double& m, double* ws) {
  wave_level<1>(s, c, m);
  wave_level<2>(s, c, m);
  wave_level<4>(s, c, m);
  wave_level<8>(s, c, m);
  wave_level<16>(s, c, m);
  const double s2 = lane32(s), c2 = lane32(c), m2 = lane32(m);
  merge3(s, c, m, s2, c2, m2);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    ws[3 * w] = s;
    ws[3 * w + 1] = c;
    ws[3 * w + 2] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) merge3(s, c, m, ws[3 * k], ws[3 * k + 1], ws[3 * k + 2]);
  }
}

// Ghost state across a non-periodic side.
__device__ __forceinline__ void wall_ghost(const double* In, int normal, double* G) {
#pragma unroll
  for (int v = 0; v < NV; ++v) G[v] = In[v];
  G[normal] = -In[normal];
}

template <int TW>
__global__ void __launch_bounds__(Tile<TW>::NT)
k_rhs(const double* __restrict__ U, double* __restrict__ dU, Geometry g, Phys ph, GhostTime gt,
      int tilesPerRow) {
  constexpr int NT = Tile<TW>::NT, NXI = Tile<TW>::NXI, NYI = Tile<TW>::NYI;
  __shared__ double sU[NV][NT];
  __shared__ double sF[NV][NT];
  __shared__ double sG[NV][NT];
  __shared__ double xFs[NV][NXI], xFl[NV][NXI], xFr[NV][NXI];
  __shared__ double xhs[NXI], xhl[NXI], xhr[NXI];
  __shared__ double yFsB[NV][NYI], yFrB[NV][NYI], yhsB[NYI], yhrB[NYI];
  __shared__ double yFsT[NV][NYI], yFlT[NV][NYI], yhsT[NYI], yhlT[NYI];

  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int ey = tile / tilesPerRow;
  const int ex0 = (tile - ey * tilesPerRow) * TW;
  const int nW = min(TW, g.Nx - ex0);
  const int nNodes = NPP * nW;
  const int t = threadIdx.x;
  const long long Nd = g.Nd;
  const long long gbase = ((long long)ey * g.Nx + ex0) * NPP;

  // ---- phase 0: nodal state and fluxes (sgn2d.cpp:188-196) -----------------
  double Un[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
  Prim q{0.0, 0.0, 0.0, 0.0};
  double Hn = 0.0;
  if (t < nNodes) {
#pragma unroll
    for (int v = 0; v < NV; ++v) Un[v] = U[(long long)v * Nd + gbase + t];
    Hn = g.Hnod[gbase + t];
    q = prims(Un);
    double F[NV], G[NV];
    flux_x(Un, q, Hn, ph, F);
    flux_y(Un, q, Hn, ph, G);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      sU[v][t] = Un[v];
      sF[v][t] = F[v];
      sG[v][t] = G[v];
    }
  }
  __syncthreads();

  // ---- phase 1: interface fluxes (sgn2d.cpp:198-239) -----------------------
  const int nXI = NP * (nW + 1);
  if (t < nXI) {
    const int i = t / NP, b = t - NP * i;
    const int ex = ex0 + i;                   // this interface is the left edge of column ex
    double UL[NV], UR[NV];
    bool ghostL = false, ghostR = false;
    if (i < nW) {
#pragma unroll
      for (int v = 0; v < NV; ++v) UR[v] = sU[v][i * NPP + NP * b] + 0.0;
    } else if (ex < g.Nx) {
      load_face(U, Nd, nidx(g, ex, ey, 0, b), UR);
    } else if (g.bc[1] == 0) {
      load_face(U, Nd, nidx(g, 0, ey, 0, b), UR);
    } else {
      ghostR = true;
    }
    if (i > 0) {
#pragma unroll
      for (int v = 0; v < NV; ++v) UL[v] = sU[v][(i - 1) * NPP + (NP - 1) + NP * b] + 0.0;
    } else if (ex > 0) {
      load_face(U, Nd, nidx(g, ex - 1, ey, NP - 1, b), UL);
    } else if (g.bc[0] == 0) {
      load_face(U, Nd, nidx(g, g.Nx - 1, ey, NP - 1, b), UL);
    } else {
      ghostL = true;
    }
    if (ghostL) {
      if (g.bc[0] == 1) {
        wall_ghost(UR, IHU, UL);
      } else {
        const double* c = g.inflow + ((long long)ey * NP + b) * (NV * 3);
#pragma unroll
        for (int v = 0; v < NV; ++v) UL[v] = (c[3 * v] + c[3 * v + 1] * gt.s) + c[3 * v + 2] * gt.c;
      }
    }
    if (ghostR) wall_ghost(UL, IHU, UR);
    const double Hf = g.HfX[((long long)ey * (g.Nx + 1) + ex) * NP + b];
    double Fs[NV], FL[NV], FR[NV];
    interface_flux<0>(UL, UR, Hf, ph, Fs, FL, FR);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      xFs[v][t] = Fs[v];
      xFl[v][t] = FL[v];
      xFr[v][t] = FR[v];
    }
    xhs[t] = 0.5 * (UL[IH] + UR[IH]);
    xhl[t] = UL[IH];
    xhr[t] = UR[IH];
  } else if (t < nXI + 2 * NP * nW) {
    const int r2 = t - nXI;
    const int side = r2 / (NP * nW);          // 0 bottom edge, 1 top edge
    const int r = r2 - side * NP * nW;
    const int j = r / NP, a = r - NP * j;
    const int ex = ex0 + j;
    double UL[NV], UR[NV];
    if (side == 0) {
#pragma unroll
      for (int v = 0; v < NV; ++v) UR[v] = sU[v][j * NPP + a] + 0.0;
      if (ey > 0) load_face(U, Nd, nidx(g, ex, ey - 1, a, NP - 1), UL);
      else if (g.bc[2] == 0) load_face(U, Nd, nidx(g, ex, g.Ny - 1, a, NP - 1), UL);
      else wall_ghost(UR, IHV, UL);
      const double Hf = g.HfY[((long long)ey * g.Nx + ex) * NP + a];
      double Fs[NV], FL[NV], FR[NV];
      interface_flux<1>(UL, UR, Hf, ph, Fs, FL, FR);
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        yFsB[v][r] = Fs[v];
        yFrB[v][r] = FR[v];
      }
      yhsB[r] = 0.5 * (UL[IH] + UR[IH]);
      yhrB[r] = UR[IH];
    } else {
#pragma unroll
      for (int v = 0; v < NV; ++v) UL[v] = sU[v][j * NPP + a + NP * (NP - 1)] + 0.0;
      if (ey + 1 < g.Ny) load_face(U, Nd, nidx(g, ex, ey + 1, a, 0), UR);
      else if (g.bc[3] == 0) load_face(U, Nd, nidx(g, ex, 0, a, 0), UR);
      else wall_ghost(UL, IHV, UR);
      const double Hf = g.HfY[((long long)(ey + 1) * g.Nx + ex) * NP + a];
      double Fs[NV], FL[NV], FR[NV];
      interface_flux<1>(UL, UR, Hf, ph, Fs, FL, FR);
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        yFsT[v][r] = Fs[v];
        yFlT[v][r] = FL[v];
      }
      yhsT[r] = 0.5 * (UL[IH] + UR[IH]);
      yhlT[r] = UL[IH];
    }
  }
  __syncthreads();

  // ---- phase 2: FR lift and assembly (sgn2d.cpp:241-288) -------------------
  if (t < nNodes) {
    const int j = t / NPP, loc = t - NPP * j;
    const int b = loc / NP, a = loc - NP * b;
    const OpTables* T = g.ops;
    double Da[NP], Db[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) { Da[k] = T->D[a * NP + k]; Db[k] = T->D[b * NP + k]; }
    const double gLa = T->gLp[a], gRa = T->gRp[a], gLb = T->gLp[b], gRb = T->gRp[b];
    const int ixL = j * NP + b, ixR = (j + 1) * NP + b, iy = j * NP + a;
    const int rowx = j * NPP + NP * b, coly = j * NPP + a;
    double out[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double dfx = 0.0, dgy = 0.0;
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        dfx = dfx + Da[k] * sF[v][rowx + k];
        dgy = dgy + Db[k] * sG[v][coly + NP * k];
      }
      const double cx = (xFs[v][ixL] - xFr[v][ixL]) * gLa + (xFs[v][ixR] - xFl[v][ixR]) * gRa;
      const double cy = (yFsB[v][iy] - yFrB[v][iy]) * gLb + (yFsT[v][iy] - yFlT[v][iy]) * gRb;
      const double dF = (dfx + cx) / g.Jx;
      const double dG = (dgy + cy) / g.Jy;
      out[v] = -dF - dG;
    }
    double dhx = 0.0, dhy = 0.0;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      dhx = dhx + Da[k] * sU[IH][rowx + k];
      dhy = dhy + Db[k] * sU[IH][coly + NP * k];
    }
    dhx = dhx + ((xhs[ixL] - xhr[ixL]) * gLa + (xhs[ixR] - xhl[ixR]) * gRa);
    dhy = dhy + ((yhsB[iy] - yhrB[iy]) * gLb + (yhsT[iy] - yhlT[iy]) * gRb);
    const double hx = dhx / g.Jx, hy = dhy / g.Jy;
    // ncpX / ncpY (sgn_flux2d.h:150-161): only the hp entry is non-zero, and
    // subtracting the zero entries leaves the other sums unchanged
    const double Bx = ph.mc2 * q.u * hx;
    const double By = ph.mc2 * q.v * hy;
    double S[NV];
    source(Un, q, Hn, g.Hx[gbase + t], g.Hy[gbase + t], ph, S);
    out[IHP] = out[IHP] - Bx - By;
#pragma unroll
    for (int v = 0; v < NV; ++v) dU[(long long)v * Nd + gbase + t] = out[v] + S[v];
  }
}

// ---- RHS, row sweep ---------------------------------------------------------
//
// A workgroup owns a strip of TW element columns and a chunk of `rows`
// element rows and walks up the chunk one element row r per iteration:
//   A  the row's nodes (loaded during the previous iteration; the loads of
//      row r+1 are issued here): stage input, primitives, nodal F and G,
//      source -> LDS / registers
//   B  interface items: the TW+1 x-interface lines of row r, and the
//      y-interface line r+1 on top of the row (the line at its bottom was
//      done in the previous iteration, or before the loop). A y-interface is
//      the x-interface arithmetic applied to the states with hu and hv
//      exchanged (fluxY(U) is fluxX of the exchanged state with the hu and hv
//      components exchanged back, and |v| is the exchanged |u|), so all items
//      follow one code path. Each interface stores the flux jumps
//      F* - F(left/lower state) and F* - F(right/upper state) and the depth
//      jumps h* - h, which are the jumps the reference forms for each of the
//      two elements (sgn2d.cpp:251-257, :268-275); y lines are double-buffered.
//      The operands an interface item takes from global memory (face depth,
//      halo node, the bottom face of row r+1) are also loaded one row ahead.
//   C  nodes: the complete RHS of row r (and, in the candidate stage, its
//      scaled embedded error).
// Every y-interface is evaluated once; the stage input is read once per chunk,
// plus the top faces of the row below the chunk and the x-halo nodes.
template <int TW>
struct Sweep {
  static constexpr int NT = ((NPP * TW + 63) / 64) * 64;
  static constexpr int NXI = NP * (TW + 1);   // x-interface nodes per row
  static constexpr int NYI = NP * TW;         // y-interface nodes per line
  static_assert(2 * (NXI + NYI) <= NT, "two lanes per interface node");
};

// Operands of the stage input at one node (5 variables each): u, and the
// increment k of u + a*k in the modes that have one.
template <int MODE>
struct StageRaw {
  static constexpr bool kHasK = MODE == kAxpy || MODE == kCand;
  double u[NV], k[kHasK ? NV : 1];
};
// Per-node operands beyond the stage input: kCand: x0 = kc; kErr: x0 = u0,
// x1 = z. Faces and halos never need them.
template <int MODE>
struct NodeExtra {
  double x0[(MODE == kCand || MODE == kErr) ? NV : 1], x1[MODE == kErr ? NV : 1];
};

// A five-variable operand by one base pointer: variable v of node n is
// base[n + v Nd] with a 32-bit index (the grid has at most 2^28 nodes per
// variable, so 5 Nd < 2^32). One base per operand instead of five plane
// pointers keeps the launch's uniform values within the scalar registers.
struct VarPtrs {
  const double* base;
  unsigned nd;
  __device__ __forceinline__ double at(int v, unsigned node) const {
    return base[node + (unsigned)v * nd];
  }
};
__device__ __forceinline__ VarPtrs var_ptrs(const double* base, long long Nd) {
  VarPtrs r;
  r.base = base;
  r.nd = (unsigned)Nd;
  return r;
}

struct StagePtrs {
  VarPtrs u, k, x0, x1;
};

template <int MODE>
__device__ __forceinline__ void stage_load(const StagePtrs& P, unsigned node, StageRaw<MODE>& R) {
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    R.u[v] = P.u.at(v, node);
    if (StageRaw<MODE>::kHasK) R.k[v] = P.k.at(v, node);
  }
}

template <int MODE>
__device__ __forceinline__ void extra_load(const StagePtrs& P, unsigned node, NodeExtra<MODE>& E) {
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    if (MODE == kCand || MODE == kErr) E.x0[v] = P.x0.at(v, node);
    if (MODE == kErr) E.x1[v] = P.x1.at(v, node);
  }
}

// rk.cpp:100, :102: the stage input u + (c dt) k from its operands (a = c dt)
template <int MODE>
__device__ __forceinline__ void stage_combine(const StageRaw<MODE>& R, double a, double* W) {
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    if (StageRaw<MODE>::kHasK) W[v] = R.u[v] + a * R.k[v];
    else W[v] = R.u[v];
  }
}

template <int MODE>
__device__ __forceinline__ void stage_face(const StagePtrs& P, double a, unsigned node,
                                           double* W) {
  StageRaw<MODE> R;
  stage_load<MODE>(P, node, R);
  stage_combine<MODE>(R, a, W);
#pragma unroll
  for (int v = 0; v < NV; ++v) W[v] = W[v] + 0.0;
}

// S[v] for a v that may not be a compile-time constant (selects, no
// indexed register access)
__device__ __forceinline__ double pick(const double* S, int v) {
  return (v == IH) ? S[IH] : (v == IHU) ? S[IHU] : (v == IHV) ? S[IHV] : (v == IHW) ? S[IHW] : S[IHP];
}

// Lane permutes within each group of four lanes (DPP quad_perm): the value
// of the even lane of the pair, of the odd lane, and of the other lane.
template <int CTRL>
__device__ __forceinline__ double lane_perm(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_even(double x) { return lane_perm<0xA0>(x); }  // [0,0,2,2]
__device__ __forceinline__ double lane_odd(double x) { return lane_perm<0xF5>(x); }   // [1,1,3,3]
__device__ __forceinline__ double lane_swap(double x) { return lane_perm<0xB1>(x); }  // [1,0,3,2]

__device__ __forceinline__ unsigned node_of(int Nx, int ex, int ey, int a, int b) {
  return ((unsigned)ey * (unsigned)Nx + (unsigned)ex) * NPP + a + NP * b;
}

// One workgroup's strip x chunk `tile` of one grid (the body of the plain and
// the batched launches).
template <int TW, int MODE, int WAVES, bool KEEP = false, bool HXR = false>
__device__ __forceinline__ void sweep_tile(const double* __restrict__ U, double* __restrict__ dU,
                                           const Geometry& g, const Phys& phIn, const GhostTime& gt,
                                           const Stage& sa, int nStrips, int tile) {
  constexpr int NT = Sweep<TW>::NT, NXI = Sweep<TW>::NXI, NYI = Sweep<TW>::NYI;
  // the row sweep runs without the Manning friction term (the launchers
  // refuse it; solvers with friction use the element-row tiles), so its
  // branch is compiled out
  Phys ph = phIn;
  ph.nManning = 0.0;
  __shared__ double sU[NV][NT];
  __shared__ double sF[NV][NT];
  __shared__ double sG[NV][NT];
  // interface jumps by side s (0 = left / lower, 1 = right / upper):
  // F* - F(side s) and h* - h(side s); y lines double-buffered
  __shared__ double xJ[2][NV][NXI], xhJ[2][NXI];
  __shared__ double yJ[2][2][NV][NYI], yhJ[2][2][NYI];
  __shared__ double sD[NPP], sgL[NP], sgR[NP];
  // HXR: the still depth at the row's nodes, from which phase C forms Hx, Hy
  // (sgn2d.cpp:54-71) instead of reading them
  __shared__ double sH[HXR ? NT : 1];

  if (threadIdx.x < NPP) sD[threadIdx.x] = g.ops->D[threadIdx.x];
  if (threadIdx.x < NP) {
    sgL[threadIdx.x] = g.ops->gLp[threadIdx.x];
    sgR[threadIdx.x] = g.ops->gRp[threadIdx.x];
  }

  const int chunk = tile / nStrips;
  const int ex0 = (tile - chunk * nStrips) * TW;
  const int nW = min(TW, g.Nx - ex0);
  const int ey0 = chunk * g.rows;
  const int ey1 = min(g.Ny, ey0 + g.rows);
  const int nNodes = NPP * nW;
  const int t = threadIdx.x;
  const long long Nd = g.Nd;
  const int Nx = g.Nx, Ny = g.Ny;
  StagePtrs P;
  P.u = var_ptrs(MODE == kPlain ? U : sa.u, Nd);
  P.k = var_ptrs(sa.k1, Nd);
  P.x0 = var_ptrs(MODE == kCand ? sa.kc : sa.u0, Nd);
  P.x1 = var_ptrs(sa.z, Nd);

  // node work item: node (a, b) of element j of the strip
  const int j = t / NPP, loc = t - NPP * j;
  const int b = loc / NP, a = loc - NP * b;
  const bool isNode = t < nNodes;
  // Interface half-items: two lanes per interface node, lane 2k + side
  // taking one side (0 = left / lower, 1 = right / upper) of interface k.
  // Items 0 .. 2 nX - 1 are the x-interface nodes of the row (k = i NP + bb:
  // line i on the left edge of strip column i, node bb); the next 2 NP nW items
  // the y-interface nodes of the line above the row (k = nX + jj NP + aa).
  // Each lane evaluates its own side (primitives, one-sided flux, wave speed)
  // and the pair exchanges them with lane permutes, so the interface work is
  // spread over every wave of the workgroup.
  const int nX = NP * (nW + 1);
  const int side = t & 1, kk = t >> 1;
  const bool isX = t < 2 * nX;
  const bool isY = !isX && t < 2 * nX + 2 * NP * nW;
  const int xi = isX ? kk / NP : 0, xb = isX ? kk - NP * xi : 0;
  const int ky = isY ? kk - nX : 0;
  const int jj = ky / NP, aa = ky - NP * jj;

  // The global node this lane's side takes for row r, if any: the halo node
  // of the strip's outer x-lines, or the bottom face of row r+1 for the upper
  // side of a y item (false at a wall, where a ghost replaces it, and for
  // every side that lies in the workgroup's row).
  auto face_node = [&](int r, unsigned& n) -> bool {
    if (isX) {
      if (side == 0 && xi == 0) {
        if (ex0 > 0) { n = node_of(Nx, ex0 - 1, r, NP - 1, xb); return true; }
        if (g.bc[0] == 0) { n = node_of(Nx, Nx - 1, r, NP - 1, xb); return true; }
        return false;
      }
      if (side == 1 && xi == nW) {
        const int ex = ex0 + nW;
        if (ex < Nx) { n = node_of(Nx, ex, r, 0, xb); return true; }
        if (g.bc[1] == 0) { n = node_of(Nx, 0, r, 0, xb); return true; }
        return false;
      }
      return false;
    }
    if (side == 0) return false;
    const int y = r + 1;
    if (y < Ny) { n = node_of(Nx, ex0 + jj, y, aa, 0); return true; }
    if (g.bc[3] == 0) { n = node_of(Nx, ex0 + jj, 0, aa, 0); return true; }
    return false;
  };

  // One interface node evaluated by a lane pair. W is this lane's side's
  // state in the interface frame (a y interface has hu and hv exchanged:
  // fluxY(U) is fluxX of the exchanged state with hu and hv of the result
  // exchanged back, and |v| is the exchanged |u|). Returns this side's flux
  // jump F* - F(W) (frame order) and depth jump h* - h(W), with the Rusanov
  // flux of sgn_flux2d.h:169-179: F* = 0.5 (FL + FR) - 0.5 max(sL, sR) (UR - UL).
  auto half_iface = [&](const double* W, double Hf, double* jf, double& jh) {
    const Prim q = prims(W);
    double F[NV];
    flux_x(W, q, Hf, ph, F);
    const double so = fabs(q.u) + signal_speed(W[IH], q, ph);
    const double lam = fmax(lane_even(so), lane_odd(so));
    const double hl = 0.5 * lam;
    double hL = 0.0, hR = 0.0;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double ul = lane_even(W[v]), ur = lane_odd(W[v]);
      // FL + FR in either lane order (addition commutes exactly)
      const double Fs = 0.5 * (F[v] + lane_swap(F[v])) - hl * (ur - ul);
      jf[v] = Fs - F[v];
      if (v == IH) { hL = ul; hR = ur; }
    }
    const double hs = 0.5 * (hL + hR);
    jh = hs - W[IH];
  };
  auto to_frame = [&](double* W) {             // exchange hu, hv on y lanes
    const double u_ = W[IHU], v_ = W[IHV];
    W[IHU] = isY ? v_ : u_;
    W[IHV] = isY ? u_ : v_;
  };

  // prologue: the y line at the bottom of the chunk, both sides from global
  // memory (the bottom wall's ghost mirrors the upper side)
  if (isY) {
    double W[NV];
    const int y = ey0;
    const unsigned nUp = node_of(Nx, ex0 + jj, y, aa, 0);
    if (side == 1) {
      stage_face<MODE>(P, sa.a, nUp, W);
    } else if (y > 0) {
      stage_face<MODE>(P, sa.a, node_of(Nx, ex0 + jj, y - 1, aa, NP - 1), W);
    } else if (g.bc[2] == 0) {
      stage_face<MODE>(P, sa.a, node_of(Nx, ex0 + jj, Ny - 1, aa, NP - 1), W);
    } else {
      double Wu[NV];
      stage_face<MODE>(P, sa.a, nUp, Wu);
      wall_ghost(Wu, IHV, W);
    }
    to_frame(W);
    const double Hf = g.HfY[((long long)y * Nx + ex0 + jj) * NP + aa];
    double jf[NV], jh;
    half_iface(W, Hf, jf, jh);
    to_frame(jf);
#pragma unroll
    for (int v = 0; v < NV; ++v) yJ[y & 1][side][v][ky] = jf[v];
    yhJ[y & 1][side][ky] = jh;
  }

  auto face_node_now = [&](int r) -> bool {
    unsigned n;
    return face_node(r, n);
  };
  // The interface lanes' global operands of row r + 1: the halo nodes and
  // the bottom face of row r + 2. The two-wave stage launches issue them
  // after their phase A of row r. The plain launch and the three-wave stage-2
  // launch issue them at the start of phase C, which leaves phase C and the
  // next phase A to cover their latency, brings the full loads of row r + 2
  // (issued after that phase A) close enough behind that fewer bottom-face
  // lines have left L2, and shortens their register lifetime: the plain
  // launch measured 4 % faster, and the stage-2 build fits three waves per
  // SIMD (with Hx, Hy read) with 3 registers spilled instead of 15. The
  // two-wave stage launches were not faster this way.
  constexpr bool kLateFace = MODE == kPlain || (MODE == kAxpy && WAVES == 3);
  StageRaw<MODE> fR;
  double fH = 0.0;
  auto face_prefetch = [&](int r) {
    if (isX || isY) {
      unsigned n;
      if (face_node(r, n)) stage_load<MODE>(P, n, fR);
      fH = isX ? g.HfX[((long long)r * (Nx + 1) + ex0 + xi) * NP + xb]
               : g.HfY[((long long)(r + 1) * Nx + ex0 + jj) * NP + aa];
    }
  };
  face_prefetch(ey0);

  // loads of row ey0's node data
  StageRaw<MODE> pR;
  NodeExtra<MODE> pE;
  double pH = 0.0, pHx = 0.0, pHy = 0.0;
  if (isNode) {
    const unsigned gn = node_of(Nx, ex0, ey0, 0, 0) + t;
    stage_load<MODE>(P, gn, pR);
    extra_load<MODE>(P, gn, pE);
    pH = g.Hnod[gn];
    if (!HXR) { pHx = g.Hx[gn]; pHy = g.Hy[gn]; }
  }
  // scaled embedded error of this work-item's nodes (kErr)
  double eS = 0.0, eC = 0.0, eM = 0.0;

  for (int r = ey0; r < ey1; ++r) {
    const unsigned rowBase = node_of(Nx, ex0, r, 0, 0);
    // ---- A ---------------------------------------------------------------------
    double S[NV], qu = 0.0, qv = 0.0, qw = 0.0;
    SrcPart sp{0.0, 0.0, 0.0};
    // per-node operands phase C needs beyond the RHS (rk.cpp:106-108, :119-120,
    // sums evaluated left to right as there):
    //   kCand: c0 = u, c1 = 2/9 kc + 1/3 k, c2 = 7/24 kc + 0.25 k
    //   kErr:  c0 = u0, c2 = z
    double c0[NV], c1[NV], c2[NV];
    // the interface lane's global node state and face depth for this row
    // (the plain launch takes them at the start of phase B instead)
    double SG[NV];
    double Hf = 0.0;
    const bool haveG = (isX || isY) && face_node_now(r);
    if (!kLateFace && (isX || isY)) {
      Hf = fH;
      if (haveG) {
        stage_combine<MODE>(fR, sa.a, SG);
#pragma unroll
        for (int v = 0; v < NV; ++v) SG[v] = SG[v] + 0.0;
      }
    }
    if (isNode) {
      double Un[NV];
      stage_combine<MODE>(pR, sa.a, Un);
      if (KEEP && sa.yOut) {
        // the stage input, kept for a later reverse sweep
#pragma unroll
        for (int v = 0; v < NV; ++v) sa.yOut[rowBase + t + (unsigned)v * (unsigned)Nd] = Un[v];
      }
      const double Hn = pH, Hxn = pHx, Hyn = pHy;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        if (MODE == kCand) {
          c0[v] = pR.u[v];
          c1[v] = 2.0 / 9.0 * pE.x0[v] + 1.0 / 3.0 * pR.k[v];
          c2[v] = 7.0 / 24.0 * pE.x0[v] + 0.25 * pR.k[v];
        } else if (MODE == kErr) {
          c0[v] = pE.x0[v];
          c2[v] = pE.x1[v];
        }
      }
      const Prim q = prims(Un);
      qu = q.u;
      qv = q.v;
      if (HXR) {
        // the gradient terms wait for phase C
        sp = source_part(Un, q, Hn, ph);
        qw = q.w;
        sH[t] = Hn;
      } else {
        source(Un, q, Hn, Hxn, Hyn, ph, S);
      }
      double F[NV], G[NV];
      flux_x(Un, q, Hn, ph, F);
      flux_y(Un, q, Hn, ph, G);
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        sU[v][t] = Un[v];
        sF[v][t] = F[v];
        sG[v][t] = G[v];
      }
    }
    // row r+1's node operands, consumed one iteration later. Issued after
    // the last use of this row's, and fenced against the scheduler hoisting
    // them above those uses: the loads then land in the registers the loop
    // carries, so no copy (and no wait for them) is needed at the back-edge.
    __builtin_amdgcn_sched_barrier(0);
    if (!kLateFace && r + 1 < ey1) face_prefetch(r + 1);
    if (isNode && r + 1 < ey1) {
      const unsigned gn = rowBase + (unsigned)Nx * NPP + t;
      stage_load<MODE>(P, gn, pR);
      extra_load<MODE>(P, gn, pE);
      pH = g.Hnod[gn];
      if (!HXR) { pHx = g.Hx[gn]; pHy = g.Hy[gn]; }
    }
    __syncthreads();

    // ---- B ---------------------------------------------------------------------
    if (isX || isY) {
      double W[NV];
      if (kLateFace) {
        Hf = fH;
        if (haveG) {
          stage_combine<MODE>(fR, sa.a, W);
#pragma unroll
          for (int v = 0; v < NV; ++v) W[v] = W[v] + 0.0;
        }
      } else if (haveG) {
#pragma unroll
        for (int v = 0; v < NV; ++v) W[v] = SG[v];
      }
      if (isX) {
        // own side: a node of this row, the halo node, or a ghost of the
        // other side's node (a wall mirrors it; the inflow side is the table)
        const int li = (side == 0) ? xi - 1 : xi;         // strip column of the node
        const bool inRow = (side == 0) ? xi > 0 : xi < nW;
        if (inRow) {
          const int n = li * NPP + (side == 0 ? NP - 1 : 0) + NP * xb;
#pragma unroll
          for (int v = 0; v < NV; ++v) W[v] = sU[v][n] + 0.0;
        } else if (!haveG) {
          if (side == 0 && g.bc[0] == 2) {
            const double* c = g.inflow + ((long long)r * NP + xb) * (NV * 3);
#pragma unroll
            for (int v = 0; v < NV; ++v)
              W[v] = (c[3 * v] + c[3 * v + 1] * gt.s) + c[3 * v + 2] * gt.c;
          } else {
            const int n = (side == 0) ? NP * xb : (nW - 1) * NPP + (NP - 1) + NP * xb;
            double Wo[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) Wo[v] = sU[v][n] + 0.0;
            wall_ghost(Wo, IHU, W);
          }
        }
      } else {
        const int n = jj * NPP + aa + NP * (NP - 1);    // top face of row r
        if (side == 0 || !haveG) {
          double Wo[NV];
#pragma unroll
          for (int v = 0; v < NV; ++v) Wo[v] = sU[v][n] + 0.0;
          if (side == 0) {
#pragma unroll
            for (int v = 0; v < NV; ++v) W[v] = Wo[v];
          } else {
            wall_ghost(Wo, IHV, W);                   // the top wall
          }
        }
      }
      to_frame(W);
      double jf[NV], jh;
      half_iface(W, Hf, jf, jh);
      to_frame(jf);
      if (isX) {
#pragma unroll
        for (int v = 0; v < NV; ++v) xJ[side][v][xi * NP + xb] = jf[v];
        xhJ[side][xi * NP + xb] = jh;
      } else {
        const int buf = (r + 1) & 1;
#pragma unroll
        for (int v = 0; v < NV; ++v) yJ[buf][side][v][ky] = jf[v];
        yhJ[buf][side][ky] = jh;
      }
    }
    __syncthreads();

    // ---- C: dU of row r ---------------------------------------------------------
    if (kLateFace) {
      if (r + 1 < ey1) face_prefetch(r + 1);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (isNode) {
      double Da[NP], Db[NP];
#pragma unroll
      for (int k = 0; k < NP; ++k) { Da[k] = sD[a * NP + k]; Db[k] = sD[b * NP + k]; }
      const double gLa = sgL[a], gRa = sgR[a], gLb = sgL[b], gRb = sgR[b];
      const int ixL = j * NP + b, ixR = (j + 1) * NP + b, iy = j * NP + a;
      const int bot = r & 1, top = (r + 1) & 1;
      const int rowx = j * NPP + NP * b, coly = j * NPP + a;
      double dhx = 0.0, dhy = 0.0, dHx = 0.0, dHy = 0.0;
      // the three-wave stage-2 build keeps this loop rolled: unrolled, its
      // hoisted LDS operands cost the 3 registers it otherwise spills (1 %
      // faster without the spill; rolled, the other builds ran 3-8 % slower)
      constexpr int kDepthUnroll = (WAVES == 3 && MODE == kAxpy) ? 1 : NP;
#pragma unroll kDepthUnroll
      for (int k = 0; k < NP; ++k) {
        dhx = dhx + Da[k] * sU[IH][rowx + k];
        dhy = dhy + Db[k] * sU[IH][coly + NP * k];
        if (HXR) {
          // sgn2d.cpp:54-71: (D H) / J along the node's x and y lines
          dHx = dHx + Da[k] * sH[rowx + k];
          dHy = dHy + Db[k] * sH[coly + NP * k];
        }
      }
      if (HXR) source_rest(sp, qu, qv, qw, div_const(dHx, g.dx), div_const(dHy, g.dy), ph, S);
      dhx = dhx + (xhJ[1][ixL] * gLa + xhJ[0][ixR] * gRa);
      dhy = dhy + (yhJ[bot][1][iy] * gLb + yhJ[top][0][iy] * gRb);
      const double hx = div_const(dhx, g.dx), hy = div_const(dhy, g.dy);
      // the hp entries of the non-conservative products (the others are zero,
      // and subtracting zero changes no sum)
      const double Bx = ph.mc2 * qu * hx;
      const double By = ph.mc2 * qv * hy;
      const unsigned gn = rowBase + t;
      auto variable = [&](int v) {
        double dfx = 0.0, dgy = 0.0;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
          dfx = dfx + Da[k] * sF[v][rowx + k];
          dgy = dgy + Db[k] * sG[v][coly + NP * k];
        }
        const double cx = xJ[1][v][ixL] * gLa + xJ[0][v][ixR] * gRa;
        const double cy = yJ[bot][1][v][iy] * gLb + yJ[top][0][v][iy] * gRb;
        double o = -div_const(dfx + cx, g.dx) - div_const(dgy + cy, g.dy);
        if (v == IHP) o = o - Bx - By;
        const double k4 = o + pick(S, v);
        const unsigned vi = gn + (unsigned)v * (unsigned)Nd;
        if (MODE == kCand) {
          // rk.cpp:106-108 and the first three terms of :119-120 (k4 is k3 here)
          sa.uNew[vi] = pick(c0, v) + sa.dt * (pick(c1, v) + 4.0 / 9.0 * k4);
          sa.zOut[vi] = pick(c2, v) + 1.0 / 3.0 * k4;
        } else {
          dU[vi] = k4;
        }
        if (MODE == kAxpy && KEEP && sa.y3Out) {
          // the stage-3 input u + (3/4 dt) k2 of a recomputed step
          // (rk.cpp:104), formed as the elementwise combination forms it
          sa.y3Out[vi] = P.u.at(v, gn) + sa.a3 * k4;
        }
        if (MODE == kErr) {
          // rk.cpp:119-129 for this entry
          const double u0 = pick(c0, v), un = sU[v][t];
          const double uEmb = u0 + sa.dt * (pick(c2, v) + 0.125 * k4);
          const double e = un - uEmb;
          const double au = fabs(u0), an = fabs(un);
          const double r_ = e / (sa.atol + sa.rtol * ((au < an) ? an : au));
          acc_add(eS, eC, r_ * r_);
          const double ar = fabs(r_);
          eM = (eM < ar) ? ar : eM;
        }
      };
      if (WAVES == 3 && MODE == kAxpy) {
        // one variable per trip keeps the LDS operands of a single variable
        // live (the register budget of three waves per SIMD)
#pragma unroll 1
        for (int v = 0; v < NV; ++v) variable(v);
      } else {
        // unrolled, so that the per-node operands stay in registers without
        // selects. Scheduling fences bound how far the LDS operands of later
        // variables are hoisted: the stage builds take two groups, (the depth
        // gradient, h, hu) and (hv, hw, hp), which lets one variable's LDS
        // reads overlap another's arithmetic within their two-wave register
        // budget (k4 2.5 %, stage 3 1.5 %, the two-wave stage 2 2.4 % faster,
        // same bits); the plain build keeps the depth gradient and each
        // variable in a group of its own (grouped, it ran slower)
        constexpr bool kGrouped = MODE != kPlain;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          if (kGrouped ? v == 2 : true) __builtin_amdgcn_sched_barrier(0);
          variable(v);
        }
      }
    }
    __syncthreads();
  }

  if (MODE == kErr) {
    // fixed-order reduction over the workgroup (LDS of the last row reused)
    block_reduce(eS, eC, eM, &sF[0][0]);
    if (t == 0) {
      sa.part[3 * tile] = eS;
      sa.part[3 * tile + 1] = eC;
      sa.part[3 * tile + 2] = eM;
    }
  }
}

template <int TW, int MODE, int WAVES, bool KEEP = false, bool HXR = false>
__global__ void __launch_bounds__(Sweep<TW>::NT, WAVES)
k_rhs_sweep(const double* __restrict__ U, double* __restrict__ dU, Geometry g, Phys ph,
            GhostTime gt, Stage sa, int nStrips) {
  sweep_tile<TW, MODE, WAVES, KEEP, HXR>(U, dU, g, ph, gt, sa, nStrips,
                                         xcd_tile(blockIdx.x, gridDim.x));
}

// Many grids in one launch: the workgroups of every member's tiles, packed
// back to back (tileStart[m] = first packed index of member m, ascending);
// each member keeps its own tile decomposition, so its error partials, and
// everything else, equal its own launch's.
template <int TW, int MODE, int WAVES, bool KEEP = false, bool HXR = false>
__global__ void __launch_bounds__(Sweep<TW>::NT, WAVES)
k_rhs_sweep_batch(const BatchMember* __restrict__ mem, const long long* __restrict__ tileStart,
                  int nMem) {
  const long long w = xcd_tile(blockIdx.x, gridDim.x);
  int lo = 0, hi = nMem - 1;                  // last member with tileStart <= w
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tileStart[mid] <= w) lo = mid; else hi = mid - 1;
  }
  const BatchMember& M = mem[lo];
  const long long tile = w - tileStart[lo];
  // a grid larger than the members' tiles (never launched so) does nothing
  if (tile >= M.nTiles) return;
  sweep_tile<TW, MODE, WAVES, KEEP, HXR>(M.U, M.dU, M.g, M.ph, M.gt, M.sa, M.nStrips, (int)tile);
}

// Hx, Hy at every node (sgn2d.cpp:54-71).
__global__ void __launch_bounds__(kThreads)
k_bathy_grad(const double* __restrict__ H, double* __restrict__ Hx, double* __restrict__ Hy,
             const OpTables* __restrict__ T, long long Nd, double Jx, double Jy) {
  const long long gn = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (gn >= Nd) return;
  const long long e = gn / NPP;
  const int loc = (int)(gn - e * NPP);
  const int b = loc / NP, a = loc - NP * b;
  const double* He = H + e * NPP;
  double dHx = 0.0, dHy = 0.0;
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    dHx = dHx + T->D[a * NP + k] * He[k + NP * b];
    dHy = dHy + T->D[b * NP + k] * He[a + NP * k];
  }
  Hx[gn] = dHx / Jx;
  Hy[gn] = dHy / Jy;
}

// Whether Hx, Hy hold (D Hnod) / J (sgn2d.cpp:54-71, the arithmetic of
// k_bathy_grad) bit for bit at every node: any difference sets *bad.
__global__ void __launch_bounds__(kThreads)
k_bathy_check(const double* __restrict__ H, const double* __restrict__ Hx,
              const double* __restrict__ Hy, const OpTables* __restrict__ T, long long Nd,
              double Jx, double Jy, int* bad) {
  const long long gn = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (gn >= Nd) return;
  const long long e = gn / NPP;
  const int loc = (int)(gn - e * NPP);
  const int b = loc / NP, a = loc - NP * b;
  const double* He = H + e * NPP;
  double dHx = 0.0, dHy = 0.0;
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    dHx = dHx + T->D[a * NP + k] * He[k + NP * b];
    dHy = dHy + T->D[b * NP + k] * He[a + NP * k];
  }
  if (__double_as_longlong(dHx / Jx) != __double_as_longlong(Hx[gn]) ||
      __double_as_longlong(dHy / Jy) != __double_as_longlong(Hy[gn]))
    atomicOr(bad, 1);
}

// ---- RK stage arithmetic (src/solver/rk.cpp) ----------------------------------

__global__ void __launch_bounds__(kThreads)
k_comb1(double* __restrict__ out, const double* __restrict__ u, double a1,
        const double* __restrict__ k1, size_t n) {
  const size_t stride = (size_t)gridDim.x * kThreads;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
    out[i] = u[i] + a1 * k1[i];
}

__global__ void __launch_bounds__(kThreads)
k_bs3_update(double* __restrict__ out, const double* __restrict__ u, double dt,
             const double* __restrict__ k1, const double* __restrict__ k2,
             const double* __restrict__ k3, size_t n) {
  const size_t stride = (size_t)gridDim.x * kThreads;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
    out[i] = u[i] + dt * (2.0 / 9.0 * k1[i] + 1.0 / 3.0 * k2[i] + 4.0 / 9.0 * k3[i]);
}

__global__ void __launch_bounds__(kThreads)
k_rk4_update(double* __restrict__ u, double dt, const double* __restrict__ k1,
             const double* __restrict__ k2, const double* __restrict__ k3,
             const double* __restrict__ k4, size_t n) {
  const size_t stride = (size_t)gridDim.x * kThreads;
  const double d6 = dt / 6.0;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
    u[i] += d6 * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
}

// rk.cpp:118-131: r = (uNew - uEmb) / (atol + rtol*max(|u|, |uNew|)).
__global__ void __launch_bounds__(kThreads)
k_err_norm(const double* __restrict__ u, const double* __restrict__ uNew,
           const double* __restrict__ k1, const double* __restrict__ k2,
           const double* __restrict__ k3, const double* __restrict__ k4, double dt, double atol,
           double rtol, size_t n, double* __restrict__ part) {
  __shared__ double ws[3 * kThreads / 64];
  double s = 0.0, c = 0.0, m = 0.0;
  const size_t stride = (size_t)gridDim.x * kThreads;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    const double uEmb =
        u[i] + dt * (7.0 / 24.0 * k1[i] + 0.25 * k2[i] + 1.0 / 3.0 * k3[i] + 0.125 * k4[i]);
    const double e = uNew[i] - uEmb;
    const double au = fabs(u[i]), an = fabs(uNew[i]);
    const double sc_ = atol + rtol * ((au < an) ? an : au);   // std::max(|u|, |uNew|)
    const double r = e / sc_;
    acc_add(s, c, r * r);
    const double ar = fabs(r);
    m = (m < ar) ? ar : m;                                     // std::max(errMax, |r|)
  }
  block_reduce(s, c, m, ws);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = s;
    part[3 * blockIdx.x + 1] = c;
    part[3 * blockIdx.x + 2] = m;
  }
}

// Second level: one block folds the nb partial triples in a fixed order.
// out[0] = sum, out[1] = max.
__device__ __forceinline__ void finish_block(const double* __restrict__ part, int nb,
                                             double* __restrict__ out) {
  __shared__ double ws[3 * kThreads / 64];
  double s = 0.0, c = 0.0, m = 0.0;
  for (int i = threadIdx.x; i < nb; i += kThreads) {
    acc_add(s, c, part[3 * i]);
    c = c + part[3 * i + 1];
    m = fmax(m, part[3 * i + 2]);
  }
  block_reduce(s, c, m, ws);
  if (threadIdx.x == 0) {
    out[0] = s + c;
    out[1] = m;
  }
}

__global__ void __launch_bounds__(kThreads)
k_finish(const double* __restrict__ part, int nb, double* __restrict__ out) {
  finish_block(part, nb, out);
}

// k_finish writing straight into host memory mapped for the device: the
// (sum, max) pair, then, after a system-scope fence, the attempt's sequence
// number, for which the host polls.
__global__ void __launch_bounds__(kThreads)
k_finish_post(const double* __restrict__ part, int nb, double* __restrict__ host,
              unsigned long long seq) {
  __shared__ double pair[2];
  finish_block(part, nb, pair);
  if (threadIdx.x == 0) {
    host[0] = pair[0];
    host[1] = pair[1];
    __threadfence_system();
    reinterpret_cast<volatile unsigned long long*>(host)[2] = seq;
  }
}

// one block per member: member m's partials into out[2m], out[2m+1]
__global__ void __launch_bounds__(kThreads)
k_finish_batch(const BatchMember* __restrict__ mem, int nMem, double* __restrict__ out) {
  const BatchMember& M = mem[blockIdx.x];
  finish_block(M.sa.part, M.nTiles, out + 2 * blockIdx.x);
}

// Max of eta = h - H over the nodes inside the open window (x0, x1) x (y0, y1)
// and the lowest node index attaining it (the result of a sequential scan with
// a strict ">", apps/shoal_focus_mpi.cu:199-204): per block, (value, index)
// partials in part[2 b], part[2 b + 1]; empty: (-1e300, -1). Node coordinates
// are formed as the host forms them, ex hx + xo[a] with xo[a] = 0.5 hx (xi_a + 1).
__global__ void __launch_bounds__(kThreads)
k_eta_max(const double* __restrict__ U, const double* __restrict__ Hn, Geometry g, double hx,
          double hy, EtaWindow w, double* __restrict__ part) {
  __shared__ double sv[kThreads];
  __shared__ long long si[kThreads];
  double best = -1e300;
  long long bi = -1;
  const long long Nd = g.Nd;
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < Nd; i += stride) {
    const long long e = i / NPP;
    const int loc = (int)(i - e * NPP);
    const int b = loc / NP, a = loc - NP * b;
    const int ey = (int)(e / g.Nx), ex = (int)(e - (long long)ey * g.Nx);
    const double x = ex * hx + w.xo[a];
    const double y = ey * hy + w.yo[b];
    if (!(x > w.x0 && x < w.x1 && y > w.y0 && y < w.y1)) continue;
    const double eta = U[i] - Hn[i];
    if (eta > best) { best = eta; bi = i; }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      const double v2 = sv[threadIdx.x + off];
      const long long i2 = si[threadIdx.x + off];
      const double v1 = sv[threadIdx.x];
      const long long i1 = si[threadIdx.x];
      if (v2 > v1 || (v2 == v1 && i2 >= 0 && (i1 < 0 || i2 < i1))) {
        sv[threadIdx.x] = v2;
        si[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sv[0];
    part[2 * blockIdx.x + 1] = (double)si[0];
  }
}

// eta = h - H at n given nodes
__global__ void __launch_bounds__(kThreads)
k_gather_eta(const double* __restrict__ U, const double* __restrict__ Hn,
             const long long* __restrict__ nodes, int n, double* __restrict__ out) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k < n) out[k] = U[nodes[k]] - Hn[nodes[k]];
}

// Element means of eta, u, v, p, H (src/io/vtk2d.h:39-44, :67-81): for each
// field, 0.25 * sum over b, a of qw[a] qw[b] f at the element's nodes, in that
// order; out[f * Ne + e].
__global__ void __launch_bounds__(kThreads)
k_elem_means(const double* __restrict__ U, const double* __restrict__ Hn,
             const OpTables* __restrict__ T, long long Nd, long long Ne,
             double* __restrict__ out) {
  const long long e = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (e >= Ne) return;
  double m[5];
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    double s = 0.0;
    for (int b = 0; b < NP; ++b)
      for (int a = 0; a < NP; ++a) {
        const long long i = e * NPP + a + NP * b;
        const double h = U[i];
        double v;
        if (f == 0) v = h - Hn[i];
        else if (f == 1) v = U[(long long)IHU * Nd + i] / h;
        else if (f == 2) v = U[(long long)IHV * Nd + i] / h;
        else if (f == 3) v = U[(long long)IHP * Nd + i] / h;
        else v = Hn[i];
        s += T->qw[a] * T->qw[b] * v;
      }
    m[f] = 0.25 * s;
  }
#pragma unroll
  for (int f = 0; f < 5; ++f) out[f * Ne + e] = m[f];
}

// Diagnostics over nodes: 0 = max wave speed (sgn2d.cpp:295-304), 1 = mass
// integrand qw_a qw_b h (:319-332), 2 = energy integrand (:334-350).
__global__ void __launch_bounds__(kThreads)
k_diag(const double* __restrict__ U, Geometry g, Phys ph, int which, double* __restrict__ part) {
  __shared__ double ws[3 * kThreads / 64];
  double s = 0.0, c = 0.0, m = 0.0;
  const long long Nd = g.Nd;
  const size_t stride = (size_t)gridDim.x * kThreads;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < (size_t)Nd; i += stride) {
    double Un[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) Un[v] = U[(long long)v * Nd + i];
    if (which == 0) {
      const Prim q = prims(Un);
      const double ce = signal_speed(Un[IH], q, ph);
      const double s0 = fabs(q.u) + ce, s1 = fabs(q.v) + ce;
      m = fmax(m, fmax(s0, s1));
    } else {
      const int loc = (int)(i % NPP);
      const int b = loc / NP, a = loc - NP * b;
      const double wq = g.ops->qw[a] * g.ops->qw[b];
      double val;
      if (which == 1) {
        val = wq * Un[IH];
      } else {
        const double h = Un[IH];
        const double u = Un[IHU] / h, v = Un[IHV] / h, w = Un[IHW] / h;
        const double ed = 0.5 * h *
                          (u * u + v * v + (1.0 / ph.gamma) * w * w + ph.g * (h - 2.0 * g.Hnod[i]));
        val = wq * ed;
      }
      acc_add(s, c, val);
    }
  }
  block_reduce(s, c, m, ws);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = s;
    part[3 * blockIdx.x + 1] = c;
    part[3 * blockIdx.x + 2] = m;
  }
}

inline unsigned elementwise_blocks(size_t n) {
  const size_t want = (n + kThreads - 1) / kThreads;
  return (unsigned)(want < 16384 ? (want ? want : 1) : 16384);
}

constexpr int kTileW = kStripW;   // elements per RHS tile / sweep strip

}  // namespace

long long sweep_tiles(const Geometry& g) {
  const long long strips = (g.Nx + kTileW - 1) / kTileW;
  const long long chunks = (g.Ny + g.rows - 1) / g.rows;
  return strips * chunks;
}

namespace {

// one row-sweep launch of a (mode, waves, keep) build, with the still-depth
// gradients read (hxr false) or formed from Hnod (hxr true)
template <int M, int W, bool K>
void sweep_launch(bool hxr, dim3 grid, dim3 block, hipStream_t st, const double* U, double* dU,
                  const Geometry& g, const Phys& ph, GhostTime gt, const Stage& sa, int strips) {
  if (hxr)
    hipLaunchKernelGGL((k_rhs_sweep<kTileW, M, W, K, true>), grid, block, 0, st, U, dU, g, ph, gt,
                       sa, strips);
  else
    hipLaunchKernelGGL((k_rhs_sweep<kTileW, M, W, K, false>), grid, block, 0, st, U, dU, g, ph, gt,
                       sa, strips);
}

template <int M, int W, bool K>
void sweep_launch_batch(bool hxr, dim3 grid, dim3 block, hipStream_t st, const BatchMember* mem,
                        const long long* tileStart, int nMem) {
  if (hxr)
    hipLaunchKernelGGL((k_rhs_sweep_batch<kTileW, M, W, K, true>), grid, block, 0, st, mem,
                       tileStart, nMem);
  else
    hipLaunchKernelGGL((k_rhs_sweep_batch<kTileW, M, W, K, false>), grid, block, 0, st, mem,
                       tileStart, nMem);
}

}  // namespace

hipError_t launch_rhs_stage(const Geometry& g, const Phys& ph, int mode, const Stage& sa,
                            double* dU, GhostTime gt, hipStream_t st) {
  const int strips = (g.Nx + kTileW - 1) / kTileW;
  const long long tiles = sweep_tiles(g);
  if (tiles > 0x7fffffffLL || g.rows < 1) return hipErrorInvalidValue;
  const dim3 grid((unsigned)tiles), block(Sweep<kTileW>::NT);
  // waves per SIMD the register allocation targets, per stage mode (bit m of
  // g.waves3 set: 3 waves, else 2)
  const bool w3 = (g.waves3 >> mode) & 1;
  if (mode < kPlain || mode > kErr || ph.nManning != 0.0) return hipErrorInvalidValue;
  // Hx, Hy formed from Hnod (when the device check allowed it) in the
  // two-wave stage builds; the plain build and the three-wave stage-2 build
  // have no registers for it and read them
  const bool hxr = g.hxr != 0 && mode != kPlain && !(mode == kAxpy && w3);
#define SGNX_SWEEP(M, W, K) sweep_launch<M, W, K>(hxr, grid, block, st, sa.u, dU, g, ph, gt, sa, strips)
  const bool keep = sa.yOut != nullptr;
  if (mode == kPlain) { if (w3) SGNX_SWEEP(kPlain, 3, false); else SGNX_SWEEP(kPlain, 2, false); }
  else if (mode == kAxpy) {
    if (keep) { if (w3) SGNX_SWEEP(kAxpy, 3, true); else SGNX_SWEEP(kAxpy, 2, true); }
    else { if (w3) SGNX_SWEEP(kAxpy, 3, false); else SGNX_SWEEP(kAxpy, 2, false); }
  } else if (mode == kCand) {
    if (keep) { if (w3) SGNX_SWEEP(kCand, 3, true); else SGNX_SWEEP(kCand, 2, true); }
    else { if (w3) SGNX_SWEEP(kCand, 3, false); else SGNX_SWEEP(kCand, 2, false); }
  } else { if (w3) SGNX_SWEEP(kErr, 3, false); else SGNX_SWEEP(kErr, 2, false); }
#undef SGNX_SWEEP
  return hipGetLastError();
}

int sweep_strips(const Geometry& g) { return (g.Nx + kTileW - 1) / kTileW; }

hipError_t launch_rhs_stage_batch(const BatchMember* mem, const long long* tileStart, int nMem,
                                  long long totalTiles, int mode, bool waves3, hipStream_t st,
                                  bool keep, bool hxr) {
  if (nMem < 1 || totalTiles < 1 || totalTiles > 0x7fffffffLL) return hipErrorInvalidValue;
  if (mode < kPlain || mode > kErr) return hipErrorInvalidValue;
  const dim3 grid((unsigned)totalTiles), block(Sweep<kTileW>::NT);
  hxr = hxr && mode != kPlain && !(mode == kAxpy && waves3);
#define SGNX_SWEEPB(M, W, K) sweep_launch_batch<M, W, K>(hxr, grid, block, st, mem, tileStart, nMem)
  if (mode == kPlain) { if (waves3) SGNX_SWEEPB(kPlain, 3, false); else SGNX_SWEEPB(kPlain, 2, false); }
  else if (mode == kAxpy && keep) { if (waves3) SGNX_SWEEPB(kAxpy, 3, true); else SGNX_SWEEPB(kAxpy, 2, true); }
  else if (mode == kAxpy) { if (waves3) SGNX_SWEEPB(kAxpy, 3, false); else SGNX_SWEEPB(kAxpy, 2, false); }
  else if (mode == kCand && keep) { if (waves3) SGNX_SWEEPB(kCand, 3, true); else SGNX_SWEEPB(kCand, 2, true); }
  else if (mode == kCand) { if (waves3) SGNX_SWEEPB(kCand, 3, false); else SGNX_SWEEPB(kCand, 2, false); }
  else { if (waves3) SGNX_SWEEPB(kErr, 3, false); else SGNX_SWEEPB(kErr, 2, false); }
#undef SGNX_SWEEPB
  return hipGetLastError();
}

hipError_t launch_finish_batch(const BatchMember* mem, int nMem, double* out, hipStream_t st) {
  if (nMem < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_finish_batch, dim3((unsigned)nMem), dim3(kThreads), 0, st, mem, nMem, out);
  return hipGetLastError();
}

hipError_t launch_finish_post(const double* part, long long nb, double* host,
                              unsigned long long seq, hipStream_t st) {
  hipLaunchKernelGGL(k_finish_post, dim3(1), dim3(kThreads), 0, st, part, (int)nb, host, seq);
  return hipGetLastError();
}

hipError_t launch_finish(const double* part, long long nb, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(kThreads), 0, st, part, (int)nb, out);
  return hipGetLastError();
}

hipError_t launch_rhs(const Geometry& g, const Phys& ph, const double* U, double* dU,
                      GhostTime gt, hipStream_t st) {
  if (g.variant == 1) {
    Stage sa{};
    sa.u = U;
    return launch_rhs_stage(g, ph, kPlain, sa, dU, gt, st);
  }
  const int tpr = (g.Nx + kTileW - 1) / kTileW;
  const long long tiles = (long long)tpr * g.Ny;
  if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_rhs<kTileW>, dim3((unsigned)tiles), dim3(Tile<kTileW>::NT), 0, st, U, dU,
                     g, ph, gt, tpr);
  return hipGetLastError();
}

hipError_t launch_bathy_check(const Geometry& g, const double* Hnod, const double* Hx,
                              const double* Hy, int* bad, hipStream_t st) {
  const long long nb = (g.Nd + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(k_bathy_check, dim3((unsigned)nb), dim3(kThreads), 0, st, Hnod, Hx, Hy,
                     g.ops, g.Nd, g.Jx, g.Jy, bad);
  return hipGetLastError();
}

hipError_t launch_bathy_grad(const Geometry& g, const double* Hnod, double* Hx, double* Hy,
                             hipStream_t st) {
  const long long nb = (g.Nd + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(k_bathy_grad, dim3((unsigned)nb), dim3(kThreads), 0, st, Hnod, Hx, Hy,
                     g.ops, g.Nd, g.Jx, g.Jy);
  return hipGetLastError();
}

hipError_t launch_comb1(double* out, const double* u, double a1, const double* k1, size_t n,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_comb1, dim3(elementwise_blocks(n)), dim3(kThreads), 0, st, out, u, a1, k1,
                     n);
  return hipGetLastError();
}

hipError_t launch_bs3_update(double* out, const double* u, double dt, const double* k1,
                             const double* k2, const double* k3, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_bs3_update, dim3(elementwise_blocks(n)), dim3(kThreads), 0, st, out, u,
                     dt, k1, k2, k3, n);
  return hipGetLastError();
}

hipError_t launch_rk4_update(double* u, double dt, const double* k1, const double* k2,
                             const double* k3, const double* k4, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_rk4_update, dim3(elementwise_blocks(n)), dim3(kThreads), 0, st, u, dt, k1,
                     k2, k3, k4, n);
  return hipGetLastError();
}

hipError_t launch_err_norm(const double* u, const double* uNew, const double* k1,
                           const double* k2, const double* k3, const double* k4, double dt,
                           double atol, double rtol, size_t n, double* part, double* out,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_err_norm, dim3(kReduceBlocks), dim3(kThreads), 0, st, u, uNew, k1, k2, k3,
                     k4, dt, atol, rtol, n, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(kThreads), 0, st, part, kReduceBlocks, out);
  return hipGetLastError();
}

hipError_t launch_eta_max(const Geometry& g, const double* U, double hx, double hy,
                          const EtaWindow& w, double* part, hipStream_t st) {
  hipLaunchKernelGGL(k_eta_max, dim3(kReduceBlocks), dim3(kThreads), 0, st, U, g.Hnod, g, hx, hy,
                     w, part);
  return hipGetLastError();
}

hipError_t launch_gather_eta(const Geometry& g, const double* U, const long long* nodes, int n,
                             double* out, hipStream_t st) {
  if (n < 1) return hipSuccess;
  hipLaunchKernelGGL(k_gather_eta, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, st, U,
                     g.Hnod, nodes, n, out);
  return hipGetLastError();
}

hipError_t launch_elem_means(const Geometry& g, const double* U, double* out, hipStream_t st) {
  const long long Ne = (long long)g.Nx * g.Ny;
  hipLaunchKernelGGL(k_elem_means, dim3((unsigned)((Ne + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, st, U, g.Hnod, g.ops, g.Nd, Ne, out);
  return hipGetLastError();
}

hipError_t launch_diag(const Geometry& g, const Phys& ph, const double* U, int which,
                       double* part, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_diag, dim3(kReduceBlocks), dim3(kThreads), 0, st, U, g, ph, which, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(kThreads), 0, st, part, kReduceBlocks, out);
  return hipGetLastError();
}

}  // namespace sgnx
