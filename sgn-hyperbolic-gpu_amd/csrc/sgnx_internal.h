// =============================================================================
// sgnx_internal.h -- device-side data of a solver handle and the kernel
// launchers shared by the C-ABI layer (sgnx_api.hip) and the kernels
// (sgnx_kernels.hip).
// =============================================================================
#ifndef SGNX_INTERNAL_H
#define SGNX_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdlib>

#include "sgnx_fastdiv.h"
#include "sgnx_phys.h"

namespace sgnx {

constexpr int NP = 5;     // nodes per direction at P = 4
constexpr int NPP = 25;   // nodes per element

// Per-lane operator tables (uploaded once): D row-major, gL', gR'.
struct OpTables {
  double D[NPP];
  double gLp[NP], gRp[NP];
  double qw[NP];
};

// Element columns per RHS tile / row-sweep strip. Five columns make a
// 128-work-item workgroup (125 node items, two waves), so a compute unit holds
// four workgroups of the two-wave builds and six of the three-wave one.
constexpr int kStripW = 5;

// The waves mask of SGNX_SWEEP_WAVES3 (host): 0x-prefixed hexadecimal or
// decimal text ("0x2" and "2" alike; a leading 0 alone is not octal), low 4 bits.
inline int waves_mask_from_text(const char* w) {
  const bool hex = w[0] == '0' && (w[1] == 'x' || w[1] == 'X');
  return (int)std::strtol(w, nullptr, hex ? 16 : 10) & 0xf;
}

// Bijective remap of workgroup ids so that XCD x (ids x, x+8, x+16, ...)
// receives the contiguous tile range [start_x, start_x + count_x).
__device__ __forceinline__ int xcd_tile(int bid, int nb) {
  constexpr int kXcd = 8;
  const int q = nb / kXcd, r = nb - q * kXcd;
  const int x = bid - (bid / kXcd) * kXcd, i = bid / kXcd;
  return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// What one RHS launch needs besides the stage input and output.
struct Geometry {
  int Nx, Ny;
  long long Nd;              // nodes per variable
  double Jx, Jy;
  int bc[4];                 // left, right, bottom, top
  const OpTables* ops;       // device
  const double* Hnod;        // device, Nd
  const double* Hx;          // device, Nd
  const double* Hy;          // device, Nd
  const double* HfX;         // device, (Nx+1)*Ny*5
  const double* HfY;         // device, Nx*(Ny+1)*5
  const double* inflow;      // device, Ny*5*15 or null
  ConstDiv dx, dy;           // division by Jx, Jy
  int variant;               // RHS kernel: 0 element-row tiles, 1 row sweep
  int rows;                  // row sweep: element rows per workgroup
  int waves3;                // row sweep: bit m set = stage mode m built for 3 waves/SIMD
  int hxr;                   // Hx, Hy equal (D Hnod) / J bit for bit: the row sweep forms
                             // them from Hnod instead of reading them
};

// Inflow ghost time factors for one RHS evaluation: sin(w t), cos(w t),
// evaluated on the host with the C library, as the oracle does.
struct GhostTime { double s, c; };

// Stage input formed while loading (row-sweep kernel), with the arithmetic of
// src/solver/rk.cpp:99-131 (BS3 stages, candidate and embedded error):
//   kPlain   U                                   dU = R(U)
//   kAxpy    u + a*k1                            dU = R(u + a k1)   (stage 2)
//   kCand    u + a*k1, a = 0.75 dt, k1 = k2      stage 3; at each node the
//            launch forms k3 = R(.) and writes, instead of k3,
//              uNew = u + dt*((2/9 kc + 1/3 k1) + 4/9 k3)   (kc = the k1 of rk.cpp)
//              z    = (7/24 kc + 0.25 k1) + 1/3 k3
//            (the candidate and the first three terms of the embedded sum)
//   kErr     uNew                                dU = k4 = R(uNew), and the
//            scaled embedded error with u0 = u and z:
//              r = (uNew - (u0 + dt*(z + 0.125 k4))) / (atol + rtol*max|.|),
//            one (sum r^2, carry, max |r|) triple per workgroup in `part`
// Together kCand and kErr read and write what the separate candidate update
// would, but every face and halo state of the last RHS is a single operand.
enum StageMode { kPlain = 0, kAxpy = 1, kCand = 2, kErr = 3 };
struct Stage {
  const double* u;    // stage input (kErr: uNew)
  const double* k1;   // kAxpy, kCand: the increment operand
  double a;           // kAxpy, kCand: its coefficient
  const double* kc;   // kCand: k1 of the step
  double dt;          // kCand, kErr
  double* uNew;       // kCand: candidate out
  double* zOut;       // kCand: partial embedded sum out
  const double* u0;   // kErr: u of the step
  const double* z;    // kErr: partial embedded sum
  double atol, rtol;  // kErr
  double* part;       // kErr
  double* yOut;       // kAxpy, kCand (optional): the stage input u + a*k1 out
  double* y3Out;      // kAxpy with yOut (optional): the next stage's input u + a3*R out
  double a3;          // its coefficient
};

// One grid of a batched launch: its operands, geometry, constants, inflow
// time factors and stage, and its row-sweep decomposition (strips per chunk
// row, tiles in all).
struct BatchMember {
  const double* U;
  double* dU;
  Geometry g;
  Phys ph;
  GhostTime gt;
  Stage sa;
  int nStrips;
  int nTiles;
};

// ---- launchers (sgnx_kernels.hip); all asynchronous on `st` ----------------

// dU = R(U)
hipError_t launch_rhs(const Geometry& g, const Phys& ph, const double* U, double* dU,
                      GhostTime gt, hipStream_t st);
// dU = R(stage input) with the row-sweep kernel (kCand writes sa.uNew and
// sa.zOut instead of dU; kErr also writes sa.part, 3 doubles per workgroup,
// folded by launch_finish). kAxpy and kCand with sa.yOut set also write the
// stage input they form (the batched form: keep = some member's yOut is set).
hipError_t launch_rhs_stage(const Geometry& g, const Phys& ph, int mode, const Stage& sa,
                            double* dU, GhostTime gt, hipStream_t st);
// out[0] = sum, out[1] = max of nb (sum, carry, max) partial triples, fixed order
hipError_t launch_finish(const double* part, long long nb, double* out, hipStream_t st);
// the same into host memory mapped for the device (device address `host`):
// host[0] = sum, host[1] = max, then host[2] = seq (as 64-bit integer bits)
hipError_t launch_finish_post(const double* part, long long nb, double* host,
                              unsigned long long seq, hipStream_t st);
// The row sweep over every member of a batch in one launch (device arrays:
// mem[nMem]; tileStart[m] = sum of the tiles of members before m, totalTiles
// their sum); `mode` as launch_rhs_stage, for all members.
hipError_t launch_rhs_stage_batch(const BatchMember* mem, const long long* tileStart, int nMem,
                                  long long totalTiles, int mode, bool waves3, hipStream_t st,
                                  bool keep = false, bool hxr = false);
// launch_finish for every member: member m's partials (mem[m].sa.part, its
// nTiles) into out[2m], out[2m + 1]
hipError_t launch_finish_batch(const BatchMember* mem, int nMem, double* out, hipStream_t st);
// row-sweep strips per chunk row of a grid
int sweep_strips(const Geometry& g);
// workgroups of one row-sweep launch (sizes the partial buffer)
long long sweep_tiles(const Geometry& g);
// Hx, Hy = (D . Hnod) / J per element line (sgn2d.cpp:54-71)
hipError_t launch_bathy_grad(const Geometry& g, const double* Hnod, double* Hx, double* Hy,
                             hipStream_t st);
// *bad |= 1 where Hx, Hy differ from (D . Hnod) / J in any bit (bad: device int,
// cleared by the caller)
hipError_t launch_bathy_check(const Geometry& g, const double* Hnod, const double* Hx,
                              const double* Hy, int* bad, hipStream_t st);
// out = u + a1*k1
hipError_t launch_comb1(double* out, const double* u, double a1, const double* k1, size_t n,
                        hipStream_t st);
// out = u + dt*((2/9 k1 + 1/3 k2) + 4/9 k3)   (rk.cpp:106-108)
hipError_t launch_bs3_update(double* out, const double* u, double dt, const double* k1,
                             const double* k2, const double* k3, size_t n, hipStream_t st);
// u += (dt/6)*(((k1 + 2k2) + 2k3) + k4)        (rk.cpp:36-37)
hipError_t launch_rk4_update(double* u, double dt, const double* k1, const double* k2,
                             const double* k3, const double* k4, size_t n, hipStream_t st);
// Scaled embedded error (rk.cpp:118-131): out[0] = sum r^2, out[1] = max |r|,
// reduced in a fixed order with compensated sums; `part` holds
// 3*kReduceBlocks doubles of scratch.
constexpr int kReduceBlocks = 1024;
hipError_t launch_err_norm(const double* u, const double* uNew, const double* k1,
                           const double* k2, const double* k3, const double* k4, double dt,
                           double atol, double rtol, size_t n, double* part, double* out,
                           hipStream_t st);
// Diagnostics of one state: which = 0 max wave speed (bitwise), 1 mass, 2 energy.
hipError_t launch_diag(const Geometry& g, const Phys& ph, const double* U, int which,
                       double* part, double* out, hipStream_t st);
// eta = h - H diagnostics (apps/shoal_focus_mpi.cu:196-231,
// apps/tsunami_margin_mpi.cu:255-278, src/io/vtk2d.h:39-81)
struct EtaWindow {
  double x0, x1, y0, y1;       // open window
  double xo[NP], yo[NP];       // 0.5 h (xi_a + 1): the node offsets in an element
};
// kReduceBlocks (max, node) partial pairs into part
hipError_t launch_eta_max(const Geometry& g, const double* U, double hx, double hy,
                          const EtaWindow& w, double* part, hipStream_t st);
hipError_t launch_gather_eta(const Geometry& g, const double* U, const long long* nodes, int n,
                             double* out, hipStream_t st);
// out[f * Nx * Ny + e], f = eta, u, v, p, H
hipError_t launch_elem_means(const Geometry& g, const double* U, double* out, hipStream_t st);

}  // namespace sgnx

#endif  // SGNX_INTERNAL_H
