// =============================================================================
// sgnx_fastdiv.h -- correctly rounded a / b for a divisor b known in advance.
//
// The FR lift divides by the element Jacobians Jx, Jy, which are fixed for a
// mesh. With rh = RN(1/b) and rl = RN(1/b - rh) computed once,
//   q1 = RN(a*rh + RN(a*rl))      a faithful rounding of a/b (rh + rl is 1/b
//                                 to about 2^-106 relative),
//   r  = a - b*q1                 exact (one fma; the remainder of a faithful
//                                 quotient is representable),
//   q  = RN(q1 + r*rh)            = RN(a/b) by Markstein's theorem (rh within
//                                 half an ulp of 1/b, q1 faithful),
// which is the IEEE quotient a/b bit for bit. The last step is evaluated as
// -fma(-r, rh, -q1) so that a = -0 yields -0 like the division does.
// Preconditions (checked): b > 0 with 2^-100 <= b <= 2^100 (else the caller
// keeps plain division), and |a| in [2^-900, 2^800] or a == 0. Then |a/b| >=
// 2^-1000 is normal, the remainder's last bit (about 2^(e_a - 105)) is above
// 2^-1074, and the rounding error of a*rl (at most 2^-1075 when it is
// subnormal) is far below half an ulp of q1, so every step above holds. Any
// other a (tiny, huge, infinite, NaN) takes a / b.
//
// Three fmas and one multiply (full-rate) against the ~10-instruction
// scaled-reciprocal division sequence with its quarter-rate v_rcp_f64.
// The header is compiled for host and device; tests check it against a / b.
// =============================================================================
#ifndef SGNX_FASTDIV_H
#define SGNX_FASTDIV_H

#include <cmath>

#if defined(__HIPCC__)
#define SGNX_HD __host__ __device__
#else
#define SGNX_HD
#endif

namespace sgnx {

struct ConstDiv {
  double b, rh, rl;
  int fast;        // 0: always use a / b
};

// Host: the reciprocal pair of b (computed with IEEE division and an exact
// fma residual: e = 1 - b*rh is exact for rh = RN(1/b), so rl = RN(e/b)).
inline ConstDiv make_const_div(double b) {
  ConstDiv d{b, 0.0, 0.0, 0};
  if (b > 0.0 && b >= 0x1p-100 && b <= 0x1p+100) {
    d.rh = 1.0 / b;
    const double e = std::fma(-b, d.rh, 1.0);
    d.rl = e / b;
    d.fast = 1;
  }
  return d;
}

// a is in the range the fast sequence is exact for
SGNX_HD inline bool div_fast_ok(double a) {
  const double aa = fabs(a);
  return aa <= 0x1p+800 && (aa >= 0x1p-900 || aa == 0.0);
}

SGNX_HD inline double div_fast(double a, const ConstDiv& d) {
  const double t = a * d.rl;
  const double q1 = fma(a, d.rh, t);
  const double r = fma(-q1, d.b, a);
  return -fma(-r, d.rh, -q1);
}

SGNX_HD inline double div_const(double a, const ConstDiv& d) {
  if (!d.fast || !div_fast_ok(a)) return a / d.b;
  return div_fast(a, d);
}

}  // namespace sgnx

#endif  // SGNX_FASTDIV_H
