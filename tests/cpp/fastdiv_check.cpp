// Checks sgnx::div_const against IEEE division on random and edge inputs.
// usage: fastdiv_check <n_random> <b1> [b2 ...]; prints mismatches, exit 1 on any.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "../../sgn-hyperbolic-gpu_amd/csrc/sgnx_fastdiv.h"

static uint64_t bits(double x) { uint64_t u; std::memcpy(&u, &x, 8); return u; }
static double from_bits(uint64_t u) { double x; std::memcpy(&x, &u, 8); return x; }

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 1000000;
  long bad = 0, tested = 0;
  std::mt19937_64 rng(12345);
  for (int k = 2; k < argc; ++k) {
    const double b = std::atof(argv[k]);
    const sgnx::ConstDiv d = sgnx::make_const_div(b);
    auto check = [&](double a) {
      ++tested;
      const double want = a / b, got = sgnx::div_const(a, d);
      if (bits(want) != bits(got) && !(std::isnan(want) && std::isnan(got))) {
        if (bad < 10) std::printf("b=%a a=%a want=%a got=%a\n", b, a, want, got);
        ++bad;
      }
    };
    // random doubles across the whole exponent range, random signs
    for (long i = 0; i < n; ++i) check(from_bits(rng()));
    // values of the magnitudes the lift produces
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (long i = 0; i < n; ++i) check(U(rng) * std::ldexp(1.0, (int)(rng() % 80) - 40));
    // neighbours of exact multiples of b (quotients near integers and halfway cases)
    for (long i = 0; i < n / 4; ++i) {
      const double m = (double)(int64_t)(rng() % 2000001) - 1000000.0;
      double a = m * b;
      for (int s = -2; s <= 2; ++s) check(std::nextafter(a, s < 0 ? -INFINITY : INFINITY) + 0.0 * s);
      check(a);
      check(std::nextafter(a, INFINITY));
      check(std::nextafter(a, -INFINITY));
      check((m + 0.5) * b);
    }
    // tiny dividends (the scaled path and its subnormal-quotient cut-off):
    // random significands with exponents from the subnormal range to 2^-880
    for (long i = 0; i < n; ++i) {
      const uint64_t mant = rng() & 0x000fffffffffffffULL, sign = rng() & 0x8000000000000000ULL;
      const uint64_t ex = rng() % 145;          // biased exponents 0 (subnormal) .. 144 (2^-879)
      check(from_bits(sign | (ex << 52) | mant));
    }
    // around the quotient's normal / subnormal boundary: a near 2^-1022 b
    for (long i = 0; i < n / 4; ++i) {
      double a = std::ldexp(b, -1022) * (1.0 + std::ldexp(U(rng), -(int)(rng() % 60)));
      check(a);
      check(std::nextafter(a, INFINITY));
      check(-std::nextafter(a, 0.0));
    }
    for (double a : {0.0, -0.0, (double)INFINITY, -(double)INFINITY, (double)NAN, 0x1p-1074, -0x1p-1074, 0x1p-1022,
                     0x1p-800, 0x1.fffffffffffffp-801, 0x1p-900, 0x1.fffffffffffffp-901,
                     -0x1.0000000000001p-900, 0x1p+800, 0x1.0000000000001p+800,
                     0x1.fffffffffffffp+1023, 1.0, -1.0, b, -b, std::ldexp(b, -1022),
                     std::ldexp(b, -1021), std::ldexp(b, -1023)})
      check(a);
  }
  std::printf("fastdiv: %ld tested, %ld mismatches\n", tested, bad);
  return bad ? 1 : 0;
}
