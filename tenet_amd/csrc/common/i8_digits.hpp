// INT8 digits of the split-INT8 GEMM operands, one element at a time.
//
// An element v of a line with exponent e is a = v 2^(7 - e), |a| < 126.5, and
// is stored as S INT8 digits with
//     a ~= d_0 + sum_{p=1..S-1} d_p 256^-p,  d_p in [-128, 127] (p >= 1),
// the value rounded to a multiple of 256^-(S-1) (round to nearest, ties to
// even).  With the lower digits in [-128, 127] that representation is unique,
// so any way of computing it gives the same digits.  Two ways are here:
//
//  * i8_digits_stepwise: d_0 = rint(a), then S - 1 times "remainder times
//    256, round"; a digit that reaches 128 becomes -128 and carries one into
//    the digit above, from the last digit up.  Every digit before the last
//    contributes a multiple of 256 units of the last one, so the last
//    rounding -- ties included -- is the one-step rounding below, and the two
//    give the same digits for every input.
//  * i8_digits_fixed (S <= 5): d_0 = rint(a); the remainder r = a - d_0
//    (|r| <= 1/2, exact) is rounded once to R = rint(r 2^F), F = 8 (S - 1)
//    <= 32 fractional bits; U = R + 0x80...80 (128 added to every byte) lies
//    in [0, 2^(F+1)); U >= 2^F means the top fractional digit would be 128,
//    so 2^F moves into d_0; then byte j of U minus 128 -- byte j XOR 0x80 as
//    an INT8 -- is digit S - 1 - j.  About ten operations per element
//    instead of ~50.
//
// Both run on the host too (the CPU tests compare them on random, tie and
// carry cases).
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define TNT_HD __host__ __device__
#else
#define TNT_HD
#endif

namespace tnt_digits {

// digits[0..S-1], d_0 first.
template <int S>
TNT_HD inline void i8_digits_stepwise(double a, int (&d)[S]) {
  double r = ::rint(a);
  d[0] = (int)r;
  a -= r;
  for (int p = 1; p < S; ++p) {
    a *= 256.0;
    r = ::rint(a);
    d[p] = (int)r;
    a -= r;
  }
  int carry = 0;
  for (int p = S - 1; p >= 1; --p) {
    d[p] += carry;
    carry = d[p] > 127 ? 1 : 0;
    d[p] -= carry << 8;
  }
  d[0] += carry;
}

// d_0 (as an int) and the fractional digits packed as bytes of *w: byte j
// holds digit S - 1 - j as an INT8 (two's complement).  S in [2, 5].
template <int S>
TNT_HD inline int i8_digits_fixed(double a, uint32_t* w) {
  static_assert(S >= 2 && S <= 5, "fractional digits must fit 32 bits");
  constexpr int F = 8 * (S - 1);
  constexpr double kScale = (double)(uint64_t(1) << F);
  constexpr uint32_t kBias = (uint32_t)(0x80808080ull >> (32 - F));
  double d0 = ::rint(a);
  const double r = a - d0;                      // exact, |r| <= 1/2
  double u = ::rint(r * kScale) + (double)kBias;  // exact integer in [0, 2^(F+1))
  if (u >= kScale) {                            // top fractional digit 128: carry
    u -= kScale;
    d0 += 1.0;
  }
  *w = (uint32_t)u ^ kBias;
  return (int)d0;
}

}  // namespace tnt_digits
