// CPU check of the two INT8 digit formulations (tenet_amd/csrc/common/i8_digits.hpp):
// on random elements, exact rounding ties of the last digit, half-integer
// leading digits and carry cases, the fixed-point digits are a valid
// representation (digits in range, value within half a unit of the last
// digit, exactly half a unit at a tie) and equal the stepwise digits.  Prints
// one line per digit count and exits non-zero on a failure.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../tenet_amd/csrc/common/i8_digits.hpp"

template <int S>
static int check(double a, long* mism, long* ties) {
  constexpr int F = 8 * (S - 1);
  int ds[S];
  tnt_digits::i8_digits_stepwise<S>(a, ds);
  uint32_t w = 0;
  const int d0 = tnt_digits::i8_digits_fixed<S>(a, &w);
  int df[S];
  df[0] = d0;
  for (int p = 1; p < S; ++p) df[p] = (int)(int8_t)(uint8_t)(w >> (8 * (S - 1 - p)));
  if constexpr (F < 32)
    if ((w >> F) != 0) { std::printf("FAIL high bytes a=%.17g\n", a); return 1; }
  // ranges
  if (df[0] < -127 || df[0] > 127) { std::printf("FAIL d0 a=%.17g d0=%d\n", a, df[0]); return 1; }
  // value within half a unit of the last digit (exact in long double on
  // x86: 64-bit significand covers 7 + 32 + 1 bits of these values)
  long double v = df[0];
  for (int p = 1; p < S; ++p) v += (long double)df[p] * std::ldexp(1.0L, -8 * p);
  const long double err = std::fabs(v - (long double)a), half = std::ldexp(1.0L, -F - 1);
  if (err > half) { std::printf("FAIL value a=%.17g err=%Lg\n", a, err); return 1; }
  bool same = true;
  for (int p = 0; p < S; ++p) same = same && ds[p] == df[p];
  if (!same) {
    ++*mism;
    std::printf("FAIL digits differ from the stepwise digits a=%.17g\n", a);
    return 1;
  }
  if (err == half) ++*ties;
  return 0;
}

template <int S>
static int run(long n) {
  constexpr int F = 8 * (S - 1);
  std::mt19937_64 g(12345 + S);
  std::uniform_real_distribution<double> U(-126.49, 126.49);
  long mism = 0, ties = 0, cases = 0;
  auto one = [&](double a) { ++cases; return check<S>(a, &mism, &ties); };
  for (long i = 0; i < n; ++i)
    if (one(U(g))) return 1;
  // exact ties of the last digit, half-integer d_0 ties, carry boundaries
  std::uniform_int_distribution<int64_t> I(-126, 126);
  std::uniform_int_distribution<int64_t> FR(0, (int64_t(1) << F) - 1);
  for (long i = 0; i < n / 4; ++i) {
    const double k = (double)I(g), f = std::ldexp((double)FR(g), -F);
    if (one(k + f + std::ldexp(1.0, -F - 1)) || one(-(k + f + std::ldexp(1.0, -F - 1)))) return 1;
    if (one(k + 0.5) || one(k - 0.5)) return 1;
    if (one(k + 0.5 - std::ldexp(1.0, -F - 1)) || one(k + 0.5 - std::ldexp(1.0, -F - 2))) return 1;
    if (one(k - 0.5 + std::ldexp(1.0, -F - 2)) || one(k + f)) return 1;
  }
  for (double a : {0.0, -0.0, 126.4999999, -126.4999999, 0.5, -0.5, 1.5, -1.5, 126.0, -126.0})
    if (one(a)) return 1;
  std::printf("S=%d cases %ld mismatches %ld ties %ld\n", S, cases, mism, ties);
  return 0;
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 1000000;
  return run<2>(n) || run<3>(n) || run<4>(n) || run<5>(n);
}
