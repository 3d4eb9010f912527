// fft_roots.hpp -- compile-time tables of the roots of unity used by the
// register-resident FFT stages: kTwL<L>.c[t] + i kTwL<L>.s[t] = exp(-2 pi i t / L).
// The values come from Taylor series in double precision (16 terms after the
// angle is reduced to [-pi, pi], error far below float rounding), so every
// index that is a constant after unrolling becomes an immediate operand.
#pragma once

namespace cgs {
namespace roots {

constexpr double kPiD = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x) {
  double t = x, s = x;
  for (int i = 1; i < 16; ++i) {
    t *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
    s += t;
  }
  return s;
}
constexpr double taylor_cos(double x) {
  double t = 1.0, s = 1.0;
  for (int i = 1; i < 16; ++i) {
    t *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
    s += t;
  }
  return s;
}

// exp(-2 pi i t / L) for t in [0, L): angle reduced to [-pi, pi]
template <int L>
struct TwL {
  float c[L], s[L];
  constexpr TwL() : c(), s() {
    for (int t = 0; t < L; ++t) {
      double a = -2.0 * kPiD * t / L;
      if (a < -kPiD) a += 2.0 * kPiD;
      c[t] = (float)taylor_cos(a);
      s[t] = (float)taylor_sin(a);
    }
  }
};

}  // namespace roots

template <int L>
__device__ constexpr roots::TwL<L> kTwL = roots::TwL<L>();

}  // namespace cgs
