// Certified sine for the q units of the table builder.
//
// The reference converts 2theta to q with the platform libm,
// q = (4 pi / (lambda * 1e10 or 1e9)) * sin(2theta / 2)
// (RadialConverter::convert, integration.rs:171-180). glibc's sin is not
// correctly rounded, yet bin indices and weights in q must match the
// reference bit for bit. As for atan2 (atan2_cert.hpp), the argument is
// evaluated in double-double and the double result is accepted only when
// the exact value is provably far from a rounding boundary; everything else
// is recomputed by the host with glibc.
//
// Method, for 0 <= x <= 2 (x = 2theta / 2 lies in [0, pi/2]):
//   x = c + u with c = k/64 the nearest table point and |u| <= 1/128
//   (u = x - c is exact); sin(x) = sin(c) cos(u) + cos(c) sin(u) with
//   sin(c), cos(c) tabulated in double-double (sin_table.inc, from exact
//   rational series) and sin(u) = u (1 - u^2/3! + ... - u^10/11!),
//   cos(u) = 1 - u^2/2! + ... + u^12/12! in double-double (truncation below
//   2^-110 relative). The evaluation error is far below 2^-80 relative.
//   hi = RN(value) is flagged uncertain when the double-double tail lies
//   within kSinMargin ulp of half an ulp (glibc's sin stays within 0.52 ulp
//   of the exact value, so it rounds the other way only inside 0.02 ulp of
//   a midpoint), when hi is a power of two, and for arguments outside
//   (2^-20, 2]; sin(+0) = +0 exactly.
#pragma once
#include "dd.hpp"

namespace dio {

#include "sin_table.inc"

constexpr double kSinMargin = 0.0625;  // in ulps

struct SinResult {
  double value;
  bool uncertain;
};

__host__ __device__ inline double ulp_above(double v) {
  int e;
  std::frexp(v, &e);  // v = m * 2^e, 0.5 <= |m| < 1
  return std::ldexp(1.0, e - 53);
}

__host__ __device__ inline SinResult sin_cert(double x) {
  if (x == 0.0 && !std::signbit(x)) return {0.0, false};
  if (!(x > 0x1p-20 && x <= 2.0)) return {0.0, true};
  const int k = static_cast<int>(x * 64.0 + 0.5);
  const double c = static_cast<double>(k) * 0.015625;
  const double u = x - c;  // exact: |u| <= 1/128 and c is within a factor 2 of x (k >= 1)
  const dd s = two_prod(u, u);
  auto coef = [](int n, bool neg) {  // (-1)^neg / n!
    const dd v = {kInvFact[n - 2][0], kInvFact[n - 2][1]};
    return neg ? dd_neg(v) : v;
  };
  // sin(u) / u = 1 + s (-1/3! + s (1/5! + s (-1/7! + s (1/9! + s (-1/11!)))))
  dd ps = coef(11, true);
  ps = dd_add(coef(9, false), dd_mul(s, ps));
  ps = dd_add(coef(7, true), dd_mul(s, ps));
  ps = dd_add(coef(5, false), dd_mul(s, ps));
  ps = dd_add(coef(3, true), dd_mul(s, ps));
  ps = dd_add(dd{1.0, 0.0}, dd_mul(s, ps));
  const dd sin_u = dd_mul(dd{u, 0.0}, ps);
  // cos(u) = 1 + s (-1/2! + s (1/4! + s (-1/6! + s (1/8! + s (-1/10! + s / 12!)))))
  dd pc = coef(12, false);
  pc = dd_add(coef(10, true), dd_mul(s, pc));
  pc = dd_add(coef(8, false), dd_mul(s, pc));
  pc = dd_add(coef(6, true), dd_mul(s, pc));
  pc = dd_add(coef(4, false), dd_mul(s, pc));
  pc = dd_add(coef(2, true), dd_mul(s, pc));
  const dd cos_u = dd_add(dd{1.0, 0.0}, dd_mul(s, pc));
  const dd sc = {kSinTable[k][0], kSinTable[k][1]}, cc = {kCosTable[k][0], kCosTable[k][1]};
  dd v = dd_add(dd_mul(sc, cos_u), dd_mul(cc, sin_u));
  v = fast_two_sum(v.hi, v.lo);
  const double hi = v.hi;
  int e;
  if (std::fabs(std::frexp(hi, &e)) == 0.5) return {hi, true};  // power of two
  const double frac = std::fabs(v.lo) / ulp_above(hi);           // in [0, 0.5]
  return {hi, std::fabs(frac - 0.5) < kSinMargin};
}

}  // namespace dio
