// Certified atan2 for the table builder.
//
// The reference computes every angle with the platform libm (Rust f64::atan2,
// geometry.rs:684-692), which on glibc 2.35 is not correctly rounded: about
// 4 in 10^4 random arguments differ from the correctly rounded value, always
// where the exact angle lies within ~0.023 ulp of a rounding midpoint.
// Bin indices and weights must nevertheless match the reference bit for bit.
//
// Method (Ziv's rounding test on a double-double evaluation):
//   1. Evaluate atan2(y, x) in double-double with relative error < 2^-98:
//      t = min(|x|,|y|)/max(|x|,|y|), reduced by the nearest c = i/64 through
//      atan(t) = atan(c) + atan((t - c)/(1 + t*c)), |u| <= 1/128, then the
//      odd Taylor series of atan(u) (terms to u^17), tabulated atan(c) and
//      pi, pi/2 in double-double, and the quadrant identities.
//   2. hi = RN(value) is the correctly rounded result unless the exact value
//      may lie near the midpoint between hi and a neighbour. The result is
//      flagged "uncertain" when |lo| is within kMidpointMargin ulp of half an
//      ulp (a margin of 1/16 ulp, ~2.7x the largest libm deviation measured),
//      when hi is a power of two (asymmetric ulp), or for non-finite or
//      extreme arguments.
//   3. Unflagged results equal libm; flagged ones are recomputed with the
//      host libm by the caller, so the final angles equal the reference's.
#pragma once
#include "dd.hpp"

namespace dio {

#include "atan_table.inc"

constexpr double kMidpointMargin = 0.0625;  // in ulps

struct Atan2Result {
  double value;
  bool uncertain;
};

// ulp of |v| for a normal finite v (spacing above |v|).
__host__ __device__ inline double ulp_of(double v) {
  int e;
  std::frexp(v, &e);  // v = m * 2^e, 0.5 <= |m| < 1
  return std::ldexp(1.0, e - 53);
}

__host__ __device__ inline bool is_power_of_two(double v) {
  int e;
  double m = std::frexp(v, &e);
  return std::fabs(m) == 0.5;
}

__host__ __device__ inline Atan2Result atan2_cert(double y, double x) {
  // Exact special cases, identical to IEEE/libm atan2.
  if (!(std::isfinite(y) && std::isfinite(x))) {
    return {0.0, true};
  }
  if (y == 0.0) {
    // atan2(+-0, x>0 or +0) = +-0; atan2(+-0, x<0 or -0) = +-pi.
    bool neg_x = std::signbit(x);
    double r = neg_x ? kPi[0] : 0.0;
    return {std::copysign(r, y), false};
  }
  if (x == 0.0) {
    return {std::copysign(kHalfPi[0], y), false};
  }
  double ay = std::fabs(y), ax = std::fabs(x);
  bool swap = ay > ax;
  double num = swap ? ax : ay;
  double den = swap ? ay : ax;
  // Guard ranges where the double-double steps could underflow or overflow.
  if (den > 0x1p+500 || den < 0x1p-500 || num < 0x1p-400 * den) {
    return {0.0, true};
  }
  // t = num/den as double-double: the remainder of the division is exact.
  double th = num / den;
  double rem = std::fma(-th, den, num);
  dd t = {th, rem / den};
  t = fast_two_sum(t.hi, t.lo);

  int i = static_cast<int>(t.hi * 64.0 + 0.5);
  if (i > 64) i = 64;
  double c = static_cast<double>(i) * 0.015625;
  // u = (t - c) / (1 + t*c); t.hi - c is exact (Sterbenz) for i >= 1.
  dd u_num = two_sum(t.hi - c, t.lo);
  dd tc = dd_mul(t, dd{c, 0.0});
  dd u_den = dd_add(dd{1.0, 0.0}, tc);
  dd u = dd_div(u_num, u_den);

  // atan(u) = u * P(s), s = u^2,
  // P = 1 + s(-1/3 + s(1/5 + s(-1/7 + s*T))), T = 1/9 - s/11 + ... (double).
  dd s = dd_mul(u, u);
  double sh = s.hi;
  double T = 1.0 / 9.0 - sh * (1.0 / 11.0 - sh * (1.0 / 13.0 - sh * (1.0 / 15.0 - sh * (1.0 / 17.0))));
  dd p = {T, 0.0};
  p = dd_add(dd{-kInvOdd[2][0], -kInvOdd[2][1]}, dd_mul(s, p));
  p = dd_add(dd{kInvOdd[1][0], kInvOdd[1][1]}, dd_mul(s, p));
  p = dd_add(dd{-kInvOdd[0][0], -kInvOdd[0][1]}, dd_mul(s, p));
  p = dd_add(dd{1.0, 0.0}, dd_mul(s, p));
  dd at_u = dd_mul(u, p);

  dd theta = dd_add(dd{kAtanTable[i][0], kAtanTable[i][1]}, at_u);
  if (swap) theta = dd_sub(dd{kHalfPi[0], kHalfPi[1]}, theta);
  if (std::signbit(x)) theta = dd_sub(dd{kPi[0], kPi[1]}, theta);
  theta = fast_two_sum(theta.hi, theta.lo);

  double hi = theta.hi;
  double lo = theta.lo;
  bool uncertain = false;
  if (hi == 0.0 || is_power_of_two(hi)) {
    uncertain = true;
  } else {
    double frac = std::fabs(lo) / ulp_of(hi);  // in [0, 0.5]
    if (std::fabs(frac - 0.5) < kMidpointMargin) uncertain = true;
  }
  double r = std::signbit(y) ? -hi : hi;
  return {r, uncertain};
}

}  // namespace dio
