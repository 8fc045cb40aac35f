// Double-double (unevaluated sum hi + lo) arithmetic for the certified
// transcendental evaluations of the table builder.
//
// Every routine uses only IEEE-754 round-to-nearest +, -, *, / and fused
// multiply-add, so the host and the gfx950 device produce identical bits
// when the translation unit is compiled with -ffp-contract=off.
// Algorithms: TwoSum (Knuth), FastTwoSum (Dekker), TwoProd via FMA, and the
// accurate double-double add/mul/div of Joldes, Muller & Popescu,
// "Tight and rigorous error bounds for basic building blocks of double-word
// arithmetic", ACM TOMS 44(2), 2017 (Algorithms 6, 12 and 17 / 18).
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace dio {

struct dd {
  double hi;
  double lo;
};

__host__ __device__ inline dd two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}

__host__ __device__ inline dd fast_two_sum(double a, double b) {
  double s = a + b;
  double e = b - (s - a);
  return {s, e};
}

__host__ __device__ inline dd two_prod(double a, double b) {
  double p = a * b;
  double e = std::fma(a, b, -p);
  return {p, e};
}

// DWPlusDW (accurate version, relative error <= 3u^2 + 13u^3).
__host__ __device__ inline dd dd_add(dd x, dd y) {
  dd s = two_sum(x.hi, y.hi);
  dd t = two_sum(x.lo, y.lo);
  double c = s.lo + t.hi;
  dd v = fast_two_sum(s.hi, c);
  double w = t.lo + v.lo;
  return fast_two_sum(v.hi, w);
}

__host__ __device__ inline dd dd_neg(dd x) { return {-x.hi, -x.lo}; }

__host__ __device__ inline dd dd_sub(dd x, dd y) { return dd_add(x, dd_neg(y)); }

// DWTimesDW3 (relative error <= 4u^2).
__host__ __device__ inline dd dd_mul(dd x, dd y) {
  dd c = two_prod(x.hi, y.hi);
  double tl0 = x.lo * y.lo;
  double tl1 = std::fma(x.hi, y.lo, tl0);
  double cl2 = std::fma(x.lo, y.hi, tl1);
  double cl3 = c.lo + cl2;
  return fast_two_sum(c.hi, cl3);
}

// DWDivDW3 (relative error <= 9.8u^2).
__host__ __device__ inline dd dd_div(dd x, dd y) {
  double th = 1.0 / y.hi;
  double rh = std::fma(-y.hi, th, 1.0);
  double rl = -(y.lo * th);
  dd e = fast_two_sum(rh, rl);
  dd d = dd_mul(e, dd{th, 0.0});
  dd m = dd_add(d, dd{th, 0.0});  // m ~ 1/y
  return dd_mul(x, m);
}

}  // namespace dio
