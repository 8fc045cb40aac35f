// Detector geometry shared by the gfx950 builder kernels and the host patch
// step. Compiled with -ffp-contract=off: every expression below evaluates
// left to right exactly as the reference writes it, so the device, the host
// and the reference produce the same bits.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

#include "atan2_cert.hpp"
#include "sin_cert.hpp"

namespace dio {

constexpr double kEps = 2.220446049250313080847263336181640625e-16;  // f64::EPSILON
constexpr double kPiD = 3.141592653589793115997963468544185161590576171875;
constexpr double kTauD = 6.28318530717958623199592693708837032318115234375;

// Poni parameters plus sin/cos of the rotations. The sines and cosines are
// computed once on the host with the platform libm, as the reference does
// (geometry.rs:640-660), and passed to the kernels.
struct Geo {
  double pixel1, pixel2, poni1, poni2, distance;
  double s1, c1, s2, c2, s3, c3;
  double wavelength;
  double q_scale;  // 4*pi/(wavelength * 1e10 or 1e9) for q units
  int unit;        // 0 tth rad, 1 tth deg, 2 q 1/A, 3 q 1/nm
  unsigned dim1, dim2;
};

// geometry.rs:663-681.
__host__ __device__ inline double f_x1(double p1, double p2, double p3, const Geo& g) {
  return p1 * g.c2 * g.c3 + p2 * (g.c3 * g.s1 * g.s2 - g.c1 * g.s3) -
         p3 * (g.c1 * g.c3 * g.s2 + g.s1 * g.s3);
}
__host__ __device__ inline double f_x2(double p1, double p2, double p3, const Geo& g) {
  return p1 * g.c2 * g.s3 + p2 * (g.c1 * g.c3 + g.s1 * g.s2 * g.s3) -
         p3 * (-(g.c3 * g.s1) + g.c1 * g.s2 * g.s3);
}
__host__ __device__ inline double f_x3(double p1, double p2, double p3, const Geo& g) {
  return p1 * g.s2 - p2 * g.c2 * g.s1 + p3 * g.c1 * g.c2;
}

// Sample-frame position of detector point (d1 rows, d2 columns, pixel units):
// p1 = d1 * pixel1 - poni1, p2 = d2 * pixel2 - poni2 (geometry.rs:288-365).
__host__ __device__ inline void point_xyz(const Geo& g, double d1, double d2, double& x,
                                          double& y, double& z) {
  double p1 = d1 * g.pixel1 - g.poni1;
  double p2 = d2 * g.pixel2 - g.poni2;
  x = f_x1(p1, p2, g.distance, g);
  y = f_x2(p1, p2, g.distance, g);
  z = f_x3(p1, p2, g.distance, g);
}

// Arguments of tth = atan2(sqrt(y*y + x*x), z) and chi = atan2(x, y)
// (geometry.rs:684-692).
struct AngleArgs {
  double ty, tx, cy, cx;
};
__host__ __device__ inline AngleArgs angle_args(const Geo& g, double d1, double d2) {
  double x, y, z;
  point_xyz(g, d1, d2, x, y, z);
  return {std::sqrt(y * y + x * x), z, x, y};
}

// Both angles, certified. uncertain bit 1: tth, bit 2: chi.
struct Angles {
  double tth, chi;
  int uncertain;
};
__host__ __device__ inline Angles angles_cert(const Geo& g, double d1, double d2) {
  AngleArgs a = angle_args(g, d1, d2);
  Atan2Result t = atan2_cert(a.ty, a.tx);
  Atan2Result c = atan2_cert(a.cy, a.cx);
  return {t.value, c.value, (t.uncertain ? 1 : 0) | (c.uncertain ? 2 : 0)};
}

// integration.rs:171-180 (RadialConverter::convert): 2theta in the radial
// unit: radians as they are, degrees as f64::to_degrees, q as
// q_scale * sin(2theta * 0.5). For q the device takes the certified sine
// (sin_cert.hpp): `uncertain` marks values the host recomputes with glibc
// (unit_host), `sin` is the sine whose neighbouring doubles bound glibc's.
struct UnitVal {
  double value;
  double sin;
  bool uncertain;
};
__host__ __device__ inline UnitVal unit_cert(const Geo& g, double tth) {
  if (g.unit == 0) return {tth, 0.0, false};
  if (g.unit == 1) return {tth * (180.0 / kPiD), 0.0, false};
  const SinResult r = sin_cert(tth * 0.5);
  return {g.q_scale * r.value, r.value, r.uncertain};
}

// The same conversion on the host with glibc, the reference's libm.
inline double unit_host(const Geo& g, double tth) {
  if (g.unit == 0) return tth;
  if (g.unit == 1) return tth * (180.0 / kPiD);
  return g.q_scale * ::sin(tth * 0.5);
}

// Rust `f64 as isize` / `as usize`: saturating, NaN -> 0.
__host__ __device__ inline long long sat_isize(double v) {
  if (v != v) return 0;
  if (v >= 9223372036854775807.0) return 9223372036854775807LL;
  if (v <= -9223372036854775808.0) return -9223372036854775807LL - 1;
  return static_cast<long long>(v);
}
__host__ __device__ inline unsigned long long sat_usize(double v) {
  if (!(v > 0.0)) return 0;
  if (v >= 18446744073709551615.0) return 18446744073709551615ULL;
  return static_cast<unsigned long long>(v);
}

// Rust f64::rem_euclid.
__host__ __device__ inline double rem_euclid(double a, double b) {
  double r = std::fmod(a, b);
  return r < 0.0 ? r + std::fabs(b) : r;
}
// pyfai_1d.rs:180-186.
__host__ __device__ inline double wrap_tau(double x) {
  if (x >= 0.0 && x < kTauD) return x;
  return rem_euclid(x, kTauD);
}

}  // namespace dio
