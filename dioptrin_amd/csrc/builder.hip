// Device lookup-table builder (gfx950). Compiled with -ffp-contract=off.
//
// Pipeline for a 1D table (reference: src/lut/pixel_split.rs:190-455,
// src/lut/counting_csr.rs:54-134, src/lut/pyfai_1d.rs:188-252,
// src/lut/prepared.rs:163-262):
//   1. corner-grid and pixel-centre angles with the certified atan2; the few
//      uncertain angles are recomputed with the host libm and patched in;
//   2. data-derived radial range by device min/max reduction;
//   3. per-pixel entry counts (pixel splitting: Sutherland-Hodgman clip of
//      the pixel quadrilateral against each touched bin); beam pixels are
//      set aside and integrated by 128x128x4 detector-plane sampling;
//   4. exclusive scan of the counts and a pixel-major fill of
//      (bin, pixel, FP64 weight);
//   5. stable radix sort by bin: rows list pixels in ascending order, the
//      order of the reference's builders; exact row pointers from a bin
//      histogram and a scan;
//   6. mask compaction (order kept), per-pixel normalization and strict
//      sequential FP64 denominators per row.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

#include "builder.hpp"

namespace dio {

namespace {

constexpr unsigned kThreads = 256;
constexpr int kClipCap = 20;  // 4 stages on a quadrilateral: <= 19 vertices
constexpr unsigned kSamplesPerPixel = 128 * 128 * 4;
constexpr unsigned kMaxSingular = 4096;

// ------------------------------------------------------------ angle grids

__global__ void k_grid_angles(Geo g, uint32_t rows, uint32_t cols, double off, double* tth,
                              double* chi, uint32_t* flist, uint32_t* fcount) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t n = (uint64_t)rows * cols;
  if (i >= n) return;
  uint32_t r = (uint32_t)(i / cols), c = (uint32_t)(i % cols);
  Angles a = angles_cert(g, (double)r + off, (double)c + off);
  tth[i] = a.tth;
  chi[i] = a.chi;
  if (a.uncertain) flist[atomicAdd(fcount, 1u)] = (uint32_t)i;
}

__global__ void k_patch(const uint32_t* idx, const double* tv, const double* cv, uint32_t n,
                        double* tth, double* chi) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  tth[idx[i]] = tv[i];
  chi[idx[i]] = cv[i];
}

// 2theta in the plan's radial unit; q values whose certified sine is
// uncertain are listed for the host (out may alias tth).
__global__ void k_to_unit(Geo g, const double* tth, double* out, uint64_t n, uint32_t* flist,
                          uint32_t* fcount) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const UnitVal u = unit_cert(g, tth[i]);
  out[i] = u.value;
  if (u.uncertain) flist[atomicAdd(fcount, 1u)] = (uint32_t)i;
}

__global__ void k_gather_f64(const double* v, const uint32_t* idx, uint32_t n, double* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = v[idx[i]];
}

__global__ void k_scatter_f64(const uint32_t* idx, const double* val, uint32_t n, double* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[idx[i]] = val[i];
}

// Block-level min/max partials (src/math/mod.rs:8-32 semantics for finite
// data: the order of a min/max reduction does not change its value).
__global__ void k_minmax(const double* v, uint64_t n, double* pmin, double* pmax) {
  __shared__ double smin[kThreads], smax[kThreads];
  double lo = INFINITY, hi = -INFINITY;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    double x = v[i];
    lo = x < lo ? x : lo;
    hi = x > hi ? x : hi;
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (unsigned s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      double a = smin[threadIdx.x + s], b = smax[threadIdx.x + s];
      smin[threadIdx.x] = a < smin[threadIdx.x] ? a : smin[threadIdx.x];
      smax[threadIdx.x] = b > smax[threadIdx.x] ? b : smax[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pmin[blockIdx.x] = smin[0];
    pmax[blockIdx.x] = smax[0];
  }
}

// ------------------------------------------------------- pixel splitting

struct PsParams {
  Geo g;
  const double* rc;     // corner radial (unit)
  const double* ac;     // corner chi
  const double* rad;    // centre radial (unit)
  const double* chi;    // centre chi
  uint32_t n;           // bins
  double min, step;
  int grid_ok;          // 1D sampling grid representable
  double glo, ghi;      // sampling grid radial range (min, min + n*step)
  int clamp_low;
};

// pixel_split.rs:14-25.
__device__ inline double polygon_area(const double (*poly)[2], int n) {
  if (n < 3) return 0.0;
  double area = 0.0;
  for (int i = 0; i < n; i++) {
    int j = (i + 1 == n) ? 0 : i + 1;
    area += poly[i][0] * poly[j][1] - poly[j][0] * poly[i][1];
  }
  return 0.5 * area;
}

// pixel_split.rs:28-45.
__device__ inline void intersect(const double* p0, const double* p1, int axis, double e,
                                 double* out) {
  int o = 1 - axis;
  double d = p1[axis] - p0[axis];
  if (fabs(d) < kEps) {
    out[axis] = e;
    out[o] = p0[o];
    return;
  }
  double t = (e - p0[axis]) / d;
  out[axis] = e;
  out[o] = p0[o] + t * (p1[o] - p0[o]);
}

// One Sutherland-Hodgman stage (pixel_split.rs:48-137).
__device__ inline int clip_stage(const double (*in)[2], int n, int axis, bool keep_ge, double e,
                                 double (*out)[2]) {
  if (n == 0) return 0;
  int m = 0;
  const double* prev = in[n - 1];
  bool prev_in = keep_ge ? (prev[axis] >= e) : (prev[axis] <= e);
  for (int i = 0; i < n; i++) {
    const double* cur = in[i];
    bool cur_in = keep_ge ? (cur[axis] >= e) : (cur[axis] <= e);
    if (cur_in) {
      if (!prev_in && m < kClipCap) intersect(prev, cur, axis, e, out[m++]);
      if (m < kClipCap) {
        out[m][0] = cur[0];
        out[m][1] = cur[1];
        m++;
      }
    } else if (prev_in && m < kClipCap) {
      intersect(prev, cur, axis, e, out[m++]);
    }
    prev = cur;
    prev_in = cur_in;
  }
  return m;
}

// pixel_split.rs:140-163: left, right, bottom, top.
__device__ inline double clipped_area(const double (*poly)[2], double xmin, double xmax,
                                      double ymin, double ymax) {
  double a[kClipCap][2], b[kClipCap][2];
  int m = clip_stage(poly, 4, 0, true, xmin, a);
  if (m == 0) return 0.0;
  m = clip_stage(a, m, 0, false, xmax, b);
  if (m == 0) return 0.0;
  m = clip_stage(b, m, 1, true, ymin, a);
  if (m == 0) return 0.0;
  m = clip_stage(a, m, 1, false, ymax, b);
  return fabs(polygon_area(b, m));
}

// pixel_split.rs:166-175.
__device__ inline double unwrap_angle(double angle, double reference) {
  double a = angle;
  double delta = a - reference;
  if (delta > kPiD) a -= 2.0 * kPiD;
  else if (delta < -kPiD) a += 2.0 * kPiD;
  return a;
}

// pyfai_1d.rs:188-201.
__device__ inline bool singular4(const double (*poly)[2]) {
  double winding = 0.0;
  for (int i = 0; i < 4; i++) {
    if (poly[i][0] == 0.0) return true;
    double delta = wrap_tau(poly[(i + 1) & 3][1] - poly[i][1] + kPiD) - kPiD;
    if (fabs(delta) >= kPiD - 64.0 * kEps) return true;
    winding += delta;
  }
  return fabs(winding) > kPiD;
}

// The entries of one pixel, in the reference's order (pixel_split.rs:208-326).
// Returns false for a beam pixel (handled by sampling).
template <class Emit>
__device__ bool ps_pixel(const PsParams& P, uint32_t idx, Emit& emit) {
  uint32_t dim2 = P.g.dim2;
  uint64_t w = (uint64_t)dim2 + 1;
  uint64_t b = (uint64_t)(idx / dim2) * w + idx % dim2;
  uint64_t i00 = b, i01 = b + 1, i11 = b + w + 1, i10 = b + w;
  double r00 = P.rc[i00], r01 = P.rc[i01], r10 = P.rc[i10], r11 = P.rc[i11];
  if (P.grid_ok) {
    double raw[4][2] = {{r00, P.ac[i00]}, {r01, P.ac[i01]}, {r11, P.ac[i11]}, {r10, P.ac[i10]}};
    if (singular4(raw)) return false;
  }
  double ca = P.chi[idx];
  double a00 = unwrap_angle(P.ac[i00], ca);
  double a01 = unwrap_angle(P.ac[i01], ca);
  double a10 = unwrap_angle(P.ac[i10], ca);
  double a11 = unwrap_angle(P.ac[i11], ca);
  double poly[4][2] = {{r00, a00}, {r01, a01}, {r11, a11}, {r10, a10}};
  double area_total = fabs(polygon_area(poly, 4));
  if (area_total <= 0.0) {
    double pos = (P.rad[idx] - P.min) / P.step;
    if (pos < 0.0) return true;
    long long bin = sat_isize(floor(pos));
    if (bin >= 0 && (unsigned long long)bin < P.n) emit((uint32_t)bin, 1.0);
    return true;
  }
  double rmin = r00, rmax = r00;
  double rv[3] = {r01, r10, r11};
  for (int k = 0; k < 3; k++) {
    if (rv[k] < rmin) rmin = rv[k];
    if (rv[k] > rmax) rmax = rv[k];
  }
  long long start = sat_isize(floor((rmin - P.min) / P.step));
  long long end = sat_isize(floor((rmax - P.min) / P.step));
  if (end < 0 || start >= (long long)P.n) return true;
  if (start < 0) start = 0;
  if (end >= (long long)P.n) end = (long long)P.n - 1;
  double amin = a00, amax = a00;
  double av[3] = {a01, a10, a11};
  for (int k = 0; k < 3; k++) {
    if (av[k] < amin) amin = av[k];
    if (av[k] > amax) amax = av[k];
  }
  for (long long bb = start; bb <= end; bb++) {
    double bin_min = P.min + (double)bb * P.step;
    double bin_max = bin_min + P.step;
    double area = clipped_area(poly, bin_min, bin_max, amin, amax);
    if (area > 0.0) emit((uint32_t)bb, area / area_total);
  }
  return true;
}

struct CountEmit {
  uint32_t n = 0;
  __device__ void operator()(uint32_t, double) { n++; }
};
struct FillEmit {
  uint32_t* bin;
  uint32_t* pix;
  double* w;
  uint64_t at;
  uint32_t p;
  __device__ void operator()(uint32_t b, double v) {
    bin[at] = b;
    pix[at] = p;
    w[at] = v;
    at++;
  }
};

__global__ void k_ps_count(PsParams P, uint32_t npix, uint32_t* cnt, uint32_t* sing,
                           uint32_t* nsing) {
  uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npix) return;
  CountEmit e;
  if (!ps_pixel(P, idx, e)) {
    uint32_t s = atomicAdd(nsing, 1u);
    if (s < kMaxSingular) sing[s] = idx;
    cnt[idx] = 0;
    return;
  }
  cnt[idx] = e.n;
}

// --------------------------------------------------------- beam sampling

struct SampleGrid {
  uint32_t n;
  double r0, r1;
  int clamp_low;
};

__device__ inline void sample_point(const Geo& g, uint32_t pixel, uint32_t s, double& d1,
                                    double& d2) {
  const double dy[4] = {0.25, 0.5, 0.75, 0.5};
  const double dx[4] = {0.5, 0.25, 0.5, 0.75};
  uint32_t row = s >> 9, col = (s >> 2) & 127u, o = s & 3u;
  d1 = (double)(pixel / g.dim2) + ((double)row + dy[o]) / 128.0;
  d2 = (double)(pixel % g.dim2) + ((double)col + dx[o]) / 128.0;
}

__global__ void k_sing_samples(Geo g, const uint32_t* sing, uint32_t nsing, double* stth,
                               double* schi, uint32_t* flist, uint32_t* fcount) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)nsing * kSamplesPerPixel) return;
  uint32_t slot = (uint32_t)(i / kSamplesPerPixel), s = (uint32_t)(i % kSamplesPerPixel);
  double d1, d2;
  sample_point(g, sing[slot], s, d1, d2);
  Angles a = angles_cert(g, d1, d2);
  stth[i] = a.tth;
  schi[i] = a.chi;
  if (a.uncertain) flist[atomicAdd(fcount, 1u)] = (uint32_t)i;
}

// pyfai_1d.rs:228-251 binning of the samples on the 1D grid (n radial bins,
// one azimuthal bin over (-pi, pi)).
__global__ void k_sing_bin(SampleGrid G, uint32_t nsing, const double* srad,
                           const double* schi, uint32_t* hits) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)nsing * kSamplesPerPixel) return;
  uint32_t slot = (uint32_t)(i / kSamplesPerPixel);
  double r = srad[i];
  double a0 = -kPiD, a1 = kPiD;
  double a = a0 + rem_euclid(schi[i] - a0, kTauD);
  if ((r < G.r0 && !G.clamp_low) || r >= G.r1 || a >= a1) return;
  unsigned long long ir = r < G.r0 ? 0ULL : sat_usize((r - G.r0) / (G.r1 - G.r0) * (double)G.n);
  unsigned long long ia = sat_usize((a - a0) / (a1 - a0) * 1.0);
  if (ir < G.n && ia < 1) atomicAdd(&hits[(uint64_t)slot * G.n + ir], 1u);
}

__global__ void k_sing_count(const uint32_t* hits, uint32_t n, uint32_t nsing,
                             const uint32_t* sing, uint32_t* cnt) {
  uint32_t slot = blockIdx.x;
  if (slot >= nsing) return;
  __shared__ uint32_t part[kThreads];
  uint32_t c = 0;
  for (uint32_t b = threadIdx.x; b < n; b += blockDim.x) c += hits[(uint64_t)slot * n + b] != 0;
  part[threadIdx.x] = c;
  __syncthreads();
  for (unsigned s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) cnt[sing[slot]] = part[0];
}

// Entries of beam pixels: bins ascending, weight hits / 65536.
__global__ void k_sing_fill(const uint32_t* hits, uint32_t n, uint32_t nsing,
                            const uint32_t* sing, const uint32_t* off, uint32_t* bin,
                            uint32_t* pix, double* w) {
  uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nsing) return;
  uint32_t p = sing[slot];
  uint64_t at = off[p];
  for (uint32_t b = 0; b < n; b++) {
    uint32_t h = hits[(uint64_t)slot * n + b];
    if (h) {
      bin[at] = b;
      pix[at] = p;
      w[at] = (double)h / (double)kSamplesPerPixel;
      at++;
    }
  }
}

__global__ void k_ps_fill(PsParams P, uint32_t npix, const uint32_t* off, uint32_t* bin,
                          uint32_t* pix, double* w) {
  uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npix) return;
  FillEmit e{bin, pix, w, off[idx], idx};
  ps_pixel(P, idx, e);
}

// ------------------------------------------------------------- counting

// counting_csr.rs:54-134: one unweighted entry per in-range pixel centre.
__global__ void k_count_bins(const double* rad, uint32_t npix, uint32_t n, double min,
                             double max, double step, int has_range, uint32_t* cnt,
                             uint32_t* pbin) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  double r = rad[i];
  if (has_range && !(r >= min && r <= max)) {
    cnt[i] = 0;
    return;
  }
  unsigned long long b = sat_usize((r - min) / step);
  if (b > n - 1) b = n - 1;
  cnt[i] = 1;
  pbin[i] = (uint32_t)b;
}

__global__ void k_count_fill(const uint32_t* cnt, const uint32_t* pbin, const uint32_t* off,
                             uint32_t npix, uint32_t* bin, uint32_t* pix, double* w) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix || !cnt[i]) return;
  uint32_t at = off[i];
  bin[at] = pbin[i];
  pix[at] = i;
  w[at] = 1.0;
}

// ------------------------------------------------------------ 2D grids

// integration/grid.rs:26-222 (IntegrationGrid2d) with its clamp flags.
struct Grid2 {
  uint32_t nr, na;
  double r0, r1, a0, a1;
  int clamp_radial_end, clamp_index_rounding, clamp_radial_outside, clamp_azimuthal_outside;
  __host__ __device__ double rstep() const { return (r1 - r0) / (double)nr; }
  __host__ __device__ double astep() const { return (a1 - a0) / (double)na; }
};

// grid.rs:176-221 (IntegrationGrid2d::bin); -1 when the value has no bin.
__host__ __device__ inline long long grid_bin(const Grid2& g, double radial, double az) {
  if (!std::isfinite(radial) || !std::isfinite(az)) return -1;
  if (!g.clamp_radial_outside &&
      (radial < g.r0 || radial > g.r1 || (radial == g.r1 && !g.clamp_radial_end)))
    return -1;
  double offset;
  if (g.clamp_azimuthal_outside) {
    offset = az - g.a0;
    double hi = g.a1 - g.a0;
    offset = offset < 0.0 ? 0.0 : (offset > hi ? hi : offset);
  } else if (g.clamp_index_rounding) {
    double wrapped = (az >= g.a0 && az < g.a1) ? az : g.a0 + rem_euclid(az - g.a0, kTauD);
    if (wrapped >= g.a1 && g.a1 - g.a0 < kTauD) return -1;
    offset = wrapped - g.a0;
  } else {
    double a = g.a0 + rem_euclid(az - g.a0, kTauD);
    if (a >= g.a1) return -1;
    offset = a - g.a0;
  }
  unsigned long long r = sat_usize((radial - g.r0) / g.rstep());
  unsigned long long a = sat_usize(offset / g.astep());
  if (g.clamp_radial_end || g.clamp_index_rounding || g.clamp_radial_outside)
    if (r > g.nr - 1) r = g.nr - 1;
  if (g.clamp_index_rounding || g.clamp_azimuthal_outside)
    if (a > g.na - 1) a = g.na - 1;
  if (r < g.nr && a < g.na) return (long long)(a * g.nr + r);
  return -1;
}

struct Ps2Params {
  Geo g;
  Grid2 G;
  const double* rc;
  const double* ac;
  const double* rad;
  const double* chi;
};

// shared_grid.rs:152-200 (split_polygon): the three periodic copies of the
// polygon, bins (ar, rr) in order; a bin hit by several copies gets the sum
// of its weights in copy order (-2pi, 0, +2pi), as the reference merges them.
template <class Emit>
__device__ void split2d(const Ps2Params& P, const double (*poly)[2], double area, Emit& emit) {
  const Grid2& G = P.G;
  double r0 = G.r0, a0 = G.a0, dr = G.rstep(), da = G.astep();
  double rmin = INFINITY, rmax = -INFINITY;
  for (int i = 0; i < 4; i++) {
    rmin = fmin(rmin, poly[i][0]);
    rmax = fmax(rmax, poly[i][0]);
  }
  long long rs = sat_isize(floor((rmin - r0) / dr));
  long long re = sat_isize(floor((rmax - r0) / dr));
  if (rs < 0) rs = 0;
  if (re > (long long)G.nr - 1) re = (long long)G.nr - 1;
  const double shifts[3] = {-kTauD, 0.0, kTauD};
  long long as[3], ae[3];
  long long lo = 0x7fffffffffffffffLL, hi = -1;
  for (int k = 0; k < 3; k++) {
    double amin = INFINITY, amax = -INFINITY;
    for (int i = 0; i < 4; i++) {
      double a = poly[i][1] + shifts[k];
      amin = fmin(amin, a);
      amax = fmax(amax, a);
    }
    as[k] = sat_isize(floor((amin - a0) / da));
    ae[k] = sat_isize(floor((amax - a0) / da));
    if (as[k] < 0) as[k] = 0;
    if (ae[k] > (long long)G.na - 1) ae[k] = (long long)G.na - 1;
    if (as[k] <= ae[k]) {
      lo = as[k] < lo ? as[k] : lo;
      hi = ae[k] > hi ? ae[k] : hi;
    }
  }
  for (long long ar = lo; ar <= hi; ar++) {
    bool any = false;
    for (int k = 0; k < 3; k++) any |= (as[k] <= ar && ar <= ae[k]);
    if (!any) continue;
    for (long long rr = rs; rr <= re; rr++) {
      double r = r0 + (double)rr * dr;
      double a = a0 + (double)ar * da;
      double acc = 0.0;
      bool found = false;
      for (int k = 0; k < 3; k++) {
        if (!(as[k] <= ar && ar <= ae[k])) continue;
        double ps[4][2];
        for (int i = 0; i < 4; i++) {
          ps[i][0] = poly[i][0];
          ps[i][1] = poly[i][1] + shifts[k];
        }
        double w = clipped_area(ps, r, r + dr, a, a + da) / area;
        if (w > 0.0) {
          acc = found ? acc + w : w;
          found = true;
        }
      }
      if (found) emit((uint32_t)((uint64_t)ar * G.nr + (uint64_t)rr), acc);
    }
  }
}

// shared_grid.rs:58-130 for one pixel of a PixelSplit 2D table. Returns false
// for a pixel around or on the polar axis (integrated by sampling).
template <class Emit>
__device__ bool ps2d_pixel(const Ps2Params& P, uint32_t idx, Emit& emit) {
  uint32_t dim2 = P.g.dim2;
  uint64_t w = (uint64_t)dim2 + 1;
  uint64_t b = (uint64_t)(idx / dim2) * w + idx % dim2;
  uint64_t ci[4] = {b, b + 1, b + w + 1, b + w};
  bool axis_on_edge = false, zero_r = false;
  double winding = 0.0;
  for (int i = 0; i < 4; i++) {
    double delta = P.ac[ci[(i + 1) & 3]] - P.ac[ci[i]];
    delta = rem_euclid(delta + kPiD, kTauD) - kPiD;
    axis_on_edge |= fabs(delta) >= kPiD - 64.0 * kEps;
    winding += delta;
    zero_r |= P.rc[ci[i]] == 0.0;
  }
  if (fabs(winding) > kPiD || axis_on_edge || zero_r) return false;
  double center = P.G.a0 + rem_euclid(P.chi[idx] - P.G.a0, kTauD);
  double poly[4][2];
  for (int i = 0; i < 4; i++) {
    poly[i][0] = P.rc[ci[i]];
    poly[i][1] = center + rem_euclid(P.ac[ci[i]] - center + kPiD, kTauD) - kPiD;
  }
  double area = fabs(polygon_area(poly, 4));
  if (area > 0.0) {
    split2d(P, poly, area, emit);
  } else {
    long long bin = grid_bin(P.G, P.rad[idx], P.chi[idx]);
    if (bin >= 0) emit((uint32_t)bin, 1.0);
  }
  return true;
}

__global__ void k_ps2d_count(Ps2Params P, uint32_t npix, uint32_t* cnt, uint32_t* sing,
                             uint32_t* nsing) {
  uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npix) return;
  CountEmit e;
  if (!ps2d_pixel(P, idx, e)) {
    uint32_t s = atomicAdd(nsing, 1u);
    if (s < kMaxSingular) sing[s] = idx;
    cnt[idx] = 0;
    return;
  }
  cnt[idx] = e.n;
}

__global__ void k_ps2d_fill(Ps2Params P, uint32_t npix, const uint32_t* off, uint32_t* bin,
                            uint32_t* pix, double* w) {
  uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npix) return;
  FillEmit e{bin, pix, w, off[idx], idx};
  ps2d_pixel(P, idx, e);
}

// shared_grid.rs:202-232: 2D sample coordinates d1 = row + (i + o1)/128,
// d2 = ((col * 128 + j) + o2)/128.
__host__ __device__ inline void sample_point_2d(uint32_t dim2, uint32_t pixel, uint32_t s,
                                                double& d1, double& d2) {
  const double o1[4] = {0.25, 0.5, 0.75, 0.5};
  const double o2[4] = {0.5, 0.25, 0.5, 0.75};
  uint32_t i = s >> 9, j = (s >> 2) & 127u, o = s & 3u;
  uint32_t row = pixel / dim2, col = pixel % dim2;
  d1 = (double)row + ((double)i + o1[o]) / 128.0;
  d2 = ((double)((uint64_t)col * 128 + j) + o2[o]) / 128.0;
}

__global__ void k_sing_samples2d(Geo g, const uint32_t* sing, uint32_t nsing, double* stth,
                                 double* schi, uint32_t* flist, uint32_t* fcount) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)nsing * kSamplesPerPixel) return;
  uint32_t slot = (uint32_t)(i / kSamplesPerPixel), s = (uint32_t)(i % kSamplesPerPixel);
  double d1, d2;
  sample_point_2d(g.dim2, sing[slot], s, d1, d2);
  Angles a = angles_cert(g, d1, d2);
  stth[i] = a.tth;
  schi[i] = a.chi;
  if (a.uncertain) flist[atomicAdd(fcount, 1u)] = (uint32_t)i;
}

__global__ void k_sing_bin2d(Grid2 G, uint32_t nsing, const double* srad,
                             const double* schi, uint32_t* hits) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)nsing * kSamplesPerPixel) return;
  uint32_t slot = (uint32_t)(i / kSamplesPerPixel);
  long long b = grid_bin(G, srad[i], schi[i]);
  if (b >= 0) atomicAdd(&hits[(uint64_t)slot * G.nr * G.na + (uint64_t)b], 1u);
}

// Counting 2D (shared_grid.rs:52-57): the bin of the pixel centre.
__global__ void k_count2d_bins(Grid2 G, const double* rad, const double* chi, uint32_t npix,
                               uint32_t* cnt, uint32_t* pbin) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  long long b = grid_bin(G, rad[i], chi[i]);
  cnt[i] = b >= 0;
  if (b >= 0) pbin[i] = (uint32_t)b;
}

// ------------------------------------------------------------ supersampling

// Bin rule of one subpixel (supersampled.rs:517-595, 1381-1477): 1D clamped
// (automatic range) or in-range (explicit range), 2D the grid's bin.
struct SsBin {
  Geo g;
  int twod;
  uint32_t n1;
  double min1, max1, step1;
  int clamped1;
  Grid2 G;
  uint32_t nss;
  int inv;  // 1: rf + (i + 0.5) * inv_ss coordinates; 0: (i + 0.5) / n_ss
};

// r: the subpixel's radial value in the plan's unit.
__host__ __device__ inline long long ss_bin(const SsBin& B, double r, double chi) {
  if (B.twod) return grid_bin(B.G, r, chi);
  if (!B.clamped1 && !(r >= B.min1 && r <= B.max1)) return -1;
  unsigned long long u = sat_usize((r - B.min1) / B.step1);
  if (u > B.n1 - 1) u = B.n1 - 1;
  return (long long)u;
}

// supersampled.rs:39-43 (/ n_ss) and 310-311, 438-440 (* inv_ss).
__host__ __device__ inline void ss_point(const SsBin& B, uint64_t sub, double& d1, double& d2) {
  uint32_t nss = B.nss, nss2 = nss * nss;
  uint64_t pixel = sub / nss2;
  uint32_t k = (uint32_t)(sub % nss2), i = k / nss, j = k % nss;
  uint64_t row = pixel / B.g.dim2, col = pixel % B.g.dim2;
  double gc = (double)(col * nss + j);
  if (B.inv) {
    double inv_ss = 1.0 / (double)nss;
    d1 = (double)row + ((double)i + 0.5) * inv_ss;
    d2 = (gc + 0.5) * inv_ss;
  } else {
    d1 = (double)row + ((double)i + 0.5) / (double)nss;
    d2 = (gc + 0.5) / (double)nss;
  }
}

__device__ inline double next_up(double v) { return nextafter(v, INFINITY); }
__device__ inline double next_down(double v) { return nextafter(v, -INFINITY); }

// Bin of every subpixel. An uncertain angle matters only if one of its two
// possible roundings changes the bin; such subpixels are listed for the host.
__global__ void k_ss_bins(SsBin B, uint64_t nsub, uint32_t* bins, uint32_t* flist,
                          uint32_t* fcount, uint32_t fcap) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsub) return;
  double d1, d2;
  ss_point(B, t, d1, d2);
  Angles a = angles_cert(B.g, d1, d2);
  const UnitVal u = unit_cert(B.g, a.tth);
  long long b = ss_bin(B, u.value, a.chi);
  bool ambiguous = false;
  if (a.uncertain || u.uncertain) {
    // The host's values are within one double of the certified ones: the
    // subpixel is ambiguous when a neighbouring 2theta, chi or sine (q
    // units) changes its bin.
    double tc[3] = {a.tth, next_up(a.tth), next_down(a.tth)};
    double cc[3] = {a.chi, next_up(a.chi), next_down(a.chi)};
    int nt = (a.uncertain & 1) ? 3 : 1, nc = (B.twod && (a.uncertain & 2)) ? 3 : 1;
    for (int x = 0; x < nt && !ambiguous; x++) {
      const UnitVal ux = x == 0 ? u : unit_cert(B.g, tc[x]);
      double rc[3] = {ux.value, 0.0, 0.0};
      int nr = 1;
      if (ux.uncertain) {
        if (ux.sin == 0.0) {  // outside the certified range: the host decides
          ambiguous = true;
          break;
        }
        rc[1] = B.g.q_scale * next_up(ux.sin);
        rc[2] = B.g.q_scale * next_down(ux.sin);
        nr = 3;
      }
      for (int k = 0; k < nr; k++)
        for (int y = 0; y < nc; y++) ambiguous |= ss_bin(B, rc[k], cc[y]) != b;
    }
  }
  bins[t] = b < 0 ? 0xffffffffu : (uint32_t)b;
  if (ambiguous) {
    uint32_t pos = atomicAdd(fcount, 1u);
    if (pos < fcap) flist[pos] = (uint32_t)t;
  }
}

__global__ void k_ss_patch(const uint32_t* idx, const uint32_t* val, uint32_t n, uint32_t* bins) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) bins[idx[i]] = val[i];
}

// Distinct bins among a pixel's subpixels, in first-occurrence order.
__global__ void k_ss_count(const uint32_t* bins, uint32_t npix, uint32_t nss2, uint32_t* cnt) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  const uint32_t* sb = bins + (uint64_t)p * nss2;
  uint32_t n = 0;
  for (uint32_t k = 0; k < nss2; k++) {
    uint32_t b = sb[k];
    if (b == 0xffffffffu) continue;
    bool seen = false;
    for (uint32_t q = 0; q < k && !seen; q++) seen = sb[q] == b;
    if (!seen) n++;
  }
  cnt[p] = n;
}

__global__ void k_ss_fill(const uint32_t* bins, uint32_t npix, uint32_t nss2, const uint32_t* off,
                          uint32_t* bin, uint32_t* pix, double* w) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  const uint32_t* sb = bins + (uint64_t)p * nss2;
  uint64_t at = off[p];
  for (uint32_t k = 0; k < nss2; k++) {
    uint32_t b = sb[k];
    if (b == 0xffffffffu) continue;
    bool seen = false;
    for (uint32_t q = 0; q < k && !seen; q++) seen = sb[q] == b;
    if (seen) continue;
    uint32_t c = 0;
    for (uint32_t q = k; q < nss2; q++) c += sb[q] == b;
    bin[at] = b;
    pix[at] = p;
    w[at] = (double)c / (double)nss2;
    at++;
  }
}

__global__ void k_list_angles(Geo g, const double* d1, const double* d2, uint32_t n, double* tth,
                              double* chi, uint32_t* flist, uint32_t* fcount) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Angles a = angles_cert(g, d1[i], d2[i]);
  tth[i] = a.tth;
  chi[i] = a.chi;
  if (a.uncertain) flist[atomicAdd(fcount, 1u)] = i;
}

// ------------------------------------------------------ pyFAI-compatible

struct PfParams {
  Geo g;
  const double* rc;
  const double* ac;
  const double* rad;
  const double* chi;
  uint32_t n;           // 1D bins
  double r0, r1;        // 1D radial range
  Grid2 G;              // 2D grid
};

// pyfai_1d.rs:259-354 (split1d, sector = false, sign = +1): the float32
// line-area value of relative bin j, accumulated edge by edge in order.
__device__ inline float pf1d_value(const double (*poly)[2], size_t first, size_t j) {
  float v = 0.0f;
  for (int i = 0; i < 4; i++) {
    double x0 = poly[i][0], y0 = poly[i][1];
    double x1 = poly[(i + 1) & 3][0], y1 = poly[(i + 1) & 3][1];
    if (x0 == x1) continue;
    double slope = (y1 - y0) / (x1 - x0);
    double intercept = y0 - slope * x0;
    double sign = x1 > x0 ? 1.0 : -1.0;
    double left = fmax(fmin(x0, x1), (double)(first + j));
    double right = fmin(fmax(x0, x1), (double)(first + j + 1));
    if (right > left)
      v = v + (float)(sign * (right - left) * (0.5 * slope * (right + left) + intercept));
  }
  return v;
}

template <class Emit>
__device__ bool pf1d_pixel(const PfParams& P, uint32_t idx, Emit& emit) {
  uint32_t dim2 = P.g.dim2;
  uint64_t w = (uint64_t)dim2 + 1;
  uint64_t b = (uint64_t)(idx / dim2) * w + idx % dim2;
  uint64_t ci[4] = {b, b + w, b + w + 1, b + 1};  // pyFAI order a, d, c, b
  double poly[4][2];
  for (int i = 0; i < 4; i++) {
    poly[i][0] = P.rc[ci[i]];
    poly[i][1] = P.ac[ci[i]];
  }
  if (singular4(poly)) return false;
  const double offset = 0.0;  // full azimuth (-pi, pi)
  for (int i = 0; i < 4; i++) poly[i][1] = wrap_tau(poly[i][1] - offset + kPiD) - kPiD;
  double* a = poly[0];
  double* bb = poly[1];
  double* c = poly[2];
  double* d = poly[3];
  double area = 1.0 * (0.5 * (c[0] - a[0]) * (d[1] - bb[1]) - (c[1] - a[1]) * (d[0] - bb[0]));
  if (area > 0.0) {
    for (int i = 0; i < 4; i++)
      if (poly[i][1] < 0.0) poly[i][1] += kTauD;
    double sum = 0.0;
    for (int i = 0; i < 4; i++) sum = sum + poly[i][1];
    if (sum * 0.25 > kPiD)
      for (int i = 0; i < 4; i++) poly[i][1] -= kTauD;
  }
  for (int i = 0; i < 4; i++) poly[i][1] += offset;
  double dr = (P.r1 - P.r0) / (double)P.n;
  for (int i = 0; i < 4; i++) poly[i][0] = (poly[i][0] - P.r0) / dr;
  double lo = INFINITY, hi = -INFINITY;
  for (int i = 0; i < 4; i++) {
    lo = fmin(lo, poly[i][0]);
    hi = fmax(hi, poly[i][0]);
  }
  if (hi < 0.0 || lo >= (double)P.n) return true;
  size_t first = (size_t)sat_usize(fmax(floor(lo), 0.0));
  long long e = sat_isize(floor(hi)) + 1;
  if (e < 0) e = 0;
  if (e > (long long)P.n) e = (long long)P.n;
  if (floor(lo) == floor(hi)) {
    emit((uint32_t)first, 1.0);
    return true;
  }
  size_t n = (size_t)e > first ? (size_t)e - first : 0;
  double sum = 0.0;
  for (size_t j = 0; j < n; j++) sum += (double)pf1d_value(poly, first, j);
  if (!isfinite(sum) || sum == 0.0) return true;
  double total = 0.0;
  for (size_t j = 0; j < n; j++) {
    double wv = (double)pf1d_value(poly, first, j) / sum;
    if (isfinite(wv) && wv > 0.0) total += wv;
  }
  for (size_t j = 0; j < n; j++) {
    double wv = (double)pf1d_value(poly, first, j) / sum;
    if (isfinite(wv) && wv > 0.0) emit((uint32_t)(first + j), wv / total);
  }
  return true;
}

// pyfai_split.rs:139-167 (line) for one cell: the float32 buffer value of
// (column r, row a), lines applied in order as the reference fills it.
__device__ inline float pf2d_cell(const double (*poly)[2], size_t r, size_t a) {
  float v = 0.0f;
  for (int i = 0; i < 4; i++) {
    const double* st = poly[i];
    const double* sp = poly[(i + 1) & 3];
    if (st[0] == sp[0]) continue;
    float slope = (float)((sp[1] - st[1]) / (sp[0] - st[0]));
    float intercept = (float)(sp[1] - (double)slope * sp[0]);
    double lo = fmin(st[0], sp[0]), hi = fmax(st[0], sp[0]);
    double sign = sp[0] > st[0] ? 1.0 : -1.0;
    size_t r_begin = (size_t)sat_usize(floor(lo)), r_end = (size_t)sat_usize(ceil(hi));
    if (r < r_begin || r >= r_end) continue;
    double left = fmax(lo, (double)r), right = fmin(hi, (double)(r + 1));
    float area = (float)(sign * (right - left) *
                         (0.5 * (double)slope * (right + left) + (double)intercept));
    float width = (float)(right - left);
    float remaining = fabsf(area);
    for (size_t q = 0; q <= a; q++) {
      if (remaining <= 0.0f) break;
      float inc = fminf(width, remaining);
      if (q == a) v = v + copysignf(inc, area);
      remaining = remaining - inc;
    }
  }
  return v;
}

__device__ inline double clampd(double x, double lo, double hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}

// pyfai_split.rs:35-137 (split, sign = +1). Returns 0 for "None" (the caller
// falls back to fractional clipping), 1 otherwise.
template <class Emit>
__device__ int pf2d_split(const Grid2& G, double (*poly)[2], Emit& emit) {
  double a0 = G.a0, a1 = G.a1;
  double cstart = (a0 >= -kPiD && a1 <= kPiD) ? -kPiD : (a0 >= 0.0 && a1 <= kTauD) ? 0.0 : a0;
  double offset = cstart + kPiD;
  for (int i = 0; i < 4; i++) poly[i][1] = wrap_tau(poly[i][1] - offset + kPiD) - kPiD;
  double* a = poly[0];
  double* b = poly[1];
  double* c = poly[2];
  double* d = poly[3];
  double area = 1.0 * (0.5 * (c[0] - a[0]) * (d[1] - b[1]) - (c[1] - a[1]) * (d[0] - b[0]));
  if (area > 0.0) {
    for (int i = 0; i < 4; i++)
      if (poly[i][1] < 0.0) poly[i][1] += kTauD;
    double sum = 0.0;
    for (int i = 0; i < 4; i++) sum = sum + poly[i][1];
    if (sum * 0.25 > kPiD)
      for (int i = 0; i < 4; i++) poly[i][1] -= kTauD;
  }
  for (int i = 0; i < 4; i++) poly[i][1] += offset;
  double r0 = G.r0, r1 = G.r1;
  bool all_lo = true, all_hi = true, all_alo = true, all_ahi = true;
  for (int i = 0; i < 4; i++) {
    all_lo &= poly[i][0] < r0;
    all_hi &= poly[i][0] > r1;
    all_alo &= poly[i][1] < a0;
    all_ahi &= poly[i][1] > a1;
  }
  if (all_lo || all_hi || all_alo || all_ahi) return 1;
  double dr = G.rstep(), da = G.astep();
  for (int i = 0; i < 4; i++) {
    poly[i][0] = (clampd(poly[i][0], r0, r1) - r0) / dr;
    poly[i][1] = (clampd(poly[i][1], a0, a1) - a0) / da;
  }
  double mr = INFINITY, ma = INFINITY;
  for (int i = 0; i < 4; i++) {
    mr = fmin(mr, poly[i][0]);
    ma = fmin(ma, poly[i][1]);
  }
  size_t ir = (size_t)sat_usize(floor(mr)), ia = (size_t)sat_usize(floor(ma));
  for (int i = 0; i < 4; i++) {
    poly[i][0] -= (double)ir;
    poly[i][1] -= (double)ia;
  }
  double xr = 0.0, xa = 0.0;
  for (int i = 0; i < 4; i++) {
    xr = fmax(xr, poly[i][0]);
    xa = fmax(xa, poly[i][1]);
  }
  size_t nr = (size_t)sat_usize(ceil(xr)), na = (size_t)sat_usize(ceil(xa));
  if (nr == 0 || na == 0) return 0;
  double total = 0.0;
  for (size_t r = 0; r < nr; r++)
    for (size_t q = 0; q < na; q++) {
      double v = 1.0 * (double)pf2d_cell(poly, r, q);
      if (!isfinite(v)) return 0;
      if (v > 0.0 && ir + r < G.nr && ia + q < G.na) total += v;
    }
  if (total <= 0.0) return 0;
  for (size_t r = 0; r < nr; r++)
    for (size_t q = 0; q < na; q++) {
      double v = 1.0 * (double)pf2d_cell(poly, r, q);
      if (v > 0.0 && ir + r < G.nr && ia + q < G.na)
        emit((uint32_t)((ia + q) * G.nr + ir + r), v / total);
    }
  return 1;
}

template <class Emit>
__device__ bool pf2d_pixel(const PfParams& P, uint32_t idx, Emit& emit) {
  uint32_t dim2 = P.g.dim2;
  uint64_t w = (uint64_t)dim2 + 1;
  uint64_t b = (uint64_t)(idx / dim2) * w + idx % dim2;
  uint64_t ci[4] = {b, b + 1, b + w + 1, b + w};
  bool axis_on_edge = false, zero_r = false;
  double winding = 0.0;
  for (int i = 0; i < 4; i++) {
    double delta = P.ac[ci[(i + 1) & 3]] - P.ac[ci[i]];
    delta = rem_euclid(delta + kPiD, kTauD) - kPiD;
    axis_on_edge |= fabs(delta) >= kPiD - 64.0 * kEps;
    winding += delta;
    zero_r |= P.rc[ci[i]] == 0.0;
  }
  if (fabs(winding) > kPiD || axis_on_edge || zero_r) return false;
  uint64_t pi[4] = {b, b + w, b + w + 1, b + 1};
  double pp[4][2];
  for (int i = 0; i < 4; i++) {
    pp[i][0] = P.rc[pi[i]];
    pp[i][1] = P.ac[pi[i]];
  }
  if (pf2d_split(P.G, pp, emit)) return true;
  // Fractional fallback (shared_grid.rs:100-115).
  Ps2Params Q{P.g, P.G, P.rc, P.ac, P.rad, P.chi};
  double center = P.G.a0 + rem_euclid(P.chi[idx] - P.G.a0, kTauD);
  double poly[4][2];
  for (int i = 0; i < 4; i++) {
    poly[i][0] = P.rc[ci[i]];
    poly[i][1] = center + rem_euclid(P.ac[ci[i]] - center + kPiD, kTauD) - kPiD;
  }
  double area = fabs(polygon_area(poly, 4));
  if (area > 0.0) {
    split2d(Q, poly, area, emit);
  } else {
    long long bin = grid_bin(P.G, P.rad[idx], P.chi[idx]);
    if (bin >= 0) emit((uint32_t)bin, 1.0);
  }
  return true;
}

template <bool TWOD>
__global__ void k_pf_count(PfParams P, uint32_t npix, uint32_t* cnt, uint32_t* sing,
                           uint32_t* nsing) {
  uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npix) return;
  CountEmit e;
  bool ok = TWOD ? pf2d_pixel(P, idx, e) : pf1d_pixel(P, idx, e);
  if (!ok) {
    uint32_t s = atomicAdd(nsing, 1u);
    if (s < kMaxSingular) sing[s] = idx;
    cnt[idx] = 0;
    return;
  }
  cnt[idx] = e.n;
}

template <bool TWOD>
__global__ void k_pf_fill(PfParams P, uint32_t npix, const uint32_t* off, uint32_t* bin,
                          uint32_t* pix, double* w) {
  uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npix) return;
  FillEmit e{bin, pix, w, off[idx], idx};
  if (TWOD) pf2d_pixel(P, idx, e);
  else pf1d_pixel(P, idx, e);
}

// --------------------------------------------------------- CSR assembly

__global__ void k_iota(uint32_t* v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

__global__ void k_bin_hist(const uint32_t* bin, uint64_t n, uint32_t* hist) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&hist[bin[i]], 1u);
}

__global__ void k_gather_sorted(const uint32_t* order, const uint32_t* pix, const double* w,
                                uint64_t n, uint32_t* col, double* w64, float* w32) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t k = order[i];
  col[i] = pix[k];
  double v = w[k];
  w64[i] = v;
  w32[i] = (float)v;
}

__global__ void k_keep(const uint32_t* col, const uint8_t* mask, uint64_t n, uint32_t* keep) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keep[i] = mask[col[i]] == 0;
}

__global__ void k_compact(const uint32_t* keep, const uint32_t* pos, uint64_t n,
                          const uint32_t* col, const double* w64, const float* w32,
                          uint32_t* ocol, double* ow64, float* ow32) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !keep[i]) return;
  uint32_t j = pos[i];
  ocol[j] = col[i];
  ow64[j] = w64[i];
  ow32[j] = w32[i];
}

__global__ void k_rows_after_compact(const uint32_t* row_ptr, const uint32_t* pos,
                                     uint32_t nbins, uint32_t* out) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r <= nbins) out[r] = pos[row_ptr[r]];
}

// corrections.rs:57-93, geometry.rs:552-636, 697-717.
__global__ void k_normalization(Geo g, const double* tth, const double* chi, int has_pol,
                                double factor, double* norm) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t npix = (uint64_t)g.dim1 * g.dim2;
  if (i >= npix) return;
  uint32_t r = (uint32_t)(i / g.dim2), c = (uint32_t)(i % g.dim2);
  double p1 = ((double)r + 0.5) * g.pixel1 - g.poni1;
  double p2 = ((double)c + 0.5) * g.pixel2 - g.poni2;
  double p3 = g.distance;
  double cosa = p3 / sqrt(p1 * p1 + p2 * p2 + p3 * p3);
  double sa = cosa * cosa * cosa;
  double pol = 1.0;
  if (has_pol) {
    double ct = cos(tth[i]);
    double c2 = ct * ct;
    pol = 0.5 * (1.0 + c2 - factor * cos(2.0 * chi[i]) * (1.0 - c2));
  }
  norm[i] = pol * sa;
}

// prepared.rs:229-262, row<FAST = false>: n = n + w * norm[j], row order.
// The entries are read 16 at a time (indices and weights, then their
// normalization values) so that a thread waits for memory once per 16
// additions; the additions stay sequential in row order.
__global__ void k_denominators(const uint32_t* row_ptr, const uint32_t* col, const double* w64,
                               int weighted, const double* norm, uint32_t nbins, double* den) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nbins) return;
  constexpr uint32_t B = 16;
  double n = 0.0;
  uint32_t k = row_ptr[r];
  const uint32_t e = row_ptr[r + 1];
  for (; k + B <= e; k += B) {
    uint32_t c[B];
    double w[B], v[B];
#pragma unroll
    for (uint32_t i = 0; i < B; i++) {
      c[i] = col[k + i];
      w[i] = weighted ? w64[k + i] : 1.0;
    }
#pragma unroll
    for (uint32_t i = 0; i < B; i++) v[i] = norm[c[i]];
#pragma unroll
    for (uint32_t i = 0; i < B; i++) n = weighted ? n + w[i] * v[i] : n + v[i];
  }
  for (; k < e; k++) {
    if (weighted) n = n + w64[k] * norm[col[k]];
    else n = n + norm[col[k]];
  }
  den[r] = n;
}

// lut/integrate.rs:52-110 (row_strict, COUNT): count = count + w in row
// order over the execution table (+ 1 per entry for unweighted tables).
__global__ void k_counts(const uint32_t* row_ptr, const double* w64, int weighted,
                         uint32_t nbins, double* cnt) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nbins) return;
  double c = 0.0;
  for (uint32_t k = row_ptr[r]; k < row_ptr[r + 1]; k++) c = c + (weighted ? w64[k] : 1.0);
  cnt[r] = c;
}

// ---------------------------------------------------------------- host

struct Scratch {
  DevBuf<uint8_t> tmp;
  void* get(size_t bytes) {
    if (tmp.n < bytes) tmp.alloc(bytes);
    return tmp.p;
  }
};

void exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, hipStream_t s, Scratch& sc) {
  size_t bytes = 0;
  DIO_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s));
  DIO_HIP(hipcub::DeviceScan::ExclusiveSum(sc.get(bytes), bytes, in, out, (int)n, s));
}

template <class T>
T read1(const T* d, hipStream_t s) {
  T v;
  DIO_HIP(hipMemcpyAsync(&v, d, sizeof(T), hipMemcpyDeviceToHost, s));
  DIO_HIP(hipStreamSynchronize(s));
  return v;
}

// Recompute flagged angles with the host libm (the reference's own atan2),
// in parallel over host threads.
void host_patch(const Geo& g, const std::vector<uint32_t>& idx,
                const std::vector<std::pair<double, double>>& points, std::vector<double>& tv,
                std::vector<double>& cv) {
  size_t n = idx.size();
  tv.resize(n);
  cv.resize(n);
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; t++) {
    pool.emplace_back([&, t]() {
      for (size_t i = t; i < n; i += nt) {
        AngleArgs a = angle_args(g, points[i].first, points[i].second);
        tv[i] = ::atan2(a.ty, a.tx);
        cv[i] = ::atan2(a.cy, a.cx);
      }
    });
  }
  for (auto& th : pool) th.join();
}

// 2theta values in the plan's radial unit, q values that the certified
// sine leaves uncertain recomputed with the host's glibc (tth and out are
// distinct buffers). Returns the number recomputed.
uint64_t unit_array(const Geo& g, const double* tth, double* out, uint64_t n, hipStream_t s) {
  if (n == 0) return 0;
  DevBuf<uint32_t> flist(g.unit >= 2 ? n : 1), fcount(1);
  DIO_HIP(hipMemsetAsync(fcount.p, 0, sizeof(uint32_t), s));
  k_to_unit<<<blocks_for(n, kThreads), kThreads, 0, s>>>(g, tth, out, n, flist.p, fcount.p);
  DIO_HIP(hipGetLastError());
  const uint32_t nf = read1(fcount.p, s);
  if (nf == 0) return 0;
  DevBuf<double> dv(nf);
  k_gather_f64<<<blocks_for(nf, kThreads), kThreads, 0, s>>>(tth, flist.p, nf, dv.p);
  std::vector<uint32_t> idx(nf);
  std::vector<double> v(nf);
  DIO_HIP(hipMemcpyAsync(idx.data(), flist.p, nf * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  DIO_HIP(hipMemcpyAsync(v.data(), dv.p, nf * sizeof(double), hipMemcpyDeviceToHost, s));
  DIO_HIP(hipStreamSynchronize(s));
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (nf < 4096) nt = 1;
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; t++)
    pool.emplace_back([&, t]() {
      for (size_t i = t; i < nf; i += nt) v[i] = unit_host(g, v[i]);
    });
  for (auto& th : pool) th.join();
  DIO_HIP(hipMemcpyAsync(dv.p, v.data(), nf * sizeof(double), hipMemcpyHostToDevice, s));
  k_scatter_f64<<<blocks_for(nf, kThreads), kThreads, 0, s>>>(flist.p, dv.p, nf, out);
  DIO_HIP(hipGetLastError());
  DIO_HIP(hipStreamSynchronize(s));
  return nf;
}

// Angles of a rows x cols point grid at d = index + off, certified and
// patched. Returns the number of patched points.
uint64_t grid_angles(const Geo& g, uint32_t rows, uint32_t cols, double off, double* tth,
                     double* chi, hipStream_t s) {
  uint64_t n = (uint64_t)rows * cols;
  DevBuf<uint32_t> flist(n), fcount(1);
  DIO_HIP(hipMemsetAsync(fcount.p, 0, sizeof(uint32_t), s));
  k_grid_angles<<<blocks_for(n, kThreads), kThreads, 0, s>>>(g, rows, cols, off, tth, chi,
                                                            flist.p, fcount.p);
  DIO_HIP(hipGetLastError());
  uint32_t nf = read1(fcount.p, s);
  if (nf == 0) return 0;
  std::vector<uint32_t> idx(nf);
  DIO_HIP(hipMemcpyAsync(idx.data(), flist.p, nf * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  DIO_HIP(hipStreamSynchronize(s));
  std::vector<std::pair<double, double>> pts(nf);
  for (uint32_t i = 0; i < nf; i++)
    pts[i] = {(double)(idx[i] / cols) + off, (double)(idx[i] % cols) + off};
  std::vector<double> tv, cv;
  host_patch(g, idx, pts, tv, cv);
  DevBuf<uint32_t> di(nf);
  DevBuf<double> dt(nf), dc(nf);
  DIO_HIP(hipMemcpyAsync(di.p, idx.data(), nf * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  DIO_HIP(hipMemcpyAsync(dt.p, tv.data(), nf * sizeof(double), hipMemcpyHostToDevice, s));
  DIO_HIP(hipMemcpyAsync(dc.p, cv.data(), nf * sizeof(double), hipMemcpyHostToDevice, s));
  k_patch<<<blocks_for(nf, kThreads), kThreads, 0, s>>>(di.p, dt.p, dc.p, nf, tth, chi);
  DIO_HIP(hipGetLastError());
  DIO_HIP(hipStreamSynchronize(s));
  return nf;
}

void device_minmax(const double* v, uint64_t n, hipStream_t s, double& lo, double& hi) {
  unsigned nb = std::min<uint64_t>(1024, std::max<uint64_t>(1, blocks_for(n, kThreads)));
  DevBuf<double> pmin(nb), pmax(nb);
  k_minmax<<<nb, kThreads, 0, s>>>(v, n, pmin.p, pmax.p);
  DIO_HIP(hipGetLastError());
  std::vector<double> a(nb), b(nb);
  DIO_HIP(hipMemcpyAsync(a.data(), pmin.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
  DIO_HIP(hipMemcpyAsync(b.data(), pmax.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
  DIO_HIP(hipStreamSynchronize(s));
  lo = a[0];
  hi = b[0];
  for (unsigned i = 1; i < nb; i++) {
    if (a[i] < lo) lo = a[i];
    if (b[i] > hi) hi = b[i];
  }
}

// Pixel-major entries -> CSR with rows in ascending pixel order.
void assemble(uint32_t nbins, uint64_t nnz, DevBuf<uint32_t>& bin, DevBuf<uint32_t>& pix,
              DevBuf<double>& w, hipStream_t s, Scratch& sc, Table& t) {
  t.nbins = nbins;
  t.nnz = nnz;
  t.row_ptr.alloc(nbins + 1);
  t.col.alloc(nnz);
  t.w64.alloc(nnz);
  t.w32.alloc(nnz);
  DevBuf<uint32_t> hist(nbins + 1);
  DIO_HIP(hipMemsetAsync(hist.p, 0, (nbins + 1) * sizeof(uint32_t), s));
  if (nnz) {
    k_bin_hist<<<blocks_for(nnz, kThreads), kThreads, 0, s>>>(bin.p, nnz, hist.p);
    DIO_HIP(hipGetLastError());
  }
  exclusive_scan(hist.p, t.row_ptr.p, nbins + 1, s, sc);
  if (!nnz) return;
  DevBuf<uint32_t> order_in(nnz), order(nnz), keys(nnz);
  k_iota<<<blocks_for(nnz, kThreads), kThreads, 0, s>>>(order_in.p, nnz);
  int end_bit = 1;
  while (end_bit < 32 && (1ull << end_bit) < nbins) end_bit++;
  size_t bytes = 0;
  DIO_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, bin.p, keys.p, order_in.p, order.p,
                                             (int)nnz, 0, end_bit, s));
  DIO_HIP(hipcub::DeviceRadixSort::SortPairs(sc.get(bytes), bytes, bin.p, keys.p, order_in.p,
                                             order.p, (int)nnz, 0, end_bit, s));
  k_gather_sorted<<<blocks_for(nnz, kThreads), kThreads, 0, s>>>(order.p, pix.p, w.p, nnz,
                                                                t.col.p, t.w64.p, t.w32.p);
  DIO_HIP(hipGetLastError());
}

void compact(const Table& in, const uint8_t* dmask, hipStream_t s, Scratch& sc, Table& out) {
  out.nbins = in.nbins;
  DevBuf<uint32_t> keep(in.nnz + 1), pos(in.nnz + 1);
  DIO_HIP(hipMemsetAsync(keep.p, 0, (in.nnz + 1) * sizeof(uint32_t), s));
  if (in.nnz)
    k_keep<<<blocks_for(in.nnz, kThreads), kThreads, 0, s>>>(in.col.p, dmask, in.nnz, keep.p);
  exclusive_scan(keep.p, pos.p, in.nnz + 1, s, sc);
  uint32_t m = read1(pos.p + in.nnz, s);
  out.nnz = m;
  out.row_ptr.alloc(in.nbins + 1);
  out.col.alloc(m);
  out.w64.alloc(m);
  out.w32.alloc(m);
  k_rows_after_compact<<<blocks_for(in.nbins + 1, kThreads), kThreads, 0, s>>>(
      in.row_ptr.p, pos.p, in.nbins, out.row_ptr.p);
  if (in.nnz)
    k_compact<<<blocks_for(in.nnz, kThreads), kThreads, 0, s>>>(
        keep.p, pos.p, in.nnz, in.col.p, in.w64.p, in.w32.p, out.col.p, out.w64.p, out.w32.p);
  DIO_HIP(hipGetLastError());
  DIO_HIP(hipStreamSynchronize(s));
}

// IntegrationGrid2d::new validity for the 1D sampling grid (grid.rs:48-108).
bool sample_grid_ok(uint32_t n, double lo, double hi) {
  double bounds[2][2] = {{lo, hi}, {-kPiD, kPiD}};
  uint32_t counts[2] = {n, 1};
  if (n == 0) return false;
  for (int k = 0; k < 2; k++) {
    double l = bounds[k][0], h = bounds[k][1];
    double step = (h - l) / (double)counts[k];
    if (!std::isfinite(l) || !std::isfinite(h) || !std::isfinite(step) || step <= 0.0 ||
        l + step == l || h - step == h)
      return false;
  }
  return lo >= 0.0;
}

// Beam pixels: sample, patch uncertain sample angles, bin, count, fill.
struct Singular {
  std::vector<uint32_t> pixels;
  DevBuf<uint32_t> dpix, hits;
  uint64_t patched = 0;
};

void singular_prepare(const Geo& g, const SampleGrid& G, const Grid2* G2, Singular& S,
                      uint32_t* cnt, hipStream_t s) {
  uint32_t nbins = G2 ? G2->nr * G2->na : G.n;
  uint32_t ns = (uint32_t)S.pixels.size();
  if (!ns) return;
  std::sort(S.pixels.begin(), S.pixels.end());
  S.dpix.alloc(ns);
  DIO_HIP(hipMemcpyAsync(S.dpix.p, S.pixels.data(), ns * sizeof(uint32_t), hipMemcpyHostToDevice,
                         s));
  uint64_t total = (uint64_t)ns * kSamplesPerPixel;
  DevBuf<double> st(total), sc(total);
  DevBuf<uint32_t> flist(total), fcount(1);
  DIO_HIP(hipMemsetAsync(fcount.p, 0, sizeof(uint32_t), s));
  if (G2)
    k_sing_samples2d<<<blocks_for(total, kThreads), kThreads, 0, s>>>(g, S.dpix.p, ns, st.p,
                                                                     sc.p, flist.p, fcount.p);
  else
    k_sing_samples<<<blocks_for(total, kThreads), kThreads, 0, s>>>(g, S.dpix.p, ns, st.p, sc.p,
                                                                   flist.p, fcount.p);
  DIO_HIP(hipGetLastError());
  uint32_t nf = read1(fcount.p, s);
  if (nf) {
    std::vector<uint32_t> idx(nf);
    DIO_HIP(hipMemcpyAsync(idx.data(), flist.p, nf * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    DIO_HIP(hipStreamSynchronize(s));
    std::vector<std::pair<double, double>> pts(nf);
    const double dy[4] = {0.25, 0.5, 0.75, 0.5};
    const double dx[4] = {0.5, 0.25, 0.5, 0.75};
    for (uint32_t i = 0; i < nf; i++) {
      uint32_t slot = idx[i] / kSamplesPerPixel, smp = idx[i] % kSamplesPerPixel;
      uint32_t pixel = S.pixels[slot];
      if (G2) {
        sample_point_2d(g.dim2, pixel, smp, pts[i].first, pts[i].second);
        continue;
      }
      uint32_t row = smp >> 9, col = (smp >> 2) & 127u, o = smp & 3u;
      pts[i] = {(double)(pixel / g.dim2) + ((double)row + dy[o]) / 128.0,
                (double)(pixel % g.dim2) + ((double)col + dx[o]) / 128.0};
    }
    std::vector<double> tv, cv;
    host_patch(g, idx, pts, tv, cv);
    DevBuf<uint32_t> di(nf);
    DevBuf<double> dt(nf), dc(nf);
    DIO_HIP(hipMemcpyAsync(di.p, idx.data(), nf * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    DIO_HIP(hipMemcpyAsync(dt.p, tv.data(), nf * sizeof(double), hipMemcpyHostToDevice, s));
    DIO_HIP(hipMemcpyAsync(dc.p, cv.data(), nf * sizeof(double), hipMemcpyHostToDevice, s));
    k_patch<<<blocks_for(nf, kThreads), kThreads, 0, s>>>(di.p, dt.p, dc.p, nf, st.p, sc.p);
    DIO_HIP(hipGetLastError());
    S.patched = nf;
  }
  // Sample radial values in the plan's unit.
  const double* srad = st.p;
  DevBuf<double> sr;
  if (g.unit != 0) {
    sr.alloc(total);
    S.patched += unit_array(g, st.p, sr.p, total, s);
    srad = sr.p;
  }
  S.hits.alloc((uint64_t)ns * nbins);
  DIO_HIP(hipMemsetAsync(S.hits.p, 0, S.hits.bytes(), s));
  if (G2)
    k_sing_bin2d<<<blocks_for(total, kThreads), kThreads, 0, s>>>(*G2, ns, srad, sc.p,
                                                                 S.hits.p);
  else
    k_sing_bin<<<blocks_for(total, kThreads), kThreads, 0, s>>>(G, ns, srad, sc.p, S.hits.p);
  k_sing_count<<<ns, kThreads, 0, s>>>(S.hits.p, nbins, ns, S.dpix.p, cnt);
  DIO_HIP(hipGetLastError());
  DIO_HIP(hipStreamSynchronize(s));
}

void build_1d(const Geo& g, const dio_plan_desc& d, hipStream_t s, Scratch& sc, Built& out) {
  uint32_t dim1 = d.dim1, dim2 = d.dim2;
  uint64_t npix = (uint64_t)dim1 * dim2;
  uint32_t n = d.n_radial;
  // Pixel-centre angles (calc_tth_all / calc_azimuth_all).
  DevBuf<double> ttp(npix), chp(npix), radp;
  out.uncertain += grid_angles(g, dim1, dim2, 0.5, ttp.p, chp.p, s);
  const double* rad = ttp.p;
  if (g.unit != 0) {
    radp.alloc(npix);
    out.uncertain += unit_array(g, ttp.p, radp.p, npix, s);
    rad = radp.p;
  }
  DevBuf<uint32_t> cnt(npix + 1), off(npix + 1);
  DIO_HIP(hipMemsetAsync(cnt.p, 0, (npix + 1) * sizeof(uint32_t), s));
  DevBuf<uint32_t> ebin, epix;
  DevBuf<double> ew;
  uint64_t nnz = 0;
  if (d.method == DIO_COUNTING) {
    double min, max, step;
    if (d.has_radial_range) {
      min = d.radial_min;
      max = d.radial_max + 10.0 * kEps;
    } else {
      device_minmax(rad, npix, s, min, max);
      max = max + 10.0 * kEps;
    }
    step = (max - min) / (double)n;
    DevBuf<uint32_t> pbin(npix);
    k_count_bins<<<blocks_for(npix, kThreads), kThreads, 0, s>>>(rad, (uint32_t)npix, n, min, max,
                                                                step, d.has_radial_range, cnt.p,
                                                                pbin.p);
    exclusive_scan(cnt.p, off.p, npix + 1, s, sc);
    nnz = read1(off.p + npix, s);
    ebin.alloc(nnz);
    epix.alloc(nnz);
    ew.alloc(nnz);
    k_count_fill<<<blocks_for(npix, kThreads), kThreads, 0, s>>>(cnt.p, pbin.p, off.p,
                                                                (uint32_t)npix, ebin.p, epix.p,
                                                                ew.p);
    DIO_HIP(hipGetLastError());
    out.weighted = false;
    out.hist_min = min;
    out.hist_step = step;
  } else if (d.method == DIO_PIXEL_SPLIT) {
    uint64_t nc = (uint64_t)(dim1 + 1) * (dim2 + 1);
    DevBuf<double> tc(nc), ac(nc), rcb;
    out.uncertain += grid_angles(g, dim1 + 1, dim2 + 1, 0.0, tc.p, ac.p, s);
    const double* rc = tc.p;
    if (g.unit != 0) {
      rcb.alloc(nc);
      out.uncertain += unit_array(g, tc.p, rcb.p, nc, s);
      rc = rcb.p;
    }
    double min, max;
    if (d.has_radial_range) {
      min = d.radial_min;
      max = d.radial_max + 10.0 * kEps;
    } else {
      device_minmax(rc, nc, s, min, max);
      max += 10.0 * kEps;
    }
    double step = (max - min) / (double)n;
    PsParams P;
    P.g = g;
    P.rc = rc;
    P.ac = ac.p;
    P.rad = rad;
    P.chi = chp.p;
    P.n = n;
    P.min = min;
    P.step = step;
    P.glo = min;
    P.ghi = min + (double)n * step;
    P.grid_ok = sample_grid_ok(n, P.glo, P.ghi) ? 1 : 0;
    P.clamp_low = d.has_radial_range ? 0 : 1;
    DevBuf<uint32_t> sing(kMaxSingular), nsing(1);
    DIO_HIP(hipMemsetAsync(nsing.p, 0, sizeof(uint32_t), s));
    k_ps_count<<<blocks_for(npix, kThreads), kThreads, 0, s>>>(P, (uint32_t)npix, cnt.p, sing.p,
                                                              nsing.p);
    DIO_HIP(hipGetLastError());
    uint32_t ns = read1(nsing.p, s);
    if (ns > kMaxSingular) fail(DIO_ERR_UNSUPPORTED, "too many beam pixels for sampling");
    Singular S;
    S.pixels.resize(ns);
    if (ns) {
      DIO_HIP(hipMemcpyAsync(S.pixels.data(), sing.p, ns * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, s));
      DIO_HIP(hipStreamSynchronize(s));
    }
    SampleGrid G{n, P.glo, P.ghi, P.clamp_low};
    singular_prepare(g, G, nullptr, S, cnt.p, s);
    out.uncertain += S.patched;
    out.singular = ns;
    exclusive_scan(cnt.p, off.p, npix + 1, s, sc);
    uint64_t total = read1(off.p + npix, s);
    nnz = total;
    ebin.alloc(nnz);
    epix.alloc(nnz);
    ew.alloc(nnz);
    k_ps_fill<<<blocks_for(npix, kThreads), kThreads, 0, s>>>(P, (uint32_t)npix, off.p, ebin.p,
                                                             epix.p, ew.p);
    if (ns)
      k_sing_fill<<<blocks_for(ns, 64), 64, 0, s>>>(S.hits.p, n, ns, S.dpix.p, off.p, ebin.p,
                                                    epix.p, ew.p);
    DIO_HIP(hipGetLastError());
    DIO_HIP(hipStreamSynchronize(s));
    out.hist_min = min;
    out.hist_step = step;
  } else {
    fail(DIO_ERR_UNSUPPORTED, "method not available for 1D tables on this device build");
  }
  if (nnz > 0x7fffffffull) fail(DIO_ERR_UNSUPPORTED, "table exceeds 2^31 entries");
  out.n_radial = n;
  out.n_azimuthal = 1;
  assemble(n, nnz, ebin, epix, ew, s, sc, out.full);
  // Corrections use the (patched) pixel-centre angles.
  out.norm.alloc(npix);
  k_normalization<<<blocks_for(npix, kThreads), kThreads, 0, s>>>(
      g, ttp.p, chp.p, d.has_polarization, d.polarization_factor, out.norm.p);
  DIO_HIP(hipGetLastError());
  DIO_HIP(hipStreamSynchronize(s));
}

// Angles of an explicit point list, certified and patched.
uint64_t list_angles(const Geo& g, const std::vector<double>& d1, const std::vector<double>& d2,
                     DevBuf<double>& tth, DevBuf<double>& chi, hipStream_t s) {
  uint32_t n = (uint32_t)d1.size();
  tth.alloc(n);
  chi.alloc(n);
  if (!n) return 0;
  DevBuf<double> a(n), b(n);
  DevBuf<uint32_t> flist(n), fcount(1);
  DIO_HIP(hipMemcpyAsync(a.p, d1.data(), n * 8, hipMemcpyHostToDevice, s));
  DIO_HIP(hipMemcpyAsync(b.p, d2.data(), n * 8, hipMemcpyHostToDevice, s));
  DIO_HIP(hipMemsetAsync(fcount.p, 0, 4, s));
  k_list_angles<<<blocks_for(n, kThreads), kThreads, 0, s>>>(g, a.p, b.p, n, tth.p, chi.p,
                                                            flist.p, fcount.p);
  DIO_HIP(hipGetLastError());
  uint32_t nf = read1(fcount.p, s);
  if (!nf) return 0;
  std::vector<uint32_t> idx(nf);
  DIO_HIP(hipMemcpyAsync(idx.data(), flist.p, nf * 4, hipMemcpyDeviceToHost, s));
  DIO_HIP(hipStreamSynchronize(s));
  std::vector<std::pair<double, double>> pts(nf);
  for (uint32_t i = 0; i < nf; i++) pts[i] = {d1[idx[i]], d2[idx[i]]};
  std::vector<double> tv, cv;
  host_patch(g, idx, pts, tv, cv);
  DevBuf<uint32_t> di(nf);
  DevBuf<double> dt(nf), dc(nf);
  DIO_HIP(hipMemcpyAsync(di.p, idx.data(), nf * 4, hipMemcpyHostToDevice, s));
  DIO_HIP(hipMemcpyAsync(dt.p, tv.data(), nf * 8, hipMemcpyHostToDevice, s));
  DIO_HIP(hipMemcpyAsync(dc.p, cv.data(), nf * 8, hipMemcpyHostToDevice, s));
  k_patch<<<blocks_for(nf, kThreads), kThreads, 0, s>>>(di.p, dt.p, dc.p, nf, tth.p, chi.p);
  DIO_HIP(hipGetLastError());
  DIO_HIP(hipStreamSynchronize(s));
  return nf;
}

// histogram.rs:151-181: range over pixel centres and the hull grid
// (lut/common.rs:33-65) plus 3 eps, for radial (unit) and azimuth.
void ss_ranges(const Geo& g, uint32_t nss, const double* rad, const double* chi, uint64_t npix,
               hipStream_t s, double& rlo, double& rhi, double& alo, double& ahi,
               uint64_t& patched) {
  uint64_t dim1 = g.dim1, dim2 = g.dim2;
  double border = 0.5 / (double)nss, step = 1.0 / (double)nss;
  std::vector<double> h1, h2;
  for (uint64_t i = 0; i < dim1 * nss; i++) { h1.push_back(border + (double)i * step); h2.push_back(border); }
  for (uint64_t i = 0; i < dim1 * nss; i++) {
    h1.push_back(border + (double)i * step);
    h2.push_back((double)dim2 - border);
  }
  for (uint64_t i = 1; i < dim2 * nss; i++) { h1.push_back(border); h2.push_back(border + (double)i * step); }
  for (uint64_t i = 1; i < dim2 * nss; i++) {
    h1.push_back((double)dim1 - border);
    h2.push_back(border + (double)i * step);
  }
  DevBuf<double> ht, ha, hr;
  patched += list_angles(g, h1, h2, ht, ha, s);
  const double* hrad = ht.p;
  if (g.unit != 0) {
    hr.alloc(h1.size());
    patched += unit_array(g, ht.p, hr.p, h1.size(), s);
    hrad = hr.p;
  }
  double a, b, c, d;
  device_minmax(rad, npix, s, a, b);
  device_minmax(hrad, h1.size(), s, c, d);
  rlo = std::fmin(a, c);
  rhi = std::fmax(b, d) + 3.0 * kEps;
  if (chi) {
    device_minmax(chi, npix, s, a, b);
    device_minmax(ha.p, h1.size(), s, c, d);
    alo = std::fmin(a, c);
    ahi = std::fmax(b, d) + 3.0 * kEps;
  }
}

// Supersampled entries: per-subpixel bins, host-resolved ambiguous ones,
// then per-pixel distinct bins with weight count / n_ss^2.
uint64_t ss_entries(const SsBin& B, uint64_t npix, hipStream_t s, Scratch& sc,
                    DevBuf<uint32_t>& cnt, DevBuf<uint32_t>& off, DevBuf<uint32_t>& ebin,
                    DevBuf<uint32_t>& epix, DevBuf<double>& ew, uint64_t& patched) {
  uint32_t nss2 = B.nss * B.nss;
  uint64_t nsub = npix * nss2;
  if (nsub > 0xfffffff0ull) fail(DIO_ERR_UNSUPPORTED, "too many subpixels for 32-bit indices");
  DevBuf<uint32_t> bins(nsub), flist(1u << 20), fcount(1);
  DIO_HIP(hipMemsetAsync(fcount.p, 0, 4, s));
  k_ss_bins<<<blocks_for(nsub, kThreads), kThreads, 0, s>>>(B, nsub, bins.p, flist.p, fcount.p,
                                                           1u << 20);
  DIO_HIP(hipGetLastError());
  uint32_t nf = read1(fcount.p, s);
  if (nf > (1u << 20)) fail(DIO_ERR_DEVICE, "too many ambiguous subpixels");
  if (nf) {
    std::vector<uint32_t> idx(nf), val(nf);
    DIO_HIP(hipMemcpyAsync(idx.data(), flist.p, nf * 4, hipMemcpyDeviceToHost, s));
    DIO_HIP(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < nf; i++) {
      double d1, d2;
      ss_point(B, idx[i], d1, d2);
      AngleArgs a = angle_args(B.g, d1, d2);
      long long b = ss_bin(B, unit_host(B.g, ::atan2(a.ty, a.tx)), ::atan2(a.cy, a.cx));
      val[i] = b < 0 ? 0xffffffffu : (uint32_t)b;
    }
    DevBuf<uint32_t> di(nf), dv(nf);
    DIO_HIP(hipMemcpyAsync(di.p, idx.data(), nf * 4, hipMemcpyHostToDevice, s));
    DIO_HIP(hipMemcpyAsync(dv.p, val.data(), nf * 4, hipMemcpyHostToDevice, s));
    k_ss_patch<<<blocks_for(nf, kThreads), kThreads, 0, s>>>(di.p, dv.p, nf, bins.p);
    DIO_HIP(hipStreamSynchronize(s));
    patched += nf;
  }
  k_ss_count<<<blocks_for(npix, kThreads), kThreads, 0, s>>>(bins.p, (uint32_t)npix, nss2,
                                                            cnt.p);
  exclusive_scan(cnt.p, off.p, npix + 1, s, sc);
  uint64_t nnz = read1(off.p + npix, s);
  ebin.alloc(nnz);
  epix.alloc(nnz);
  ew.alloc(nnz);
  k_ss_fill<<<blocks_for(npix, kThreads), kThreads, 0, s>>>(bins.p, (uint32_t)npix, nss2, off.p,
                                                           ebin.p, epix.p, ew.p);
  DIO_HIP(hipGetLastError());
  DIO_HIP(hipStreamSynchronize(s));
  return nnz;
}

// integration/grid.rs:48-108 (IntegrationGrid2d::new): validity and the
// correction of azimuthal starts outside [-2pi, 2pi].
Grid2 grid_new(uint32_t nr, uint32_t na, double rlo, double rhi, double alo, double ahi) {
  double lo[2] = {rlo, alo}, hi[2] = {rhi, ahi};
  uint32_t cnt[2] = {nr, na};
  const char* names[2] = {"radial", "azimuthal"};
  for (int k = 0; k < 2; k++) {
    double step = (hi[k] - lo[k]) / (double)cnt[k];
    if (!std::isfinite(lo[k]) || !std::isfinite(hi[k]) || !std::isfinite(step) ||
        step <= 0.0 || lo[k] + step == lo[k] || hi[k] - step == hi[k])
      fail(DIO_ERR_INVALID, std::string(names[k]) +
                                " range must have finite, increasing, resolvable bin edges");
  }
  if (rlo < 0.0) fail(DIO_ERR_INVALID, "radial range must be nonnegative");
  if (ahi - alo > kTauD + 1e-12) fail(DIO_ERR_INVALID, "azimuthal span must be at most 2pi");
  if (!(alo >= -kTauD && alo <= kTauD)) {
    double start = rem_euclid(alo, kTauD);
    if (start >= kPiD) start -= kTauD;
    ahi = start + (ahi - alo);
    alo = start;
  }
  Grid2 g{};
  g.nr = nr;
  g.na = na;
  g.r0 = rlo;
  g.r1 = rhi;
  g.a0 = alo;
  g.a1 = ahi;
  return g;
}

// Default and explicit-range 2D tables of Counting and PixelSplit
// (integration/methods.rs:365-388, grid.rs:579-675, 978-1028,
// lut/shared_grid.rs:13-150).
void build_2d(const Geo& g, const dio_plan_desc& d, hipStream_t s, Scratch& sc, Built& out) {
  uint32_t dim1 = d.dim1, dim2 = d.dim2;
  uint64_t npix = (uint64_t)dim1 * dim2;
  uint32_t nr = d.n_radial, na = d.n_azimuthal;
  uint64_t nb64 = (uint64_t)nr * na;
  DevBuf<double> ttp(npix), chp(npix), radp;
  out.uncertain += grid_angles(g, dim1, dim2, 0.5, ttp.p, chp.p, s);
  const double* rad = ttp.p;
  if (g.unit != 0) {
    radp.alloc(npix);
    out.uncertain += unit_array(g, ttp.p, radp.p, npix, s);
    rad = radp.p;
  }
  bool ps = d.method == DIO_PIXEL_SPLIT;
  if (!ps && d.method != DIO_COUNTING)
    fail(DIO_ERR_UNSUPPORTED, "method not available for 2D tables on this device build");
  uint64_t nc = (uint64_t)(dim1 + 1) * (dim2 + 1);
  DevBuf<double> tc, ac, rcb;
  const double* rc = nullptr;
  if (ps) {
    tc.alloc(nc);
    ac.alloc(nc);
    out.uncertain += grid_angles(g, dim1 + 1, dim2 + 1, 0.0, tc.p, ac.p, s);
    rc = tc.p;
    if (g.unit != 0) {
      rcb.alloc(nc);
      out.uncertain += unit_array(g, tc.p, rcb.p, nc, s);
      rc = rcb.p;
    }
  }
  double rlo, rhi, alo, ahi;
  if (d.has_radial_range) {
    rlo = d.radial_min;
    rhi = d.radial_max;
  } else {
    device_minmax(ps ? rc : rad, ps ? nc : npix, s, rlo, rhi);
    rhi = rhi + 10.0 * kEps;
  }
  if (d.has_azimuthal_range) {
    alo = d.azimuthal_min;
    ahi = d.azimuthal_max;
  } else {
    device_minmax(ps ? ac.p : chp.p, ps ? nc : npix, s, alo, ahi);
    ahi = ahi + 10.0 * kEps;
    if (!(ahi - alo > 1e-12)) {
      alo = -kPiD;
      ahi = kPiD;
    }
  }
  Grid2 G = grid_new(nr, na, rlo, rhi, alo, ahi);
  G.clamp_index_rounding = 1;
  if (!d.has_radial_range && !ps) G.clamp_radial_end = 1;
  DevBuf<uint32_t> cnt(npix + 1), off(npix + 1);
  DIO_HIP(hipMemsetAsync(cnt.p, 0, (npix + 1) * sizeof(uint32_t), s));
  DevBuf<uint32_t> ebin, epix;
  DevBuf<double> ew;
  uint64_t nnz = 0;
  if (!ps) {
    DevBuf<uint32_t> pbin(npix);
    k_count2d_bins<<<blocks_for(npix, kThreads), kThreads, 0, s>>>(G, rad, chp.p, (uint32_t)npix,
                                                                  cnt.p, pbin.p);
    exclusive_scan(cnt.p, off.p, npix + 1, s, sc);
    nnz = read1(off.p + npix, s);
    ebin.alloc(nnz);
    epix.alloc(nnz);
    ew.alloc(nnz);
    k_count_fill<<<blocks_for(npix, kThreads), kThreads, 0, s>>>(cnt.p, pbin.p, off.p,
                                                                (uint32_t)npix, ebin.p, epix.p,
                                                                ew.p);
    DIO_HIP(hipGetLastError());
    out.weighted = false;
  } else {
    Ps2Params P{g, G, rc, ac.p, rad, chp.p};
    DevBuf<uint32_t> sing(kMaxSingular), nsing(1);
    DIO_HIP(hipMemsetAsync(nsing.p, 0, sizeof(uint32_t), s));
    k_ps2d_count<<<blocks_for(npix, kThreads), kThreads, 0, s>>>(P, (uint32_t)npix, cnt.p,
                                                                sing.p, nsing.p);
    DIO_HIP(hipGetLastError());
    uint32_t ns = read1(nsing.p, s);
    if (ns > kMaxSingular) fail(DIO_ERR_UNSUPPORTED, "too many beam pixels for sampling");
    Singular S;
    S.pixels.resize(ns);
    if (ns) {
      DIO_HIP(hipMemcpyAsync(S.pixels.data(), sing.p, ns * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, s));
      DIO_HIP(hipStreamSynchronize(s));
    }
    SampleGrid unused{0, 0.0, 0.0, 0};
    singular_prepare(g, unused, &G, S, cnt.p, s);
    out.uncertain += S.patched;
    out.singular = ns;
    exclusive_scan(cnt.p, off.p, npix + 1, s, sc);
    nnz = read1(off.p + npix, s);
    ebin.alloc(nnz);
    epix.alloc(nnz);
    ew.alloc(nnz);
    k_ps2d_fill<<<blocks_for(npix, kThreads), kThreads, 0, s>>>(P, (uint32_t)npix, off.p, ebin.p,
                                                               epix.p, ew.p);
    if (ns)
      k_sing_fill<<<blocks_for(ns, 64), 64, 0, s>>>(S.hits.p, (uint32_t)nb64, ns, S.dpix.p,
                                                    off.p, ebin.p, epix.p, ew.p);
    DIO_HIP(hipGetLastError());
    DIO_HIP(hipStreamSynchronize(s));
  }
  if (nnz > 0x7fffffffull) fail(DIO_ERR_UNSUPPORTED, "table exceeds 2^31 entries");
  out.n_radial = nr;
  out.n_azimuthal = na;
  out.hist_min = G.r0;
  out.hist_step = G.rstep();
  out.az_min = G.a0;
  out.az_step = G.astep();
  assemble((uint32_t)nb64, nnz, ebin, epix, ew, s, sc, out.full);
  out.norm.alloc(npix);
  k_normalization<<<blocks_for(npix, kThreads), kThreads, 0, s>>>(
      g, ttp.p, chp.p, d.has_polarization, d.polarization_factor, out.norm.p);
  DIO_HIP(hipGetLastError());
  DIO_HIP(hipStreamSynchronize(s));
}

// Supersampled tables, 1D (supersampled.rs:517-595) and the default or
// explicit-range 2D grid (methods.rs:365-388, grid.rs:579-675, 978-1028,
// supersampled.rs:1437-1477).
void build_ss(const Geo& g, const dio_plan_desc& d, hipStream_t s, Scratch& sc, Built& out) {
  uint32_t dim1 = d.dim1, dim2 = d.dim2;
  uint64_t npix = (uint64_t)dim1 * dim2;
  DevBuf<double> ttp(npix), chp(npix), radp;
  out.uncertain += grid_angles(g, dim1, dim2, 0.5, ttp.p, chp.p, s);
  const double* rad = ttp.p;
  if (g.unit != 0) {
    radp.alloc(npix);
    out.uncertain += unit_array(g, ttp.p, radp.p, npix, s);
    rad = radp.p;
  }
  SsBin B{};
  B.g = g;
  B.nss = d.n_ss;
  B.inv = 1;
  uint32_t nbins;
  if (d.ndim == 1) {
    double min, max, dummy;
    if (d.has_radial_range) {
      min = d.radial_min;
      max = d.radial_max + 10.0 * kEps;
    } else {
      ss_ranges(g, d.n_ss, rad, nullptr, npix, s, min, max, dummy, dummy, out.uncertain);
    }
    B.twod = 0;
    B.n1 = d.n_radial;
    B.min1 = min;
    B.max1 = max;
    B.step1 = (max - min) / (double)d.n_radial;
    B.clamped1 = d.has_radial_range ? 0 : 1;
    nbins = d.n_radial;
    out.n_radial = d.n_radial;
    out.n_azimuthal = 1;
    out.hist_min = min;
    out.hist_step = B.step1;
  } else {
    double rlo = d.radial_min, rhi = d.radial_max, alo = d.azimuthal_min, ahi = d.azimuthal_max;
    if (!d.has_radial_range || !d.has_azimuthal_range) {
      double a, b, c, e;
      ss_ranges(g, d.n_ss, rad, chp.p, npix, s, a, b, c, e, out.uncertain);
      if (!d.has_radial_range) {
        rlo = a;
        rhi = b;
      }
      if (!d.has_azimuthal_range) {
        alo = c;
        ahi = e;
        if (!(ahi - alo > 1e-12)) {
          alo = -kPiD;
          ahi = kPiD;
        }
      }
    }
    Grid2 G = grid_new(d.n_radial, d.n_azimuthal, rlo, rhi, alo, ahi);
    G.clamp_index_rounding = 1;
    if (!d.has_radial_range) G.clamp_radial_outside = 1;
    if (!d.has_azimuthal_range) G.clamp_azimuthal_outside = 1;
    B.twod = 1;
    B.G = G;
    // The fast builder (and its * inv_ss coordinates) serves only the default
    // clamped grid (supersampled.rs:1448-1452, 750-753); the others go through
    // build_counted_csr (/ n_ss).
    B.inv = (G.clamp_radial_outside && G.clamp_azimuthal_outside && G.astep() > 1e-9) ? 1 : 0;
    nbins = d.n_radial * d.n_azimuthal;
    out.n_radial = d.n_radial;
    out.n_azimuthal = d.n_azimuthal;
    out.hist_min = G.r0;
    out.hist_step = G.rstep();
    out.az_min = G.a0;
    out.az_step = G.astep();
  }
  DevBuf<uint32_t> cnt(npix + 1), off(npix + 1), ebin, epix;
  DIO_HIP(hipMemsetAsync(cnt.p, 0, (npix + 1) * sizeof(uint32_t), s));
  DevBuf<double> ew;
  uint64_t nnz = ss_entries(B, npix, s, sc, cnt, off, ebin, epix, ew, out.uncertain);
  if (nnz > 0x7fffffffull) fail(DIO_ERR_UNSUPPORTED, "table exceeds 2^31 entries");
  assemble(nbins, nnz, ebin, epix, ew, s, sc, out.full);
  out.norm.alloc(npix);
  k_normalization<<<blocks_for(npix, kThreads), kThreads, 0, s>>>(
      g, ttp.p, chp.p, d.has_polarization, d.polarization_factor, out.norm.p);
  DIO_HIP(hipGetLastError());
  DIO_HIP(hipStreamSynchronize(s));
}

// pyFAI-compatible tables: 1D (pyfai_1d.rs:55-109) and the 2D shared grid
// with PixelSplitMode::PyfaiHistogram (methods.rs:371-373, grid.rs:579-622).
void build_pyfai(const Geo& g, const dio_plan_desc& d, hipStream_t s, Scratch& sc, Built& out) {
  uint32_t dim1 = d.dim1, dim2 = d.dim2;
  uint64_t npix = (uint64_t)dim1 * dim2;
  uint64_t nc = (uint64_t)(dim1 + 1) * (dim2 + 1);
  DevBuf<double> ttp(npix), chp(npix), radp, tc(nc), ac(nc), rcb;
  out.uncertain += grid_angles(g, dim1, dim2, 0.5, ttp.p, chp.p, s);
  out.uncertain += grid_angles(g, dim1 + 1, dim2 + 1, 0.0, tc.p, ac.p, s);
  const double* rad = ttp.p;
  const double* rc = tc.p;
  if (g.unit != 0) {
    radp.alloc(npix);
    rcb.alloc(nc);
    out.uncertain += unit_array(g, ttp.p, radp.p, npix, s);
    out.uncertain += unit_array(g, tc.p, rcb.p, nc, s);
    rad = radp.p;
    rc = rcb.p;
  }
  double r0, r1;
  if (d.has_radial_range) {
    r0 = d.radial_min;
    r1 = d.radial_max;
  } else {
    // pyfai_1d.rs:36-46: [max(min corner, 0), max corner * (1 + f32::EPSILON)].
    double lo, hi;
    device_minmax(rc, nc, s, lo, hi);
    r0 = std::fmax(lo, 0.0);
    r1 = hi * (1.0 + 1.1920928955078125e-07);
  }
  PfParams P{};
  P.g = g;
  P.rc = rc;
  P.ac = ac.p;
  P.rad = rad;
  P.chi = chp.p;
  bool twod = d.ndim == 2;
  uint32_t nbins;
  SampleGrid SG{0, 0.0, 0.0, 0};
  Grid2 G{};
  if (twod) {
    G = grid_new(d.n_radial, d.n_azimuthal, r0, r1,
                 d.has_azimuthal_range ? d.azimuthal_min : -kPiD,
                 d.has_azimuthal_range ? d.azimuthal_max : kPiD);
    G.clamp_index_rounding = 1;
    P.G = G;
    nbins = d.n_radial * d.n_azimuthal;
    out.n_radial = d.n_radial;
    out.n_azimuthal = d.n_azimuthal;
    out.hist_min = G.r0;
    out.hist_step = G.rstep();
    out.az_min = G.a0;
    out.az_step = G.astep();
  } else {
    if (!sample_grid_ok(d.n_radial, r0, r1))
      fail(DIO_ERR_INVALID, "radial range must have finite, increasing, resolvable bin edges");
    P.n = d.n_radial;
    P.r0 = r0;
    P.r1 = r1;
    SG = SampleGrid{d.n_radial, r0, r1, d.has_radial_range ? 0 : 1};
    nbins = d.n_radial;
    out.n_radial = d.n_radial;
    out.n_azimuthal = 1;
    out.hist_min = r0;
    out.hist_step = (r1 - r0) / (double)d.n_radial;
  }
  DevBuf<uint32_t> cnt(npix + 1), off(npix + 1), sing(kMaxSingular), nsing(1), ebin, epix;
  DevBuf<double> ew;
  DIO_HIP(hipMemsetAsync(cnt.p, 0, (npix + 1) * sizeof(uint32_t), s));
  DIO_HIP(hipMemsetAsync(nsing.p, 0, sizeof(uint32_t), s));
  if (twod)
    k_pf_count<true><<<blocks_for(npix, kThreads), kThreads, 0, s>>>(P, (uint32_t)npix, cnt.p,
                                                                    sing.p, nsing.p);
  else
    k_pf_count<false><<<blocks_for(npix, kThreads), kThreads, 0, s>>>(P, (uint32_t)npix, cnt.p,
                                                                     sing.p, nsing.p);
  DIO_HIP(hipGetLastError());
  uint32_t ns = read1(nsing.p, s);
  if (ns > kMaxSingular) fail(DIO_ERR_UNSUPPORTED, "too many beam pixels for sampling");
  Singular S;
  S.pixels.resize(ns);
  if (ns) {
    DIO_HIP(hipMemcpyAsync(S.pixels.data(), sing.p, ns * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           s));
    DIO_HIP(hipStreamSynchronize(s));
  }
  singular_prepare(g, SG, twod ? &G : nullptr, S, cnt.p, s);
  out.uncertain += S.patched;
  out.singular = ns;
  exclusive_scan(cnt.p, off.p, npix + 1, s, sc);
  uint64_t nnz = read1(off.p + npix, s);
  if (!twod && nnz > 64ull * 1024 * 1024)
    fail(DIO_ERR_UNSUPPORTED,
         "pyFAI-compatible splitting exceeds the 64-million-entry memory bound; reduce binning.");
  if (nnz > 0x7fffffffull) fail(DIO_ERR_UNSUPPORTED, "table exceeds 2^31 entries");
  ebin.alloc(nnz);
  epix.alloc(nnz);
  ew.alloc(nnz);
  if (twod)
    k_pf_fill<true><<<blocks_for(npix, kThreads), kThreads, 0, s>>>(P, (uint32_t)npix, off.p,
                                                                   ebin.p, epix.p, ew.p);
  else
    k_pf_fill<false><<<blocks_for(npix, kThreads), kThreads, 0, s>>>(P, (uint32_t)npix, off.p,
                                                                    ebin.p, epix.p, ew.p);
  if (ns)
    k_sing_fill<<<blocks_for(ns, 64), 64, 0, s>>>(S.hits.p, nbins, ns, S.dpix.p, off.p, ebin.p,
                                                  epix.p, ew.p);
  DIO_HIP(hipGetLastError());
  DIO_HIP(hipStreamSynchronize(s));
  assemble(nbins, nnz, ebin, epix, ew, s, sc, out.full);
  out.norm.alloc(npix);
  k_normalization<<<blocks_for(npix, kThreads), kThreads, 0, s>>>(
      g, ttp.p, chp.p, d.has_polarization, d.polarization_factor, out.norm.p);
  DIO_HIP(hipGetLastError());
  DIO_HIP(hipStreamSynchronize(s));
}

}  // namespace

Geo make_geo(const dio_poni& p, const dio_plan_desc& d) {
  Geo g;
  g.pixel1 = p.pixel1;
  g.pixel2 = p.pixel2;
  g.poni1 = p.poni1;
  g.poni2 = p.poni2;
  g.distance = p.distance;
  // geometry.rs:640-660: host libm sin/cos of the rotations.
  g.s1 = ::sin(p.rot1);
  g.c1 = ::cos(p.rot1);
  g.s2 = ::sin(p.rot2);
  g.c2 = ::cos(p.rot2);
  g.s3 = ::sin(p.rot3);
  g.c3 = ::cos(p.rot3);
  g.wavelength = p.wavelength;
  g.unit = d.unit;
  // integration.rs:154-169 (RadialConverter::new): 4 pi / (lambda * scale).
  g.q_scale = 0.0;
  if (d.unit == DIO_UNIT_Q_A || d.unit == DIO_UNIT_Q_NM)
    g.q_scale = 4.0 * kPiD / (p.wavelength * (d.unit == DIO_UNIT_Q_A ? 1e10 : 1e9));
  g.dim1 = d.dim1;
  g.dim2 = d.dim2;
  return g;
}

void build_tables(const Geo& g, const dio_plan_desc& d, hipStream_t s, Built& out) {
  Scratch sc;
  if (d.method == DIO_SUPERSAMPLED) build_ss(g, d, s, sc, out);
  else if (d.method == DIO_PYFAI_FULL) build_pyfai(g, d, s, sc, out);
  else if (d.ndim == 2) build_2d(g, d, s, sc, out);
  else build_1d(g, d, s, sc, out);
  uint64_t npix = (uint64_t)d.dim1 * d.dim2;
  bool any_mask = false;
  if (d.mask) {
    for (uint64_t i = 0; i < npix && !any_mask; i++) any_mask = d.mask[i] != 0;
  }
  if (any_mask) {
    out.mask.alloc(npix);
    DIO_HIP(hipMemcpyAsync(out.mask.p, d.mask, npix, hipMemcpyHostToDevice, s));
    compact(out.full, out.mask.p, s, sc, out.exec);
    out.has_exec = true;
  }
  const Table& t = out.exec_table();
  out.den.alloc(t.nbins);
  k_denominators<<<blocks_for(t.nbins, 64), 64, 0, s>>>(t.row_ptr.p, t.col.p, t.w64.p,
                                                        out.weighted ? 1 : 0, out.norm.p,
                                                        t.nbins, out.den.p);
  DIO_HIP(hipGetLastError());
  DIO_HIP(hipStreamSynchronize(s));
}

void bin_counts(const Built& b, double* cnt, hipStream_t s) {
  const Table& t = b.exec_table();
  k_counts<<<blocks_for(t.nbins, 64), 64, 0, s>>>(t.row_ptr.p, t.w64.p, b.weighted ? 1 : 0,
                                                  t.nbins, cnt);
  DIO_HIP(hipGetLastError());
}

}  // namespace dio
