// CSR gather kernels (gfx950). The path is a sparse, bandwidth-bound
// reduction, so it uses the vector ALUs, not MFMA.
//
// Layout in HBM per plan: col u32[nnz], w32 f32[nnz] (or w64 f64[nnz] for
// the F64 precision), den f64[bins], segment bounds u32[nseg] x 2.
// Frames: n_frames x npix elements of the frame dtype, row-major.
//
// k_seg: one wavefront per segment (<= kSegLen entries of one row) and a tile
// of F frames. Lane l reads entries begin + l, begin + l + 64, ... once
// (coalesced column and weight streams) and applies each to all F frames
// held in registers, so the table is read once per tile instead of once per
// frame. The 64 lane sums are reduced by a fixed shuffle tree and lane 0
// writes one partial per frame. k_finalize adds a row's partials in segment
// order and divides by the FP64 denominator. Every frame's arithmetic is
// independent of F and of the launch shape, so a frame gives the same bits
// alone or in any batch (reference contract: cuda/tests.rs:288, 377).
//
// Block index -> segment mapping is XCD-aware: the hardware deals blocks to
// the 8 XCDs round robin, so block b is remapped to run segment band
// (b % 8) * per + b / 8. Each XCD then walks a contiguous band of rows
// (neighbouring rings), and the frame lines those rows share stay in that
// XCD's L2.
#include <cfloat>
#include <vector>

#include "gather.hpp"

namespace dio {

namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;

// Frame element -> arithmetic value. Float32-storage precisions see the
// frame as float32 (integers converted in register, float64 rounded), like
// the reference's typed kernels (cuda.rs:346-353).
template <class P> __device__ inline float to_f32(P v) { return static_cast<float>(v); }
template <class P> __device__ inline double to_f64_exact(P v) { return static_cast<double>(v); }

// ---------------------------------------------------------- accumulators

// MODE 0: float32 fused multiply-add (CudaPrecision::F32).
struct AccPlain {
  float s;
  __device__ void init() { s = 0.0f; }
  template <class P> __device__ void add(float w, P raw) { s = fmaf(w, to_f32(raw), s); }
  template <class P> __device__ void add_var(float w, P raw) {
    s = fmaf(w * w, fmaxf(to_f32(raw), 0.0f), s);
  }
  __device__ void merge(const AccPlain& o) { s = s + o.s; }
  __device__ AccPlain down(int off) const { return {__shfl_down(s, off, kWave)}; }
  __device__ double value() const { return static_cast<double>(s); }
};

// MODE 1: double-word (hi, lo) float32 accumulation (F32Compensated):
// exact products by FMA residual, Knuth TwoSum, renormalization; a
// non-finite total is kept as is (+-inf or NaN), as the CPU does.
struct AccComp {
  float hi, lo;
  __device__ void init() { hi = 0.0f; lo = 0.0f; }
  __device__ void add_pair(float th, float tl) {
    float s = hi + th;
    if (!(fabsf(s) <= FLT_MAX)) {
      hi = s;
      lo = 0.0f;
      return;
    }
    float bb = s - hi;
    float e = (hi - (s - bb)) + (th - bb);
    float t = (e + lo) + tl;
    float h = s + t;
    lo = t - (h - s);
    hi = h;
  }
  template <class P> __device__ void add(float w, P raw) {
    float x = to_f32(raw);
    float p = w * x;
    add_pair(p, fmaf(w, x, -p));
  }
  template <class P> __device__ void add_var(float w, P raw) {
    float x = fmaxf(to_f32(raw), 0.0f);
    float w2 = w * w;
    float w2l = fmaf(w, w, -w2);
    float p = w2 * x;
    float pl = fmaf(w2, x, -p);
    pl = fmaf(w2l, x, pl);
    add_pair(p, pl);
  }
  __device__ void merge(const AccComp& o) { add_pair(o.hi, o.lo); }
  __device__ AccComp down(int off) const {
    return {__shfl_down(hi, off, kWave), __shfl_down(lo, off, kWave)};
  }
  __device__ double value() const { return static_cast<double>(hi) + static_cast<double>(lo); }
};

// MODE 3: float32 storage, float64 register accumulation. On CDNA the FP64
// FMA issues at the FP32 rate, so this is one DFMA per entry and frame.
struct AccWide {
  double s;
  __device__ void init() { s = 0.0; }
  template <class P> __device__ void addx(float w, double x) { s = fma((double)w, x, s); }
  template <class P> __device__ void add(float w, P raw) { addx<P>(w, wide_value(raw)); }
  template <class P> __device__ void add_var(float w, P raw) {
    double wd = (double)w;
    s = fma(wd * wd, fmax(wide_value(raw), 0.0), s);
  }
  // float64 frames are rounded to float32 like the other float32-storage
  // modes; integers and float32 convert exactly.
  template <class P> __device__ static double wide_value(P raw) {
    if constexpr (sizeof(P) == 8) return (double)(float)raw;
    else return (double)raw;
  }
  __device__ void merge(const AccWide& o) { s = s + o.s; }
  __device__ AccWide down(int off) const { return {__shfl_down(s, off, kWave)}; }
  __device__ double value() const { return s; }
};

// MODE 2: float64 weights, float64 pixels, float64 accumulation
// (CudaPrecision::F64).
struct AccF64 {
  double s;
  __device__ void init() { s = 0.0; }
  template <class P> __device__ void add64(double w, P raw) { s = fma(w, (double)raw, s); }
  template <class P> __device__ void add_var64(double w, P raw) {
    s = fma(w * w, fmax((double)raw, 0.0), s);
  }
  __device__ void merge(const AccF64& o) { s = s + o.s; }
  __device__ AccF64 down(int off) const { return {__shfl_down(s, off, kWave)}; }
  __device__ double value() const { return s; }
};

template <int MODE> struct AccOf;
template <> struct AccOf<0> { using T = AccPlain; };
template <> struct AccOf<1> { using T = AccComp; };
template <> struct AccOf<2> { using T = AccF64; };
template <> struct AccOf<3> { using T = AccWide; };

__device__ inline uint32_t xcd_block(uint32_t b, uint32_t grid) {
  uint32_t per = grid / 8;  // grid is a multiple of 8
  return (b % 8) * per + b / 8;
}

template <int F, int MODE, class P, bool VAR>
__global__ __launch_bounds__(kBlock) void k_seg(const uint32_t* __restrict__ col,
                                                const float* __restrict__ w32,
                                                const double* __restrict__ w64,
                                                const uint32_t* __restrict__ seg_begin,
                                                const uint32_t* __restrict__ seg_end,
                                                uint32_t nseg, const P* __restrict__ frames,
                                                uint64_t npix, uint32_t nf,
                                                double* __restrict__ part) {
  using Acc = typename AccOf<MODE>::T;
  uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  uint32_t seg = xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + wave;
  if (seg >= nseg) return;
  Acc a[F], v[F];
#pragma unroll
  for (int f = 0; f < F; f++) {
    a[f].init();
    if constexpr (VAR) v[f].init();
  }
  // Frames past the tile end alias the last frame; their sums are dropped.
  const P* fp[F];
#pragma unroll
  for (int f = 0; f < F; f++) fp[f] = frames + (uint64_t)(f < (int)nf ? f : nf - 1) * npix;
  uint32_t end = seg_end[seg];
  for (uint32_t k = seg_begin[seg] + lane; k < end; k += kWave) {
    uint32_t c = col[k];
    if constexpr (MODE == 2) {
      double w = w64[k];
#pragma unroll
      for (int f = 0; f < F; f++) {
        P x = fp[f][c];
        a[f].add64(w, x);
        if constexpr (VAR) v[f].add_var64(w, x);
      }
    } else {
      float w = w32[k];
#pragma unroll
      for (int f = 0; f < F; f++) {
        P x = fp[f][c];
        a[f].add(w, x);
        if constexpr (VAR) v[f].add_var(w, x);
      }
    }
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
#pragma unroll
    for (int f = 0; f < F; f++) {
      a[f].merge(a[f].down(off));
      if constexpr (VAR) v[f].merge(v[f].down(off));
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int f = 0; f < F; f++) {
      if (f < (int)nf) {
        part[(uint64_t)f * nseg + seg] = a[f].value();
        if constexpr (VAR) part[(uint64_t)(F + f) * nseg + seg] = v[f].value();
      }
    }
  }
}

// prepared.rs:985-997 (normalize_bin): I = s / den, sigma =
// sqrt(max(v / den^2, 0)), both 0 where den == 0.
template <int MODE, bool VAR>
__global__ void k_finalize(const double* __restrict__ part, uint32_t nseg, uint32_t F,
                           const uint32_t* __restrict__ row_seg, const double* __restrict__ den,
                           uint32_t nbins, uint32_t nf, double* __restrict__ I,
                           double* __restrict__ S) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)nbins * nf) return;
  uint32_t r = (uint32_t)(t % nbins), f = (uint32_t)(t / nbins);
  uint32_t q0 = row_seg[r], q1 = row_seg[r + 1];
  double s, v = 0.0;
  if constexpr (MODE == 0) {
    float sf = 0.0f, vf = 0.0f;
    for (uint32_t q = q0; q < q1; q++) {
      sf = sf + (float)part[(uint64_t)f * nseg + q];
      if constexpr (VAR) vf = vf + (float)part[(uint64_t)(F + f) * nseg + q];
    }
    s = sf;
    v = vf;
  } else {
    s = 0.0;
    for (uint32_t q = q0; q < q1; q++) {
      s = s + part[(uint64_t)f * nseg + q];
      if constexpr (VAR) v = v + part[(uint64_t)(F + f) * nseg + q];
    }
  }
  double d = den[r];
  I[(uint64_t)f * nbins + r] = d == 0.0 ? 0.0 : s / d;
  if constexpr (VAR) {
    double vv = d == 0.0 ? 0.0 : v / (d * d);
    S[(uint64_t)f * nbins + r] = sqrt(fmax(vv, 0.0));
  }
}

// --------------------------------------------------------- azimuthal std

__device__ inline double wave_sum(double x) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) x = x + __shfl_down(x, off, kWave);
  return __shfl(x, 0, kWave);
}

// azimuthal.rs:48-118: weighted population spread of corrected pixels on the
// unmasked table with the mask applied per entry. One wavefront per row;
// FP64 arithmetic on float32 weights.
template <class P>
__global__ __launch_bounds__(kBlock) void k_std(const uint32_t* __restrict__ row_ptr,
                                                const uint32_t* __restrict__ col,
                                                const float* __restrict__ w32,
                                                const uint8_t* __restrict__ mask,
                                                const double* __restrict__ norm,
                                                const P* __restrict__ frames, uint64_t npix,
                                                uint32_t nbins, int errors, double* I, double* S,
                                                double* SD) {
  uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  uint32_t r = blockIdx.x * kWavesPerBlock + wave;
  if (r >= nbins) return;
  const P* fr = frames + (uint64_t)blockIdx.y * npix;
  uint64_t ob = (uint64_t)blockIdx.y * nbins + r;
  uint32_t k0 = row_ptr[r], k1 = row_ptr[r + 1];
  auto contrib = [&](uint32_t k, double& c, double& s, double& x, double& w) -> bool {
    uint32_t j = col[k];
    if (mask && mask[j]) return false;
    double n = norm[j];
    c = (double)w32[k];
    s = (double)fr[j];
    x = s / n;
    w = c * n;
    return isfinite(s) && isfinite(n) && n > 0.0 && isfinite(c) && c > 0.0 && isfinite(x) &&
           isfinite(w) && w > 0.0;
  };
  // Origin: the first valid contribution in row order.
  bool found = false;
  double origin = 0.0;
  for (uint32_t base = k0; base < k1 && !found; base += kWave) {
    uint32_t k = base + lane;
    double c, s, x, w;
    bool ok = k < k1 && contrib(k, c, s, x, w);
    unsigned long long m = __ballot(ok);
    if (m) {
      int first = __ffsll((long long)m) - 1;
      origin = __shfl(ok ? x : 0.0, first, kWave);
      found = true;
    }
  }
  if (!found) {
    if (lane == 0) {
      I[ob] = 0.0;
      if (errors) S[ob] = 0.0;
      SD[ob] = NAN;
    }
    return;
  }
  double ss = 0.0, sw = 0.0, sh = 0.0, sv = 0.0;
  for (uint32_t k = k0 + lane; k < k1; k += kWave) {
    double c, s, x, w;
    if (!contrib(k, c, s, x, w)) continue;
    ss = ss + c * s;
    sw = sw + w;
    sh = sh + w * (x - origin);
    sv = sv + c * c * fmax(s, 0.0);
  }
  ss = wave_sum(ss);
  sw = wave_sum(sw);
  sh = wave_sum(sh);
  sv = wave_sum(sv);
  double mean_shifted = sh / sw;
  double m2 = 0.0;
  for (uint32_t k = k0 + lane; k < k1; k += kWave) {
    double c, s, x, w;
    if (!contrib(k, c, s, x, w)) continue;
    double delta = (x - origin) - mean_shifted;
    m2 = m2 + w * delta * delta;
  }
  m2 = wave_sum(m2);
  if (lane == 0) {
    I[ob] = ss / sw;
    if (errors) S[ob] = sqrt(fmax(sv, 0.0)) / sw;
    SD[ob] = sqrt(fmax(m2 / sw, 0.0));
  }
}

// ------------------------------------------------------------- dispatch

using SegFn = void (*)(const uint32_t*, const float*, const double*, const uint32_t*,
                       const uint32_t*, uint32_t, const void*, uint64_t, uint32_t, double*);

template <int F, int MODE, class P, bool VAR>
void launch_seg(const GatherTable& t, const void* frames, uint32_t nf, double* part,
                hipStream_t s) {
  uint32_t blocks = (t.nseg + kWavesPerBlock - 1) / kWavesPerBlock;
  blocks = (blocks + 7) / 8 * 8;
  k_seg<F, MODE, P, VAR><<<blocks, kBlock, 0, s>>>(t.col, t.w32, t.w64, t.seg_begin.p,
                                                   t.seg_end.p, t.nseg,
                                                   static_cast<const P*>(frames), t.npix, nf,
                                                   part);
}

template <int F, int MODE, class P>
void launch_seg_v(const GatherTable& t, const void* frames, uint32_t nf, bool var, double* part,
                  hipStream_t s) {
  if (var) launch_seg<F, MODE, P, true>(t, frames, nf, part, s);
  else launch_seg<F, MODE, P, false>(t, frames, nf, part, s);
}

template <int F, int MODE>
void launch_seg_p(const GatherTable& t, const void* frames, int dtype, uint32_t nf, bool var,
                  double* part, hipStream_t s) {
  switch (dtype) {
    case DIO_U16: launch_seg_v<F, MODE, uint16_t>(t, frames, nf, var, part, s); break;
    case DIO_U32: launch_seg_v<F, MODE, uint32_t>(t, frames, nf, var, part, s); break;
    case DIO_I32: launch_seg_v<F, MODE, int32_t>(t, frames, nf, var, part, s); break;
    case DIO_F32T: launch_seg_v<F, MODE, float>(t, frames, nf, var, part, s); break;
    case DIO_F64T: launch_seg_v<F, MODE, double>(t, frames, nf, var, part, s); break;
    default: fail(DIO_ERR_INVALID, "unknown frame dtype");
  }
}

template <int F>
void launch_seg_m(const GatherTable& t, const void* frames, int dtype, uint32_t nf, bool var,
                  int precision, double* part, hipStream_t s) {
  switch (precision) {
    case DIO_F32: launch_seg_p<F, 0>(t, frames, dtype, nf, var, part, s); break;
    case DIO_F32_COMPENSATED: launch_seg_p<F, 1>(t, frames, dtype, nf, var, part, s); break;
    case DIO_F64: launch_seg_p<F, 2>(t, frames, dtype, nf, var, part, s); break;
    case DIO_F32_WIDE: launch_seg_p<F, 3>(t, frames, dtype, nf, var, part, s); break;
    default: fail(DIO_ERR_INVALID, "unknown precision");
  }
}

void launch_seg_tile(const GatherTable& t, uint32_t F, const void* frames, int dtype,
                     uint32_t nf, bool var, int precision, double* part, hipStream_t s) {
  switch (F) {
    case 1: launch_seg_m<1>(t, frames, dtype, nf, var, precision, part, s); break;
    case 2: launch_seg_m<2>(t, frames, dtype, nf, var, precision, part, s); break;
    case 4: launch_seg_m<4>(t, frames, dtype, nf, var, precision, part, s); break;
    case 8: launch_seg_m<8>(t, frames, dtype, nf, var, precision, part, s); break;
    case 16: launch_seg_m<16>(t, frames, dtype, nf, var, precision, part, s); break;
    case 32: launch_seg_m<32>(t, frames, dtype, nf, var, precision, part, s); break;
    default: fail(DIO_ERR_INVALID, "frames per tile must be 1, 2, 4, 8, 16 or 32");
  }
}

size_t dtype_size(int dtype) {
  switch (dtype) {
    case DIO_U16: return 2;
    case DIO_U32: case DIO_I32: case DIO_F32T: return 4;
    case DIO_F64T: return 8;
    default: fail(DIO_ERR_INVALID, "unknown frame dtype");
  }
}

}  // namespace

void make_segments(GatherTable& t, const uint32_t* d_row_ptr, hipStream_t s) {
  std::vector<uint32_t> rp(t.nbins + 1);
  DIO_HIP(hipMemcpyAsync(rp.data(), d_row_ptr, (t.nbins + 1) * sizeof(uint32_t),
                         hipMemcpyDeviceToHost, s));
  DIO_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> b, e, rs(t.nbins + 1);
  for (uint32_t r = 0; r < t.nbins; r++) {
    rs[r] = (uint32_t)b.size();
    for (uint32_t k = rp[r]; k < rp[r + 1]; k += kSegLen) {
      b.push_back(k);
      e.push_back(std::min(rp[r + 1], k + kSegLen));
    }
  }
  rs[t.nbins] = (uint32_t)b.size();
  t.nseg = (uint32_t)b.size();
  t.seg_begin.alloc(std::max<size_t>(1, b.size()));
  t.seg_end.alloc(std::max<size_t>(1, e.size()));
  t.row_seg.alloc(t.nbins + 1);
  if (!b.empty()) {
    DIO_HIP(hipMemcpyAsync(t.seg_begin.p, b.data(), b.size() * 4, hipMemcpyHostToDevice, s));
    DIO_HIP(hipMemcpyAsync(t.seg_end.p, e.data(), e.size() * 4, hipMemcpyHostToDevice, s));
  }
  DIO_HIP(hipMemcpyAsync(t.row_seg.p, rs.data(), rs.size() * 4, hipMemcpyHostToDevice, s));
  DIO_HIP(hipStreamSynchronize(s));
}

void gather_segments_only(const GatherTable& t, GatherScratch& sc, const void* frames, int dtype,
                          uint32_t nf, bool var, int precision, hipStream_t s) {
  uint32_t F = 1;
  while (F < nf) F <<= 1;
  sc.ensure(t, F, var);
  if (t.nseg) launch_seg_tile(t, F, frames, dtype, nf, var, precision, sc.part.p, s);
  DIO_HIP(hipGetLastError());
}

void gather_frames(const GatherTable& t, GatherScratch& sc, const void* frames, int dtype,
                   uint64_t n_frames, bool var, int precision, uint32_t tile, double* intensity,
                   double* sigma, hipStream_t s) {
  if (tile == 0 || tile > kMaxTile || (tile & (tile - 1)))
    fail(DIO_ERR_INVALID, "frames per tile must be a power of two in 1..32");
  size_t esz = dtype_size(dtype);
  sc.ensure(t, tile, var);
  const char* fb = static_cast<const char*>(frames);
  for (uint64_t f0 = 0; f0 < n_frames; f0 += tile) {
    uint32_t nf = (uint32_t)std::min<uint64_t>(tile, n_frames - f0);
    const void* fr = fb + f0 * t.npix * esz;
    if (t.nseg) launch_seg_tile(t, tile, fr, dtype, nf, var, precision, sc.part.p, s);
    uint64_t work = (uint64_t)t.nbins * nf;
    unsigned blocks = blocks_for(work, 256);
    double* I = intensity + f0 * t.nbins;
    double* S = var ? sigma + f0 * t.nbins : nullptr;
    if (precision == DIO_F32) {
      if (var) k_finalize<0, true><<<blocks, 256, 0, s>>>(sc.part.p, t.nseg, tile, t.row_seg.p,
                                                          t.den, t.nbins, nf, I, S);
      else k_finalize<0, false><<<blocks, 256, 0, s>>>(sc.part.p, t.nseg, tile, t.row_seg.p,
                                                       t.den, t.nbins, nf, I, S);
    } else {
      if (var) k_finalize<1, true><<<blocks, 256, 0, s>>>(sc.part.p, t.nseg, tile, t.row_seg.p,
                                                          t.den, t.nbins, nf, I, S);
      else k_finalize<1, false><<<blocks, 256, 0, s>>>(sc.part.p, t.nseg, tile, t.row_seg.p,
                                                       t.den, t.nbins, nf, I, S);
    }
    DIO_HIP(hipGetLastError());
  }
}

void azimuthal_std_frames(const StdTable& t, const void* frames, int dtype, uint64_t n_frames,
                          bool errors, double* intensity, double* sigma, double* std_out,
                          hipStream_t s) {
  if (n_frames == 0) return;
  dim3 grid((t.nbins + kWavesPerBlock - 1) / kWavesPerBlock, (uint32_t)n_frames);
  switch (dtype) {
#define DIO_STD(T)                                                                            \
  k_std<T><<<grid, kBlock, 0, s>>>(t.row_ptr, t.col, t.w32, t.mask, t.norm,                  \
                                   static_cast<const T*>(frames), t.npix, t.nbins, errors,     \
                                   intensity, sigma, std_out);                                 \
  break;
    case DIO_U16: DIO_STD(uint16_t)
    case DIO_U32: DIO_STD(uint32_t)
    case DIO_I32: DIO_STD(int32_t)
    case DIO_F32T: DIO_STD(float)
    case DIO_F64T: DIO_STD(double)
#undef DIO_STD
    default: fail(DIO_ERR_INVALID, "unknown frame dtype");
  }
  DIO_HIP(hipGetLastError());
}

}  // namespace dio
