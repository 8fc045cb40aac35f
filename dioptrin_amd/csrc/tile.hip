// Detector-tiled gather (gfx950): the batched integration hot path.
//
// HBM layout of a tiled plan: lidx u16[nnz] + w32 f32[nnz] (6 bytes per
// entry), segment starts u32[nseg + 1], tile -> segment u32[ntiles + 1],
// bin -> segment list u32[nseg] + u32[nbins + 1], denominators f64[nbins].
//
// k_tile: one workgroup per detector tile (16 rows x 64 columns). It copies
// the tile of F frames from HBM into LDS with coalesced row loads (each frame
// byte is read from HBM exactly once per pass), then one thread per (tile,
// bin) segment applies the segment's entries, pixels ascending, to all F
// frames held in registers, and writes one partial per frame. The table is
// read once per pass of F frames. k_tfinal adds each bin's partials in tile
// order and divides by the FP64 denominator. Neither the summation order of a
// segment nor the order of partials depends on F, so a frame gives the same
// bits alone or in any batch.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <vector>

#include "tile.hpp"

namespace dio {

namespace {

constexpr unsigned kThreads = 256;

// ---------------------------------------------------------------- layout

__global__ void k_entry_keys(const uint32_t* row_ptr, uint32_t nbins, const uint32_t* col,
                             uint64_t nnz, uint32_t dim2, uint32_t ntx, uint32_t* tile,
                             uint32_t* bin, uint16_t* lidx) {
  uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  // Row of entry k: last r with row_ptr[r] <= k.
  uint32_t lo = 0, hi = nbins;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) / 2;
    if (row_ptr[mid] <= k) lo = mid;
    else hi = mid;
  }
  uint32_t p = col[k];
  uint32_t r = p / dim2, c = p % dim2;
  tile[k] = (r / kTileRows) * ntx + c / kTileCols;
  bin[k] = lo;
  lidx[k] = (uint16_t)((r % kTileRows) * kTileCols + c % kTileCols);
}

__global__ void k_tile_gather(const uint32_t* order, const uint32_t* tile_in,
                              const uint32_t* bin_in, const uint16_t* lidx_in,
                              const float* w32_in, const double* w64_in, uint64_t nnz,
                              uint32_t* tile, uint32_t* bin, uint16_t* lidx, float* w32,
                              double* w64) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  uint32_t k = order[i];
  tile[i] = tile_in[k];
  bin[i] = bin_in[k];
  lidx[i] = lidx_in[k];
  w32[i] = w32_in[k];
  if (w64) w64[i] = w64_in[k];
}

__global__ void k_seg_flags(const uint32_t* tile, const uint32_t* bin, uint64_t nnz,
                            uint32_t* flag) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nnz) return;
  if (i == nnz) {
    flag[i] = 0;
    return;
  }
  flag[i] = (i == 0 || tile[i] != tile[i - 1] || bin[i] != bin[i - 1]) ? 1u : 0u;
}

__global__ void k_seg_fill(const uint32_t* flag, const uint32_t* segid, const uint32_t* tile,
                           const uint32_t* bin, uint64_t nnz, uint32_t* seg_start,
                           uint32_t* seg_tile, uint32_t* seg_bin) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz || !flag[i]) return;
  uint32_t s = segid[i];
  seg_start[s] = (uint32_t)i;
  seg_tile[s] = tile[i];
  seg_bin[s] = bin[i];
}

__global__ void k_count_u32(const uint32_t* key, uint32_t n, uint32_t* hist) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&hist[key[i]], 1u);
}

__global__ void k_iota32(uint32_t* v, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

__global__ void k_set1(uint32_t* p, uint32_t v) { *p = v; }

int bits_for(uint64_t n) {
  int b = 1;
  while (b < 32 && (1ull << b) < n) b++;
  return b;
}

struct Tmp {
  DevBuf<uint8_t> buf;
  void* get(size_t bytes) {
    if (buf.n < bytes) buf.alloc(bytes);
    return buf.p;
  }
};

void scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, hipStream_t s, Tmp& tmp) {
  size_t bytes = 0;
  DIO_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s));
  DIO_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.get(bytes), bytes, in, out, (int)n, s));
}

template <class K, class V>
void sort_pairs(const K* kin, K* kout, const V* vin, V* vout, uint64_t n, int bits, hipStream_t s,
                Tmp& tmp) {
  size_t bytes = 0;
  DIO_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, (int)n, 0,
                                             bits, s));
  DIO_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.get(bytes), bytes, kin, kout, vin, vout, (int)n,
                                             0, bits, s));
}

// ---------------------------------------------------------------- kernels

template <class P> __device__ inline float f32_of(P v) { return static_cast<float>(v); }
template <class P> __device__ inline double wide_of(P v) {
  if constexpr (sizeof(P) == 8) return (double)(float)v;
  else return (double)v;
}

// Float32 fused multiply-add (CudaPrecision::F32).
struct TPlain {
  float s;
  __device__ void init() { s = 0.0f; }
  template <class P> __device__ void add(float w, double, P x) { s = fmaf(w, f32_of(x), s); }
  template <class P> __device__ void add_var(float w, double, P x) {
    s = fmaf(w * w, fmaxf(f32_of(x), 0.0f), s);
  }
  __device__ double value() const { return (double)s; }
};

// Double-word float32 (CudaPrecision::F32Compensated): exact products by FMA
// residual, Knuth TwoSum, renormalisation; non-finite totals kept as is.
struct TComp {
  float hi, lo;
  __device__ void init() { hi = 0.0f; lo = 0.0f; }
  __device__ void pair(float th, float tl) {
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
  template <class P> __device__ void add(float w, double, P raw) {
    float x = f32_of(raw);
    float p = w * x;
    pair(p, fmaf(w, x, -p));
  }
  template <class P> __device__ void add_var(float w, double, P raw) {
    float x = fmaxf(f32_of(raw), 0.0f);
    float w2 = w * w;
    float w2l = fmaf(w, w, -w2);
    float p = w2 * x;
    float pl = fmaf(w2, x, -p);
    pl = fmaf(w2l, x, pl);
    pair(p, pl);
  }
  __device__ double value() const { return (double)hi + (double)lo; }
};

// Float32 storage, float64 register accumulation: one full-rate DFMA.
struct TWide {
  double s;
  __device__ void init() { s = 0.0; }
  template <class P> __device__ void add(float w, double, P x) { s = fma((double)w, wide_of(x), s); }
  template <class P> __device__ void add_var(float w, double, P x) {
    double wd = (double)w;
    s = fma(wd * wd, fmax(wide_of(x), 0.0), s);
  }
  __device__ double value() const { return s; }
};

// Float64 weights and accumulation (CudaPrecision::F64).
struct TF64 {
  double s;
  __device__ void init() { s = 0.0; }
  template <class P> __device__ void add(float, double w, P x) { s = fma(w, (double)x, s); }
  template <class P> __device__ void add_var(float, double w, P x) {
    s = fma(w * w, fmax((double)x, 0.0), s);
  }
  __device__ double value() const { return s; }
};

template <int MODE> struct TAcc;
template <> struct TAcc<0> { using T = TPlain; };
template <> struct TAcc<1> { using T = TComp; };
template <> struct TAcc<2> { using T = TF64; };
template <> struct TAcc<3> { using T = TWide; };

struct TileArgs {
  const uint16_t* lidx;
  const float* w32;
  const double* w64;
  const uint32_t* seg_start;
  const uint32_t* tile_seg;
  uint32_t nseg, ntx, dim1, dim2;
  uint64_t npix;
  uint32_t nf;
  double* part;
};

template <int F, int MODE, class P, bool VAR>
__global__ __launch_bounds__(kThreads) void k_tile(TileArgs a, const P* __restrict__ frames) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  P* lds = reinterpret_cast<P*>(lds_raw);
  using Acc = typename TAcc<MODE>::T;
  uint32_t t = blockIdx.x;
  uint32_t s0 = a.tile_seg[t], s1 = a.tile_seg[t + 1];
  if (s0 == s1) return;
  uint32_t row0 = (t / a.ntx) * kTileRows, col0 = (t % a.ntx) * kTileCols;
  // Stage the tile of every frame of the pass: consecutive threads read
  // consecutive columns of one detector row (coalesced).
  for (uint32_t e = threadIdx.x; e < (uint32_t)F * kTilePixels; e += kThreads) {
    uint32_t f = e / kTilePixels, px = e % kTilePixels;
    uint32_t r = row0 + px / kTileCols, c = col0 + px % kTileCols;
    P v = P(0);
    if (f < a.nf && r < a.dim1 && c < a.dim2) v = frames[(uint64_t)f * a.npix + (uint64_t)r * a.dim2 + c];
    lds[e] = v;
  }
  __syncthreads();
  for (uint32_t s = s0 + threadIdx.x; s < s1; s += kThreads) {
    Acc acc[F], var[F];
#pragma unroll
    for (int f = 0; f < F; f++) {
      acc[f].init();
      if constexpr (VAR) var[f].init();
    }
    uint32_t k1 = a.seg_start[s + 1];
    for (uint32_t k = a.seg_start[s]; k < k1; k++) {
      uint32_t l = a.lidx[k];
      float w = a.w32[k];
      double wd = MODE == 2 ? a.w64[k] : 0.0;
#pragma unroll
      for (int f = 0; f < F; f++) {
        P x = lds[f * kTilePixels + l];
        acc[f].add(w, wd, x);
        if constexpr (VAR) var[f].add_var(w, wd, x);
      }
    }
#pragma unroll
    for (int f = 0; f < F; f++) {
      if (f < (int)a.nf) {
        a.part[(uint64_t)f * a.nseg + s] = acc[f].value();
        if constexpr (VAR) a.part[(uint64_t)(F + f) * a.nseg + s] = var[f].value();
      }
    }
  }
}

// prepared.rs:985-997 (normalize_bin) after adding a bin's partials in tile
// order: I = s / den, sigma = sqrt(max(v / den^2, 0)), 0 where den == 0.
template <bool PLAIN, bool VAR>
__global__ void k_tfinal(const double* __restrict__ part, uint32_t nseg, uint32_t F,
                         const uint32_t* __restrict__ bin_ptr, const uint32_t* __restrict__ bin_seg,
                         const double* __restrict__ den, uint32_t nbins, uint32_t nf,
                         double* __restrict__ I, double* __restrict__ S) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)nbins * nf) return;
  uint32_t b = (uint32_t)(t % nbins), f = (uint32_t)(t / nbins);
  uint32_t q0 = bin_ptr[b], q1 = bin_ptr[b + 1];
  double s, v = 0.0;
  if constexpr (PLAIN) {
    float sf = 0.0f, vf = 0.0f;
    for (uint32_t q = q0; q < q1; q++) {
      uint32_t g = bin_seg[q];
      sf = sf + (float)part[(uint64_t)f * nseg + g];
      if constexpr (VAR) vf = vf + (float)part[(uint64_t)(F + f) * nseg + g];
    }
    s = sf;
    v = vf;
  } else {
    s = 0.0;
    for (uint32_t q = q0; q < q1; q++) {
      uint32_t g = bin_seg[q];
      s = s + part[(uint64_t)f * nseg + g];
      if constexpr (VAR) v = v + part[(uint64_t)(F + f) * nseg + g];
    }
  }
  double d = den[b];
  I[(uint64_t)f * nbins + b] = d == 0.0 ? 0.0 : s / d;
  if constexpr (VAR) {
    double vv = d == 0.0 ? 0.0 : v / (d * d);
    S[(uint64_t)f * nbins + b] = sqrt(fmax(vv, 0.0));
  }
}

// ------------------------------------------------------------- dispatch

constexpr size_t kMaxLds = 160 * 1024;

template <int F, int MODE, class P, bool VAR>
void launch_tile(const TiledTable& t, const void* frames, uint32_t nf, double* part,
                 hipStream_t s) {
  size_t lds = (size_t)F * kTilePixels * sizeof(P);
  if (lds > kMaxLds) fail(DIO_ERR_INVALID, "frames per pass exceed the LDS of one CU");
  static bool attr = false;
  if (!attr) {
    DIO_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile<F, MODE, P, VAR>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr = true;
  }
  TileArgs a{t.lidx.p, t.w32.p, t.w64.p, t.seg_start.p, t.tile_seg.p, t.nseg, t.ntx,
             t.dim1, t.dim2, (uint64_t)t.dim1 * t.dim2, nf, part};
  k_tile<F, MODE, P, VAR><<<t.ntiles, kThreads, lds, s>>>(a, static_cast<const P*>(frames));
}

template <int F, int MODE, class P>
void launch_tile_v(const TiledTable& t, const void* fr, uint32_t nf, bool var, double* part,
                   hipStream_t s) {
  if (var) launch_tile<F, MODE, P, true>(t, fr, nf, part, s);
  else launch_tile<F, MODE, P, false>(t, fr, nf, part, s);
}

template <int F, int MODE>
void launch_tile_p(const TiledTable& t, const void* fr, int dtype, uint32_t nf, bool var,
                   double* part, hipStream_t s) {
  switch (dtype) {
    case DIO_U16: launch_tile_v<F, MODE, uint16_t>(t, fr, nf, var, part, s); break;
    case DIO_U32: launch_tile_v<F, MODE, uint32_t>(t, fr, nf, var, part, s); break;
    case DIO_I32: launch_tile_v<F, MODE, int32_t>(t, fr, nf, var, part, s); break;
    case DIO_F32T: launch_tile_v<F, MODE, float>(t, fr, nf, var, part, s); break;
    case DIO_F64T:
      if constexpr (F <= 16) launch_tile_v<F, MODE, double>(t, fr, nf, var, part, s);
      else fail(DIO_ERR_INVALID, "float64 frames allow at most 16 frames per pass");
      break;
    default: fail(DIO_ERR_INVALID, "unknown frame dtype");
  }
}

template <int F>
void launch_tile_m(const TiledTable& t, const void* fr, int dtype, uint32_t nf, bool var,
                   int precision, double* part, hipStream_t s) {
  switch (precision) {
    case DIO_F32: launch_tile_p<F, 0>(t, fr, dtype, nf, var, part, s); break;
    case DIO_F32_COMPENSATED: launch_tile_p<F, 1>(t, fr, dtype, nf, var, part, s); break;
    case DIO_F64: launch_tile_p<F, 2>(t, fr, dtype, nf, var, part, s); break;
    case DIO_F32_WIDE: launch_tile_p<F, 3>(t, fr, dtype, nf, var, part, s); break;
    default: fail(DIO_ERR_INVALID, "unknown precision");
  }
}

void launch_pass(const TiledTable& t, uint32_t F, const void* fr, int dtype, uint32_t nf,
                 bool var, int precision, double* part, hipStream_t s) {
  switch (F) {
    case 1: launch_tile_m<1>(t, fr, dtype, nf, var, precision, part, s); break;
    case 2: launch_tile_m<2>(t, fr, dtype, nf, var, precision, part, s); break;
    case 4: launch_tile_m<4>(t, fr, dtype, nf, var, precision, part, s); break;
    case 8: launch_tile_m<8>(t, fr, dtype, nf, var, precision, part, s); break;
    case 16: launch_tile_m<16>(t, fr, dtype, nf, var, precision, part, s); break;
    case 32: launch_tile_m<32>(t, fr, dtype, nf, var, precision, part, s); break;
    default: fail(DIO_ERR_INVALID, "frames per pass must be 1, 2, 4, 8, 16 or 32");
  }
  DIO_HIP(hipGetLastError());
}

size_t dtype_bytes(int dtype) {
  switch (dtype) {
    case DIO_U16: return 2;
    case DIO_U32: case DIO_I32: case DIO_F32T: return 4;
    case DIO_F64T: return 8;
    default: fail(DIO_ERR_INVALID, "unknown frame dtype");
  }
}

}  // namespace

uint32_t max_frames_for(int dtype) { return dtype == DIO_F64T ? 16 : kMaxFrames; }

void build_tiled(TiledTable& t, uint32_t dim1, uint32_t dim2, uint32_t nbins, uint64_t nnz,
                 const uint32_t* row_ptr, const uint32_t* col, const float* w32,
                 const double* w64, bool keep_w64, hipStream_t s) {
  Tmp tmp;
  t.dim1 = dim1;
  t.dim2 = dim2;
  t.nbins = nbins;
  t.nnz = nnz;
  t.ntx = (dim2 + kTileCols - 1) / kTileCols;
  t.nty = (dim1 + kTileRows - 1) / kTileRows;
  t.ntiles = t.ntx * t.nty;
  t.lidx.alloc(std::max<uint64_t>(nnz, 1));
  t.w32.alloc(std::max<uint64_t>(nnz, 1));
  if (keep_w64) t.w64.alloc(std::max<uint64_t>(nnz, 1));
  t.tile_seg.alloc(t.ntiles + 1);
  t.bin_ptr.alloc(nbins + 1);
  if (nnz == 0) {
    t.nseg = 0;
    t.seg_start.alloc(1);
    t.bin_seg.alloc(1);
    DIO_HIP(hipMemsetAsync(t.seg_start.p, 0, 4, s));
    DIO_HIP(hipMemsetAsync(t.tile_seg.p, 0, (t.ntiles + 1) * 4, s));
    DIO_HIP(hipMemsetAsync(t.bin_ptr.p, 0, (nbins + 1) * 4, s));
    DIO_HIP(hipStreamSynchronize(s));
    return;
  }
  DevBuf<uint32_t> tile0(nnz), bin0(nnz), order0(nnz), order(nnz), tkey(nnz), tile1(nnz),
      bin1(nnz);
  DevBuf<uint16_t> lidx0(nnz);
  k_entry_keys<<<blocks_for(nnz, kThreads), kThreads, 0, s>>>(row_ptr, nbins, col, nnz, dim2,
                                                             t.ntx, tile0.p, bin0.p, lidx0.p);
  k_iota32<<<blocks_for(nnz, kThreads), kThreads, 0, s>>>(order0.p, (uint32_t)nnz);
  // Stable by tile: entries keep their (bin, pixel) order inside a tile.
  sort_pairs(tile0.p, tkey.p, order0.p, order.p, nnz, bits_for(t.ntiles), s, tmp);
  k_tile_gather<<<blocks_for(nnz, kThreads), kThreads, 0, s>>>(
      order.p, tile0.p, bin0.p, lidx0.p, w32, keep_w64 ? w64 : nullptr, nnz, tile1.p, bin1.p,
      t.lidx.p, t.w32.p, keep_w64 ? t.w64.p : nullptr);
  DevBuf<uint32_t> flag(nnz + 1), segid(nnz + 1);
  k_seg_flags<<<blocks_for(nnz + 1, kThreads), kThreads, 0, s>>>(tile1.p, bin1.p, nnz, flag.p);
  scan_u32(flag.p, segid.p, nnz + 1, s, tmp);
  uint32_t nseg = 0;
  DIO_HIP(hipMemcpyAsync(&nseg, segid.p + nnz, 4, hipMemcpyDeviceToHost, s));
  DIO_HIP(hipStreamSynchronize(s));
  t.nseg = nseg;
  t.seg_start.alloc(nseg + 1);
  DevBuf<uint32_t> seg_tile(nseg), seg_bin(nseg);
  k_seg_fill<<<blocks_for(nnz, kThreads), kThreads, 0, s>>>(flag.p, segid.p, tile1.p, bin1.p,
                                                           nnz, t.seg_start.p, seg_tile.p,
                                                           seg_bin.p);
  k_set1<<<1, 1, 0, s>>>(t.seg_start.p + nseg, (uint32_t)nnz);
  // Tile -> first segment.
  DevBuf<uint32_t> hist(std::max(t.ntiles, nbins) + 1);
  DIO_HIP(hipMemsetAsync(hist.p, 0, (t.ntiles + 1) * 4, s));
  k_count_u32<<<blocks_for(nseg, kThreads), kThreads, 0, s>>>(seg_tile.p, nseg, hist.p);
  scan_u32(hist.p, t.tile_seg.p, t.ntiles + 1, s, tmp);
  // Bin -> segments in tile order (stable sort of segment ids by bin).
  DevBuf<uint32_t> sid(nseg), bkey(nseg);
  t.bin_seg.alloc(nseg);
  k_iota32<<<blocks_for(nseg, kThreads), kThreads, 0, s>>>(sid.p, nseg);
  sort_pairs(seg_bin.p, bkey.p, sid.p, t.bin_seg.p, nseg, bits_for(nbins), s, tmp);
  DIO_HIP(hipMemsetAsync(hist.p, 0, (nbins + 1) * 4, s));
  k_count_u32<<<blocks_for(nseg, kThreads), kThreads, 0, s>>>(seg_bin.p, nseg, hist.p);
  scan_u32(hist.p, t.bin_ptr.p, nbins + 1, s, tmp);
  DIO_HIP(hipGetLastError());
  DIO_HIP(hipStreamSynchronize(s));
}

void tiled_pass_only(const TiledTable& t, TileScratch& sc, const void* frames, int dtype,
                     uint32_t nf, bool var, int precision, hipStream_t s) {
  uint32_t F = 1;
  while (F < nf) F <<= 1;
  sc.ensure(t, F, var);
  if (t.nseg) launch_pass(t, F, frames, dtype, nf, var, precision, sc.part.p, s);
}

void tiled_gather(const TiledTable& t, TileScratch& sc, const void* frames, int dtype,
                  uint64_t n_frames, bool var, int precision, uint32_t F, double* intensity,
                  double* sigma, hipStream_t s) {
  if (F == 0 || F > max_frames_for(dtype) || (F & (F - 1)))
    fail(DIO_ERR_INVALID, "frames per pass must be a power of two within the LDS limit");
  size_t esz = dtype_bytes(dtype);
  sc.ensure(t, F, var);
  const char* fb = static_cast<const char*>(frames);
  uint64_t npix = (uint64_t)t.dim1 * t.dim2;
  for (uint64_t f0 = 0; f0 < n_frames; f0 += F) {
    uint32_t nf = (uint32_t)std::min<uint64_t>(F, n_frames - f0);
    if (t.nseg) launch_pass(t, F, fb + f0 * npix * esz, dtype, nf, var, precision, sc.part.p, s);
    uint64_t work = (uint64_t)t.nbins * nf;
    unsigned blocks = blocks_for(work, 256);
    double* I = intensity + f0 * t.nbins;
    double* S = var ? sigma + f0 * t.nbins : nullptr;
    bool plain = precision == DIO_F32;
    if (plain && var)
      k_tfinal<true, true><<<blocks, 256, 0, s>>>(sc.part.p, t.nseg, F, t.bin_ptr.p, t.bin_seg.p,
                                                  t.den, t.nbins, nf, I, S);
    else if (plain)
      k_tfinal<true, false><<<blocks, 256, 0, s>>>(sc.part.p, t.nseg, F, t.bin_ptr.p,
                                                   t.bin_seg.p, t.den, t.nbins, nf, I, S);
    else if (var)
      k_tfinal<false, true><<<blocks, 256, 0, s>>>(sc.part.p, t.nseg, F, t.bin_ptr.p,
                                                   t.bin_seg.p, t.den, t.nbins, nf, I, S);
    else
      k_tfinal<false, false><<<blocks, 256, 0, s>>>(sc.part.p, t.nseg, F, t.bin_ptr.p,
                                                    t.bin_seg.p, t.den, t.nbins, nf, I, S);
    DIO_HIP(hipGetLastError());
  }
}

}  // namespace dio
