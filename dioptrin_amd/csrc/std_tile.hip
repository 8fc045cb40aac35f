// Azimuthal spread on the detector-tiled table (gfx950): the batched form of
// azimuthal.rs:48-118 (integrate_with_columns).
//
// The bin rows of the reference are read here through the tiled execution
// table (tile.hpp): a tile's frames and its table are staged in LDS once, and
// every (segment, frame) pair is evaluated from LDS by one lane. A segment is
// a piece of one bin row (pixels ascending), so a bin's statistics are the
// merge of its segments' statistics.
//
// The reference's filter (s, n finite, n > 0, c finite, c > 0, x = s / n
// finite, w = c * n finite and > 0) is split by what it depends on:
//  - per pixel and frame, staged with the frame: x = s * (1 / n), or NaN
//    where n is not a positive finite number with a finite reciprocal or x
//    is not finite;
//  - per entry, once per tile: w = c * n, and an entry whose c or w fails the
//    test is pointed at the tile's NaN slot (table padding at its zero slot,
//    with w = 0).
// An entry then counts for a frame exactly when its staged x is finite.
// Per segment half and frame: one shifted pass with the part's first valid x
// as the shift o (a short scan finds it first): W = sum w, A = sum w (x - o),
// B = sum w (x - o)^2; mean o + A / W, second moment B - A^2 / W (the shift
// keeps the cancellation at the scale of the spread, as the reference's
// origin does). Entries that do not count add exact zeros. The sums are
// first taken without the validity test (padding adds exact zeros as it is);
// a part that meets a NaN x gets non-finite A or B and is summed again with
// the test, so both give the same bits.
// Per bin the segments are merged in bin order by a fixed scheme, so a
// frame's result depends only on the table.
// I = mean (sum w x / sum w, which is the reference's sum c s / sum w),
// sigma = sqrt(max(sum c^2 max(s, 0), 0)) / W with c^2 s taken as c w x, and
// std = sqrt(max(M2 / W, 0)); a bin without a valid contribution gives
// (0, 0, NaN). FP64 arithmetic on float32 weights; frames are read as float64
// (exact for every pixel type). x is s times the pixel's reciprocal
// normalization (1 ulp from s / n; the filter differs from the reference only
// where |s / n| is within that of the largest double).
#include <algorithm>
#include <atomic>
#include <cfloat>

#include "gather.hpp"
#include "tile_impl.hpp"

namespace dio {

namespace {

using namespace tk;

constexpr unsigned kStdThreads = 1024;
constexpr uint32_t kStdFrames = 8;                       // frames per pass
constexpr uint32_t kStdLanes = 2 * kStdFrames;           // a segment's lanes: 2 halves x frames
constexpr uint32_t kStdSlots = kStdThreads / kStdLanes;  // segments evaluated at once
constexpr int kStdR = 2;                                 // rounds whose bounds are prefetched
constexpr uint32_t kStdCap = 1024;                       // table groups per LDS chunk
static_assert(kStdCap == kStdThreads, "one table group per thread");
static_assert(kTilePixels == kStdThreads, "a thread stages one pixel");

// LDS: x of the pass's frames [kStdFrames][kLdsStride] f64 (slot kPadSlot 0
// for table padding, slot kStdNanSlot NaN for entries that never count), the
// tile's normalization [kLdsStride] f64, then a table chunk: entry byte
// offsets into a frame row (u16, 4 per group), w = c n (f64) and, with
// errors, c w (f64).
constexpr uint32_t kStdNanSlot = kPadSlot + 1;
static_assert(kStdNanSlot < kLdsStride, "NaN slot inside the row");
constexpr size_t kStdX = 0;
constexpr size_t kStdNN = kStdX + (size_t)kStdFrames * kLdsStride * 8;
constexpr size_t kStdTL = kStdNN + (size_t)kLdsStride * 8;
constexpr size_t kStdTW = kStdTL + (size_t)kStdCap * 8;
constexpr size_t kStdTC = kStdTW + (size_t)kStdCap * 32;
constexpr size_t kStdLdsPlain = kStdTC;
constexpr size_t kStdLdsVar = kStdTC + (size_t)kStdCap * 32;
static_assert(kStdLdsVar <= 160 * 1024, "LDS plan exceeds one CU");
static_assert(kStdCap >= kTilePixels / kGroup, "a whole segment must fit one table chunk");
static_assert(kStdNN % 16 == 0 && kStdTL % 16 == 0 && kStdTW % 16 == 0 && kStdTC % 16 == 0,
              "16-byte LDS rows");
static_assert(kStdNanSlot * 8 < 65536, "byte offsets fit 16 bits");

template <bool VAR> constexpr size_t std_lds() { return VAR ? kStdLdsVar : kStdLdsPlain; }

struct StdArgs {
  const void* frames;
  uint32_t nf;  // frames of this pass (<= kStdFrames); partial rows hold nf values
  uint64_t npix;
  uint32_t dim1, dim2, ntx, ntiles, flip;
  const double* norm;  // detector order
  const uint16_t* lidx;
  const float* w32;
  const uint32_t* seg_group;
  const uint32_t* tile_meta;
  const uint32_t* seg_rank;  // a segment's row in bin order
  // Per-segment partials, [bin-order row][nf] each.
  double* pw;
  double* pmean;
  float* pm2;
  float* pv;  // errors only
};

typedef const __attribute__((address_space(4))) uint32_t CWord;
typedef const __attribute__((address_space(4))) u32x4 CQuad;

// What one thread loads for a tile ahead of its use: its pixel of the pass's
// frames and its normalization, one table group of the first chunk, and the
// bounds and bin-order rows of its slot's segments in the first kStdR rounds.
template <class P> struct StdRegs {
  P x[kStdFrames];
  double n;
  u32x2 li;
  f32x4 w;
  uint32_t lo[kStdR], hi[kStdR], rank[kStdR];
};

// Byte offset i * sz into a buffer (offsets below 4 GiB; the host checks).
template <class T> __device__ __forceinline__ const T& at32(const void* base, uint32_t off) {
  return *reinterpret_cast<const T*>(static_cast<const char*>(base) + off);
}

template <class P>
__device__ __forceinline__ void std_prefetch(const StdArgs& a, uint32_t t, uint4 m,
                                             uint32_t slot, StdRegs<P>& g) {
  // Pixel tid of the tile: row tid / 64, column tid % 64. Pixels outside the
  // detector and frames past nf read pixel 0 of the pass's first frame (a
  // valid address; no table entry points at such a pixel, no partial of
  // such a frame is written).
  const uint32_t r = (t / a.ntx) * kTileRows + threadIdx.x / kTileCols;
  const uint32_t c = (t % a.ntx) * kTileCols + threadIdx.x % kTileCols;
  const bool in = r < a.dim1 && c < a.dim2;
  const uint32_t det = in ? r * a.dim2 + c : 0u;
  const uint32_t mem = in ? (a.flip ? a.dim1 - 1 - r : r) * a.dim2 + c : 0u;
  const uint32_t last = m.y - m.x - 1;
#pragma unroll
  for (int k = 0; k < kStdR; k++) {
    const uint32_t i = m.x + min(seg_of_slot<kStdSlots>(slot, (uint32_t)k), last);
    g.lo[k] = at32<uint32_t>(a.seg_group, i * 4u);
    g.hi[k] = at32<uint32_t>(a.seg_group, i * 4u + 4u);
    g.rank[k] = at32<uint32_t>(a.seg_rank, i * 4u);
  }
  // Frame f at byte offset f * npix * sizeof(P) (a pass's frames span less
  // than 4 GiB; frames past nf read frame 0).
  const uint32_t fstride = (uint32_t)(a.npix * sizeof(P));
#pragma unroll
  for (uint32_t f = 0; f < kStdFrames; f++)
    g.x[f] = at32<P>(a.frames, mem * (uint32_t)sizeof(P) + (f < a.nf ? f * fstride : 0u));
  g.n = at32<double>(a.norm, det * 8u);
  const uint32_t ng = min(m.w - m.z, kStdCap);
  const uint32_t gi = m.z + min(threadIdx.x, ng - 1);
  g.li = at32<u32x2>(a.lidx, gi * 8u);
  g.w = at32<f32x4>(a.w32, gi * 16u);
}

// The thread's pixel of each frame as x = s / n (NaN where it does not
// count), and the zero and NaN slots; the tile's normalization must already
// be in LDS.
template <class P>
__device__ __forceinline__ void std_store(const StdRegs<P>& g, unsigned char* lds) {
  double* X = reinterpret_cast<double*>(lds + kStdX);
  const double rn = 1.0 / g.n;
  const bool pix = isfinite(g.n) && g.n > 0.0 && isfinite(rn);
#pragma unroll
  for (uint32_t f = 0; f < kStdFrames; f++) {
    const double x = (double)g.x[f] * rn;
    X[f * kLdsStride + threadIdx.x] = pix && isfinite(x) ? x : __builtin_nan("");
  }
  if (threadIdx.x < kStdFrames) {
    X[threadIdx.x * kLdsStride + kPadSlot] = 0.0;
    X[threadIdx.x * kLdsStride + kStdNanSlot] = __builtin_nan("");
  }
}

// Table group i of the chunk from its index word pair and weights: w = c n
// per entry from the tile's normalization in LDS; entries failing the
// reference's c and w tests point at the NaN slot, table padding at the zero
// slot, both with w = c w = 0.
template <bool VAR>
__device__ __forceinline__ void std_prep(unsigned char* lds, uint32_t i, u32x2 li, f32x4 cw) {
  const double* NN = reinterpret_cast<const double*>(lds + kStdNN);
  const uint32_t l[4] = {li.x & 0xffffu, li.x >> 16, li.y & 0xffffu, li.y >> 16};
  const float c[4] = {cw.x, cw.y, cw.z, cw.w};
  uint32_t off[4];
  double w[4], cwv[4];
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const double cd = (double)c[e];
    const double we = cd * NN[l[e]];
    const bool ok = (cd > 0.0) & (we > 0.0) & (we <= DBL_MAX);
    off[e] = (ok ? l[e] : l[e] == kPadSlot ? kPadSlot : kStdNanSlot) * 8u;
    w[e] = ok ? we : 0.0;
    cwv[e] = ok ? cd * we : 0.0;
  }
  reinterpret_cast<u32x2*>(lds + kStdTL)[i] = u32x2{off[0] | off[1] << 16, off[2] | off[3] << 16};
  f64x2* TW = reinterpret_cast<f64x2*>(lds + kStdTW);
  TW[2 * i] = f64x2{w[0], w[1]};
  TW[2 * i + 1] = f64x2{w[2], w[3]};
  if constexpr (VAR) {
    f64x2* TC = reinterpret_cast<f64x2*>(lds + kStdTC);
    TC[2 * i] = f64x2{cwv[0], cwv[1]};
    TC[2 * i + 1] = f64x2{cwv[2], cwv[3]};
  }
}

// Table groups [gc, gc + n) (n <= kStdCap) into the LDS chunk (later chunks
// of a large tile; waits for its loads).
template <bool VAR>
__device__ __forceinline__ void std_table(const StdArgs& a, uint32_t gc, uint32_t n,
                                          unsigned char* lds) {
  const uint32_t i = gc + min(threadIdx.x, n - 1);
  const u32x2 li = at32<u32x2>(a.lidx, i * 8u);
  const f32x4 wi = at32<f32x4>(a.w32, i * 16u);
  if (threadIdx.x < n) std_prep<VAR>(lds, threadIdx.x, li, wi);
}

// The first valid x of chunk groups [ga, gb) for frame q (0 if none): the
// shift of a segment's sums.
__device__ __forceinline__ double std_shift(const unsigned char* lds, uint32_t ga, uint32_t gb,
                                            uint32_t q) {
  const unsigned char* Xq = lds + kStdX + (size_t)q * kLdsStride * 8;
  const uint16_t* t16 = reinterpret_cast<const uint16_t*>(lds + kStdTL);
  for (uint32_t e = ga * kGroup; e < gb * kGroup; e++) {
    const double x = *reinterpret_cast<const double*>(Xq + t16[e]);
    if (__builtin_isfinite(x)) return x;
  }
  return 0.0;
}

// Shifted sums over chunk groups [ga, gb) of frame q: W = sum w,
// A = sum w (x - o), B = sum w (x - o)^2 and (VAR) v = sum c w max(x, 0) over
// the entries that count; the others add exact zeros. Without CHECK every
// entry is taken as it is: the same sums where every x is finite (padding:
// x = 0, w = 0), non-finite A and B where one is NaN.
template <bool VAR, bool CHECK>
__device__ __forceinline__ void std_sums(const unsigned char* lds, uint32_t ga, uint32_t gb,
                                         uint32_t q, double o, double& W, double& A, double& B,
                                         double& v) {
  const unsigned char* Xq = lds + kStdX + (size_t)q * kLdsStride * 8;
  const u32x2* tl = reinterpret_cast<const u32x2*>(lds + kStdTL);
  const f64x2* tw = reinterpret_cast<const f64x2*>(lds + kStdTW);
  const f64x2* tc = reinterpret_cast<const f64x2*>(lds + kStdTC);
  for (uint32_t g = ga; g < gb; g++) {
    const u32x2 lw = tl[g];
    const f64x2 w01 = tw[2 * g], w23 = tw[2 * g + 1];
    const uint32_t off[4] = {lw.x & 0xffffu, lw.x >> 16, lw.y & 0xffffu, lw.y >> 16};
    const double w[4] = {w01.x, w01.y, w23.x, w23.y};
    double x[4];
#pragma unroll
    for (int e = 0; e < 4; e++) x[e] = *reinterpret_cast<const double*>(Xq + off[e]);
    double cw[4] = {0.0, 0.0, 0.0, 0.0};
    if constexpr (VAR) {
      const f64x2 c01 = tc[2 * g], c23 = tc[2 * g + 1];
      cw[0] = c01.x, cw[1] = c01.y, cw[2] = c23.x, cw[3] = c23.y;
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      if constexpr (CHECK) {
        const bool ok = __builtin_isfinite(x[e]);
        const double wv = ok ? w[e] : 0.0;
        const double d = (ok ? x[e] : o) - o, t = wv * d;
        W = W + wv;
        A = A + t;
        B = fma(t, d, B);
      } else {
        const double d = x[e] - o, t = w[e] * d;
        W = W + w[e];
        A = A + t;
        B = fma(t, d, B);
      }
      if constexpr (VAR) v = fma(cw[e], fmax(x[e], 0.0), v);
    }
  }
}

// Chan's update of (W, mean, M2) by the statistics of another part.
__device__ __forceinline__ void chan_merge(double& W, double& mean, double& m2, double Wb,
                                           double meanb, double m2b) {
  if (Wb > 0.0) {
    if (W > 0.0) {
      const double Wt = W + Wb, d = meanb - mean, fb = Wb / Wt;
      mean = mean + d * fb;
      m2 = m2 + m2b + d * d * W * fb;
      W = Wt;
    } else {
      W = Wb;
      mean = meanb;
      m2 = m2b;
    }
  }
}

// Segment s (chunk groups [ga, gb)), frame q: both lane halves take the
// segment's first valid x as the shift; half h sums the first (h = 0) or
// second half of the groups, the halves' sums are added across the lane pair
// and the first lane writes the segment's row (when write): W, mean
// o + A / W, M2 = B - A^2 / W and sum c^2 v. Every lane of the workgroup
// calls this in each round.
template <bool VAR>
__device__ __forceinline__ void std_segment(const StdArgs& a, const unsigned char* lds,
                                            uint32_t ga, uint32_t gb, uint32_t h, uint32_t q,
                                            uint32_t row, bool write) {
  const uint32_t gm = ga + (gb - ga + 1) / 2;
  double W = 0.0, A = 0.0, B = 0.0, v = 0.0, o = 0.0;
  if (write) {
    o = std_shift(lds, ga, gb, q);
    const uint32_t g0 = h ? gm : ga, g1 = h ? gb : gm;
    std_sums<VAR, false>(lds, g0, g1, q, o, W, A, B, v);
    if (!__builtin_isfinite(A) || !__builtin_isfinite(B)) {
      W = A = B = v = 0.0;
      std_sums<VAR, true>(lds, g0, g1, q, o, W, A, B, v);
    }
  }
  // Sums of the two halves (addition is commutative, so both lanes agree).
  W = W + __shfl_xor(W, (int)kStdFrames);
  A = A + __shfl_xor(A, (int)kStdFrames);
  B = B + __shfl_xor(B, (int)kStdFrames);
  if constexpr (VAR) v = v + __shfl_xor(v, (int)kStdFrames);
  if (!write || h) return;
  double mean = 0.0, m2 = 0.0;
  if (W > 0.0) {
    const double ma = A / W;
    mean = o + ma;
    m2 = fmax(B - A * ma, 0.0);
  }
  const uint32_t at = row * a.nf + q;
  *reinterpret_cast<double*>(reinterpret_cast<char*>(a.pw) + at * 8u) = W;
  *reinterpret_cast<double*>(reinterpret_cast<char*>(a.pmean) + at * 8u) = mean;
  *reinterpret_cast<float*>(reinterpret_cast<char*>(a.pm2) + at * 4u) = (float)m2;
  if constexpr (VAR) *reinterpret_cast<float*>(reinterpret_cast<char*>(a.pv) + at * 4u) = (float)v;
}

// Persistent: workgroup b walks tiles b, b + grid, ... (XCD-grouped as the
// tile kernels do), loading the next tile into registers while it computes
// the current one. Lane (slot, half, frame) evaluates half of the slot's
// segments (snake order over rounds of kStdSlots segments, longest first).
// The next tile's normalization goes to LDS when the current tile is done.
template <class P, bool VAR>
__global__ __launch_bounds__(kStdThreads) void k_tstd(StdArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const uint32_t slot = threadIdx.x / kStdLanes;
  const uint32_t h = (threadIdx.x / kStdFrames) % 2, q = threadIdx.x % kStdFrames;
  const uint32_t grid = gridDim.x;
  CQuad* meta = (CQuad*)(a.tile_meta);
  CWord* sgr = (CWord*)(a.seg_group);
  double* NN = reinterpret_cast<double*>(lds + kStdNN);
  const uint32_t t0 = first_tile();
  if (t0 >= a.ntiles) return;
  StdRegs<P> g;
  // The first tile with table entries at or after t (workgroup-uniform).
  auto next_tile = [&](uint32_t t, u32x4& m) {
    for (; t < a.ntiles; t += grid) {
      m = meta[t];
      if (m.x != m.y) return t;
    }
    return a.ntiles;
  };
  u32x4 mn;
  uint32_t tn = next_tile(t0, mn);
  if (tn < a.ntiles) {
    std_prefetch<P>(a, tn, make_uint4(mn.x, mn.y, mn.z, mn.w), slot, g);
    NN[threadIdx.x] = g.n;
  }
  if (threadIdx.x == 0) NN[kPadSlot] = 0.0;
  while (tn < a.ntiles) {
    const uint32_t t = tn;
    const uint4 m = make_uint4(mn.x, mn.y, mn.z, mn.w);
    // The following tile's bounds, loaded here so that the staging's waits
    // cover their latency.
    uint32_t tq = t + grid;
    u32x4 mq = mn;
    if (tq < a.ntiles) mq = meta[tq];
    __syncthreads();  // the previous tile is done with the LDS; NN holds tile t's
    std_store<P>(g, lds);
    if (threadIdx.x < min(m.w - m.z, kStdCap)) std_prep<VAR>(lds, threadIdx.x, g.li, g.w);
    uint32_t lo[kStdR], hi[kStdR], rk[kStdR];
#pragma unroll
    for (int k = 0; k < kStdR; k++) {
      lo[k] = g.lo[k];
      hi[k] = g.hi[k];
      rk[k] = g.rank[k];
    }
    __syncthreads();
    while (tq < a.ntiles && mq.x == mq.y) {  // tiles without entries
      tq += grid;
      if (tq < a.ntiles) mq = meta[tq];
    }
    tn = tq;
    mn = mq;
    if (tn < a.ntiles) std_prefetch<P>(a, tn, make_uint4(mn.x, mn.y, mn.z, mn.w), slot, g);
    const uint32_t s0 = m.x, s1 = m.y, g1 = m.w;
    uint32_t sa = s0, gc = m.z;
    while (sa < s1) {
      // Largest sb with seg_group[sb] - gc <= kStdCap (whole segments).
      uint32_t sb = s1;
      if (g1 - gc > kStdCap) {
        uint32_t l0 = sa + 1, l1 = s1;
        while (l0 < l1) {
          const uint32_t mid = (l0 + l1 + 1) / 2;
          if (sgr[mid] - gc <= kStdCap) l0 = mid;
          else l1 = mid - 1;
        }
        sb = l0;
      }
      if (sa != s0) {  // a later chunk of a large tile
        __syncthreads();
        std_table<VAR>(a, gc, sgr[sb] - gc, lds);
        __syncthreads();
      }
      const uint32_t rounds = (sb - sa + kStdSlots - 1) / kStdSlots;
      uint32_t r = 0;
      if (sa == s0) {
        // Rounds with prefetched bounds. They issue no global loads: a wait
        // for one would also wait for the next tile's prefetch.
#pragma unroll
        for (int k = 0; k < kStdR; k++)
          if ((uint32_t)k < rounds) {
            const uint32_t s = sa + seg_of_slot<kStdSlots>(slot, (uint32_t)k);
            std_segment<VAR>(a, lds, lo[k] - gc, hi[k] - gc, h, q, rk[k], s < sb && q < a.nf);
          }
        r = min(rounds, (uint32_t)kStdR);
      }
      for (; r < rounds; r++) {
        const uint32_t s = sa + seg_of_slot<kStdSlots>(slot, r);
        const bool has = s < sb && q < a.nf;
        const uint32_t ga = has ? a.seg_group[s] : gc, gb = has ? a.seg_group[s + 1] : gc;
        const uint32_t row = has ? a.seg_rank[s] : 0;
        std_segment<VAR>(a, lds, ga - gc, gb - gc, h, q, row, has);
      }
      sa = sb;
      if (sa < s1) gc = sgr[sa];
    }
    // Tile t's normalization is no longer read (its chunks are prepared).
    if (tn < a.ntiles) NN[threadIdx.x] = g.n;
  }
}

constexpr unsigned kStdFinalLanes = 64;
constexpr unsigned kStdFinalWaves = 4;                            // waves per bin
constexpr unsigned kStdFinalSub = kStdFinalLanes / kStdFrames;    // row lanes per frame and wave
constexpr unsigned kStdFinalRows = kStdFinalSub * kStdFinalWaves; // row lanes per frame
constexpr uint32_t kStdFinalBatch = 8;                            // rows loaded at once

// One workgroup of 4 wavefronts per bin, lane (j, q) = (row lane, frame) of
// wave v: row lane r = 8 v + j takes the bin's rows r, r + 32, ... (tile
// order, contiguous in bin order) with its first non-empty row's mean as the
// shift (W = sum W_i, A = sum W_i (m_i - o), B = sum W_i (m_i - o)^2;
// M2 = sum M2_i + B - A^2 / W). A wave's 8 row lanes of a frame are merged by
// a fixed pairwise tree (j with j ^ 1, then j ^ 2, j ^ 4; lower lane first)
// and the 4 wave results in wave order.
template <bool VAR>
__global__ __launch_bounds__(kStdFinalLanes * kStdFinalWaves) void k_tstd_final(
    const double* pw, const double* pmean, const float* pm2, const float* pv, uint32_t nf,
    const uint32_t* bin_ptr, uint32_t nbins, double* I, double* S, double* SD) {
  __shared__ f64x2 red[kStdFinalWaves][kStdFrames][2];
  const uint32_t b = blockIdx.x;
  const uint32_t wv = threadIdx.x / kStdFinalLanes, lane = threadIdx.x % kStdFinalLanes;
  const uint32_t j = lane / kStdFrames, q = lane % kStdFrames;
  const uint32_t k0 = bin_ptr[b], k1 = bin_ptr[b + 1];
  const bool act = q < nf;
  double W = 0.0, A = 0.0, B = 0.0, M2 = 0.0, V = 0.0, o = 0.0;
  bool found = false;
  auto add = [&](double wi, double mi, double m2i, double vi) {
    o = (!found && wi > 0.0) ? mi : o;
    found = found || wi > 0.0;
    // An empty row (W_i = 0, m_i = 0) adds exact zeros.
    const double d = mi - o, t = wi * d;
    W = W + wi;
    A = A + t;
    B = fma(t, d, B);
    M2 = M2 + m2i;
    V = V + vi;
  };
  // Rows in order; loads of kStdFinalBatch rows are issued before they are
  // added.
  uint32_t k = k0 + wv * kStdFinalSub + j;
  for (; k + (kStdFinalBatch - 1) * kStdFinalRows < k1; k += kStdFinalBatch * kStdFinalRows) {
    double wi[kStdFinalBatch], mi[kStdFinalBatch];
    float m2i[kStdFinalBatch], vi[kStdFinalBatch];
#pragma unroll
    for (uint32_t u = 0; u < kStdFinalBatch; u++) {
      const uint64_t at = act ? (uint64_t)(k + u * kStdFinalRows) * nf + q : 0;
      wi[u] = act ? pw[at] : 0.0;
      mi[u] = act ? pmean[at] : 0.0;
      m2i[u] = act ? pm2[at] : 0.f;
      vi[u] = VAR && act ? pv[at] : 0.f;
    }
#pragma unroll
    for (uint32_t u = 0; u < kStdFinalBatch; u++) add(wi[u], mi[u], m2i[u], vi[u]);
  }
  for (; k < k1; k += kStdFinalRows) {
    const uint64_t at = act ? (uint64_t)k * nf + q : 0;
    add(act ? pw[at] : 0.0, act ? pmean[at] : 0.0, act ? (double)pm2[at] : 0.0,
        VAR && act ? (double)pv[at] : 0.0);
  }
  double mean = 0.0;
  if (W > 0.0) {
    const double ma = A / W;
    mean = o + ma;
    M2 = M2 + fmax(B - A * ma, 0.0);
  }
#pragma unroll
  for (uint32_t off = kStdFrames; off < kStdFinalLanes; off <<= 1) {
    const double wb = __shfl_xor(W, (int)off, kStdFinalLanes);
    const double mb = __shfl_xor(mean, (int)off, kStdFinalLanes);
    const double m2b = __shfl_xor(M2, (int)off, kStdFinalLanes);
    const double vb = VAR ? __shfl_xor(V, (int)off, kStdFinalLanes) : 0.0;
    // The lower lane of each pair is the first operand, so both lanes of a
    // pair hold the same bits.
    if (lane & off) {
      double w2 = wb, mm = mb, mm2 = m2b;
      chan_merge(w2, mm, mm2, W, mean, M2);
      W = w2;
      mean = mm;
      M2 = mm2;
      V = vb + V;
    } else {
      chan_merge(W, mean, M2, wb, mb, m2b);
      V = V + vb;
    }
  }
  if (j == 0) {
    red[wv][q][0] = f64x2{W, mean};
    red[wv][q][1] = f64x2{M2, V};
  }
  __syncthreads();
  if (wv != 0 || j != 0 || !act) return;
  for (uint32_t u = 1; u < kStdFinalWaves; u++) {
    const f64x2 r0 = red[u][q][0], r1 = red[u][q][1];
    chan_merge(W, mean, M2, r0.x, r0.y, r1.x);
    V = V + r1.y;
  }
  const uint64_t o2 = (uint64_t)q * nbins + b;
  if (W > 0.0) {
    I[o2] = mean;
    if (S) S[o2] = sqrt(fmax(V, 0.0)) / W;
    SD[o2] = sqrt(fmax(M2 / W, 0.0));
  } else {
    I[o2] = 0.0;
    if (S) S[o2] = 0.0;
    SD[o2] = NAN;
  }
}

template <bool VAR> void std_attrs() {
  const void* ks[] = {reinterpret_cast<const void*>(&k_tstd<uint16_t, VAR>),
                      reinterpret_cast<const void*>(&k_tstd<uint32_t, VAR>),
                      reinterpret_cast<const void*>(&k_tstd<int32_t, VAR>),
                      reinterpret_cast<const void*>(&k_tstd<float, VAR>),
                      reinterpret_cast<const void*>(&k_tstd<double, VAR>)};
  for (const void* k : ks)
    DIO_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)std_lds<VAR>()));
}

template <bool VAR>
void std_launch(const StdArgs& a, int dtype, uint32_t grid, hipStream_t s) {
  constexpr size_t L = std_lds<VAR>();
  switch (dtype) {
    case DIO_U16: k_tstd<uint16_t, VAR><<<grid, kStdThreads, L, s>>>(a); break;
    case DIO_U32: k_tstd<uint32_t, VAR><<<grid, kStdThreads, L, s>>>(a); break;
    case DIO_I32: k_tstd<int32_t, VAR><<<grid, kStdThreads, L, s>>>(a); break;
    case DIO_F32T: k_tstd<float, VAR><<<grid, kStdThreads, L, s>>>(a); break;
    case DIO_F64T: k_tstd<double, VAR><<<grid, kStdThreads, L, s>>>(a); break;
    default: fail(DIO_ERR_INVALID, "unknown frame dtype");
  }
}

}  // namespace

size_t std_pass_bytes(const TiledTable& t, uint32_t frames, bool errors) {
  return (size_t)t.nseg * frames * (2 * sizeof(double) + (errors ? 2 : 1) * sizeof(float)) + 256;
}

uint32_t std_frames_per_pass() { return kStdFrames; }

void azimuthal_std_tiled(const TiledTable& t, const double* norm, void* scratch,
                         size_t scratch_bytes, const void* frames, int dtype, uint64_t n_frames,
                         bool errors, double* intensity, double* sigma, double* std_out,
                         hipStream_t s) {
  if (n_frames == 0) return;
  const size_t esz = dtype_bytes(dtype);
  const uint64_t npix = (uint64_t)t.dim1 * t.dim2;
  // The kernels address a frame, the normalization, the table and the
  // partials by 32-bit byte offsets.
  if (npix * kStdFrames * 8 >= (1ull << 32) || (uint64_t)t.ngroups * 16 >= (1ull << 32) ||
      (uint64_t)t.nseg * kStdFrames * 8 >= (1ull << 32))
    fail(DIO_ERR_INVALID, "azimuthal std: detector or table too large for 32-bit offsets");
  // Frames per pass: kStdFrames, fewer when the scratch holds fewer.
  uint32_t F = kStdFrames;
  while (F > 1 && std_pass_bytes(t, F, errors) > scratch_bytes) F--;
  if (std_pass_bytes(t, F, errors) > scratch_bytes)
    fail(DIO_ERR_INVALID, "azimuthal std scratch below one frame's partials");
  static std::atomic<uint64_t> done{0};
  int dev = 0;
  DIO_HIP(hipGetDevice(&dev));
  const uint64_t bit = 1ull << (dev & 63);
  if (!(done.load() & bit)) {
    std_attrs<false>();
    std_attrs<true>();
    done.fetch_or(bit);
  }
  const uint32_t grid = std::min<uint32_t>(t.ntiles, t.max_wgs ? std::min(t.cus, t.max_wgs) : t.cus);
  for (uint64_t f0 = 0; f0 < n_frames; f0 += F) {
    const uint32_t nf = (uint32_t)std::min<uint64_t>(F, n_frames - f0);
    // Planes of nf-frame rows: W and mean (f64), then M2 and (errors)
    // sum c^2 v as float32.
    char* base = static_cast<char*>(scratch);
    const size_t rows = (size_t)t.nseg * nf;  // planes start 16-byte aligned when nf = 8
    StdArgs a{};
    a.frames = static_cast<const char*>(frames) + f0 * npix * esz;
    a.nf = nf;
    a.npix = npix;
    a.dim1 = t.dim1;
    a.dim2 = t.dim2;
    a.ntx = t.ntx;
    a.ntiles = t.ntiles;
    a.flip = t.flip ? 1u : 0u;
    a.norm = norm;
    a.lidx = t.lidx.p;
    a.w32 = t.w32.p;
    a.seg_group = t.seg_group.p;
    a.tile_meta = t.tile_meta.p;
    a.seg_rank = t.seg_rank.p;
    a.pw = reinterpret_cast<double*>(base);
    a.pmean = a.pw + rows;
    a.pm2 = reinterpret_cast<float*>(a.pmean + rows);
    a.pv = errors ? a.pm2 + rows : nullptr;
    double* I = intensity + f0 * t.nbins;
    double* S = errors ? sigma + f0 * t.nbins : nullptr;
    double* D = std_out + f0 * t.nbins;
    if (t.nseg) {
      if (errors) std_launch<true>(a, dtype, grid, s);
      else std_launch<false>(a, dtype, grid, s);
    }
    const uint32_t blocks = t.nbins;
    if (blocks == 0) continue;
    if (errors)
      k_tstd_final<true><<<blocks, kStdFinalLanes * kStdFinalWaves, 0, s>>>(
          a.pw, a.pmean, a.pm2, a.pv, nf, t.bin_ptr.p, t.nbins, I, S, D);
    else
      k_tstd_final<false><<<blocks, kStdFinalLanes * kStdFinalWaves, 0, s>>>(
          a.pw, a.pmean, a.pm2, a.pv, nf, t.bin_ptr.p, t.nbins, I, S, D);
    DIO_HIP(hipGetLastError());
  }
}

}  // namespace dio
