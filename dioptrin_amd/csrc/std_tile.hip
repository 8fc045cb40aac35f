// Azimuthal spread on the detector-tiled table (gfx950): the batched form of
// azimuthal.rs:48-118 (integrate_with_columns).
//
// The bin rows of the reference are read here through the tiled execution
// table (tile.hpp): a tile's frames, its pixel normalization and its table
// are staged in LDS once, and every (segment, frame) pair is evaluated from
// LDS by one lane. A segment is a piece of one bin row (pixels ascending), so
// a bin's statistics are the merge of its segments' statistics:
//  - per segment and frame, over the valid contributions (the reference's
//    filter: s, n finite, n > 0, c finite, c > 0, x = s / n finite,
//    w = c * n finite and > 0), one shifted pass with the segment's first
//    valid x as the shift o: W = sum w, A = sum w (x - o), B = sum w (x - o)^2,
//    sum c s and sum c^2 max(s, 0); the segment mean is o + A / W and its
//    second moment about the mean B - A^2 / W (the shift keeps the
//    cancellation at the scale of the spread, as the reference's origin does);
//  - per bin, the segments are merged in tile order with Chan's pairwise
//    update (mean and second moment of a union), by 64 segment lanes and a
//    fixed tree, so a frame's result depends only on the table.
// I = sum c s / W, sigma = sqrt(max(sum c^2 v, 0)) / W and std =
// sqrt(max(M2 / W, 0)); a bin without a valid contribution gives (0, 0, NaN).
// FP64 arithmetic on float32 weights; frames are read as float64 (exact for
// every pixel type). x is s times the pixel's reciprocal normalization (1 ulp
// from s / n; the filter differs from the reference only where |s / n| is
// within that of the largest double).
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

// LDS: frames [kStdFrames][kLdsStride] f64, then per pixel (n, 1 / n) as an
// f64 pair, then a table chunk (index words, weights).
constexpr size_t kStdX = 0;
constexpr size_t kStdNR = kStdX + (size_t)kStdFrames * kLdsStride * 8;
constexpr size_t kStdTL = kStdNR + (size_t)kLdsStride * 16;
constexpr size_t kStdTW = kStdTL + (size_t)kStdCap * 8;
constexpr size_t kStdLds = kStdTW + (size_t)kStdCap * 16;
static_assert(kStdLds <= 160 * 1024, "LDS plan exceeds one CU");
static_assert(kStdCap >= kTilePixels / kGroup, "a whole segment must fit one table chunk");
static_assert(kStdNR % 16 == 0 && kStdTL % 16 == 0, "16-byte LDS rows");

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
  float* pw;
  float* pcs;
  float* pm2;
  float* pv;  // null without errors
  double* pmean;
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
  const uint64_t det = in ? (uint64_t)r * a.dim2 + c : 0;
  const uint64_t mem = in ? (uint64_t)(a.flip ? a.dim1 - 1 - r : r) * a.dim2 + c : 0;
  const P* fr = static_cast<const P*>(a.frames);
  const uint32_t last = m.y - m.x - 1;
#pragma unroll
  for (int k = 0; k < kStdR; k++) {
    const uint32_t i = m.x + min(seg_of_slot<kStdSlots>(slot, (uint32_t)k), last);
    g.lo[k] = a.seg_group[i];
    g.hi[k] = a.seg_group[i + 1];
    g.rank[k] = a.seg_rank[i];
  }
#pragma unroll
  for (uint32_t f = 0; f < kStdFrames; f++)
    g.x[f] = fr[(uint64_t)(f < a.nf ? f : 0) * a.npix + (f < a.nf ? mem : 0)];
  g.n = a.norm[det];
  const uint32_t ng = min(m.w - m.z, kStdCap);
  const uint32_t gi = m.z + min(threadIdx.x, ng - 1);
  g.li = reinterpret_cast<const u32x2*>(a.lidx)[gi];
  g.w = reinterpret_cast<const f32x4*>(a.w32)[gi];
}

template <class P>
__device__ __forceinline__ void std_store(const StdRegs<P>& g, uint32_t ng, unsigned char* lds) {
  double* X = reinterpret_cast<double*>(lds + kStdX);
  f64x2* NR = reinterpret_cast<f64x2*>(lds + kStdNR);
#pragma unroll
  for (uint32_t f = 0; f < kStdFrames; f++) X[f * kLdsStride + threadIdx.x] = (double)g.x[f];
  // A pixel whose normalization is not a positive finite number with a
  // finite reciprocal is stored as (0, 0): its w = c * 0 then fails the
  // positive test, as the reference's n checks do.
  const double rn = 1.0 / g.n;
  const bool pix = isfinite(g.n) && g.n > 0.0 && isfinite(rn);
  const f64x2 nr = {pix ? g.n : 0.0, pix ? rn : 0.0};
  NR[threadIdx.x] = nr;
  if (threadIdx.x < kStdFrames) X[threadIdx.x * kLdsStride + kPadSlot] = 0.0;
  if (threadIdx.x == 0) {
    const f64x2 z = {0.0, 0.0};
    NR[kPadSlot] = z;
  }
  if (threadIdx.x < ng) {
    reinterpret_cast<u32x2*>(lds + kStdTL)[threadIdx.x] = g.li;
    reinterpret_cast<f32x4*>(lds + kStdTW)[threadIdx.x] = g.w;
  }
}

// Table groups [gc, gc + n) (n <= kStdCap) into the LDS chunk (later chunks
// of a large tile; waits for its loads).
__device__ __forceinline__ void std_table(const StdArgs& a, uint32_t gc, uint32_t n,
                                          unsigned char* lds) {
  const uint32_t i = gc + min(threadIdx.x, n - 1);
  const u32x2 li = reinterpret_cast<const u32x2*>(a.lidx)[i];
  const f32x4 wi = reinterpret_cast<const f32x4*>(a.w32)[i];
  if (threadIdx.x < n) {
    reinterpret_cast<u32x2*>(lds + kStdTL)[threadIdx.x] = li;
    reinterpret_cast<f32x4*>(lds + kStdTW)[threadIdx.x] = wi;
  }
}

// Statistics of a part of a segment for one frame.
struct StdPart {
  double W, mean, m2, cs, v;
};

// Chunk groups [ga, gb) of frame q: one shifted pass over the valid
// contributions (shift: the part's first valid x).
__device__ __forceinline__ StdPart std_part(const unsigned char* lds, uint32_t ga, uint32_t gb,
                                            uint32_t q) {
  const double* X = reinterpret_cast<const double*>(lds + kStdX) + q * kLdsStride;
  const f64x2* NR = reinterpret_cast<const f64x2*>(lds + kStdNR);
  const u32x2* tl = reinterpret_cast<const u32x2*>(lds + kStdTL);
  const f32x4* tw = reinterpret_cast<const f32x4*>(lds + kStdTW);
  double W = 0.0, A = 0.0, B = 0.0, cs = 0.0, v = 0.0, o = 0.0;
  bool found = false;
  for (uint32_t g = ga; g < gb; g++) {
    const u32x2 lw = tl[g];
    const f32x4 wv = tw[g];
    const uint32_t l[4] = {lw.x & 0xffffu, lw.x >> 16, lw.y & 0xffffu, lw.y >> 16};
    const float cw[4] = {wv.x, wv.y, wv.z, wv.w};
    double sg[4];
    f64x2 nr[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
      sg[e] = X[l[e]];
      nr[e] = NR[l[e]];
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      // The reference's filter: with n and 1 / n positive and finite (or
      // both stored as 0), "s, n, c finite, n, c > 0, x finite, w finite and
      // > 0" is "x finite and w positive and finite" (an infinite or NaN s
      // makes x so; c <= 0, infinite or NaN makes w so).
      const double c = (double)cw[e];
      const double x = sg[e] * nr[e].y, w = c * nr[e].x;
      const bool ok = __builtin_isfinite(x) & (w > 0.0) & (w <= DBL_MAX);
      o = (ok & !found) ? x : o;
      found = found | ok;
      if (ok) {
        const double d = x - o, t = w * d;
        W = W + w;
        A = A + t;
        B = fma(t, d, B);
        cs = fma(c, sg[e], cs);
        v = fma(c * c, fmax(sg[e], 0.0), v);
      }
    }
  }
  StdPart p{W, 0.0, 0.0, cs, v};
  if (found) {
    const double ma = A / W;
    p.mean = o + ma;
    p.m2 = fmax(B - A * ma, 0.0);
  }
  return p;
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

// Segment s (chunk groups [ga, gb)), frame q: lane half h takes the first
// (h = 0) or second half of the groups; the halves are merged (first, second)
// across the lane pair and the first lane writes the segment's row (when
// write). Every lane of the workgroup calls this in each round.
__device__ __forceinline__ void std_segment(const StdArgs& a, const unsigned char* lds,
                                            uint32_t ga, uint32_t gb, uint32_t h, uint32_t q,
                                            uint32_t row, bool write) {
  const uint32_t gm = ga + (gb - ga + 1) / 2;
  StdPart p = write ? std_part(lds, h ? gm : ga, h ? gb : gm, q) : StdPart{0, 0, 0, 0, 0};
  const StdPart o{__shfl_xor(p.W, (int)kStdFrames), __shfl_xor(p.mean, (int)kStdFrames),
                  __shfl_xor(p.m2, (int)kStdFrames), __shfl_xor(p.cs, (int)kStdFrames),
                  __shfl_xor(p.v, (int)kStdFrames)};
  if (!write || h) return;
  chan_merge(p.W, p.mean, p.m2, o.W, o.mean, o.m2);
  const uint64_t at = (uint64_t)row * a.nf + q;
  a.pw[at] = (float)p.W;
  a.pcs[at] = (float)(p.cs + o.cs);
  a.pm2[at] = (float)p.m2;
  if (a.pv) a.pv[at] = (float)(p.v + o.v);
  a.pmean[at] = p.mean;
}

// Persistent: workgroup b walks tiles b, b + grid, ... (XCD-grouped as the
// tile kernels do), loading the next tile into registers while it computes
// the current one. Lane (slot, half, frame) evaluates half of the slot's
// segments (snake order over rounds of kStdSlots segments, longest first).
template <class P>
__global__ __launch_bounds__(kStdThreads) void k_tstd(StdArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const uint32_t slot = threadIdx.x / kStdLanes;
  const uint32_t h = (threadIdx.x / kStdFrames) % 2, q = threadIdx.x % kStdFrames;
  const uint32_t grid = gridDim.x;
  CQuad* meta = (CQuad*)(a.tile_meta);
  CWord* sgr = (CWord*)(a.seg_group);
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
  if (tn < a.ntiles) std_prefetch<P>(a, tn, make_uint4(mn.x, mn.y, mn.z, mn.w), slot, g);
  while (tn < a.ntiles) {
    const uint32_t t = tn;
    const uint4 m = make_uint4(mn.x, mn.y, mn.z, mn.w);
    __syncthreads();  // the previous tile is done with the LDS
    std_store<P>(g, min(m.w - m.z, kStdCap), lds);
    uint32_t lo[kStdR], hi[kStdR], rk[kStdR];
#pragma unroll
    for (int k = 0; k < kStdR; k++) {
      lo[k] = g.lo[k];
      hi[k] = g.hi[k];
      rk[k] = g.rank[k];
    }
    __syncthreads();
    tn = next_tile(t + grid, mn);
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
        std_table(a, gc, sgr[sb] - gc, lds);
        __syncthreads();
      }
      const uint32_t rounds = (sb - sa + kStdSlots - 1) / kStdSlots;
      for (uint32_t r = 0; r < rounds; r++) {
        const uint32_t s = sa + seg_of_slot<kStdSlots>(slot, r);
        const bool has = s < sb && q < a.nf;
        uint32_t ga = 0, gb = 0, row = 0;
        if (sa == s0 && r < (uint32_t)kStdR) {
#pragma unroll
          for (int k = 0; k < kStdR; k++)
            if (r == (uint32_t)k) {
              ga = lo[k];
              gb = hi[k];
              row = rk[k];
            }
        } else if (has) {
          ga = a.seg_group[s];
          gb = a.seg_group[s + 1];
          row = a.seg_rank[s];
        }
        std_segment(a, lds, ga - gc, gb - gc, h, q, row, has);
      }
      sa = sb;
      if (sa < s1) gc = sgr[sa];
    }
  }
}

constexpr unsigned kStdFinalLanes = 64;
constexpr unsigned kStdFinalWaves = 4;

// One wavefront per bin: segment lane l merges the bin's rows l, l + 64, ...
// (tile order, contiguous in bin order) for every frame of the pass, then the
// 64 lane results are merged by a fixed pairwise tree (l with l ^ 1, then
// l ^ 2, ... l ^ 32).
__global__ __launch_bounds__(kStdFinalLanes * kStdFinalWaves) void k_tstd_final(
    const float* pw, const float* pcs, const float* pm2, const float* pv, const double* pmean,
    uint32_t nf, const uint32_t* bin_ptr, uint32_t nbins, double* I, double* S, double* SD) {
  const uint32_t b = blockIdx.x * kStdFinalWaves + threadIdx.x / kStdFinalLanes;
  const uint32_t l = threadIdx.x % kStdFinalLanes;
  if (b >= nbins) return;
  const uint32_t k0 = bin_ptr[b], k1 = bin_ptr[b + 1];
  double W[kStdFrames], M[kStdFrames], M2[kStdFrames], C[kStdFrames], V[kStdFrames];
#pragma unroll
  for (uint32_t j = 0; j < kStdFrames; j++) W[j] = M[j] = M2[j] = C[j] = V[j] = 0.0;
  for (uint32_t k = k0 + l; k < k1; k += kStdFinalLanes) {
    const uint64_t at = (uint64_t)k * nf;
    float w[kStdFrames], m2[kStdFrames], cs[kStdFrames], v[kStdFrames];
    double mn[kStdFrames];
    if (nf == kStdFrames) {  // whole rows: 16-byte loads (rows start 32-byte aligned)
#pragma unroll
      for (uint32_t j = 0; j < kStdFrames; j += 4) {
        const f32x4 a = *reinterpret_cast<const f32x4*>(pw + at + j);
        const f32x4 b = *reinterpret_cast<const f32x4*>(pm2 + at + j);
        const f32x4 c = *reinterpret_cast<const f32x4*>(pcs + at + j);
        const f32x4 d = pv ? *reinterpret_cast<const f32x4*>(pv + at + j) : f32x4{0, 0, 0, 0};
        const f64x2 e = *reinterpret_cast<const f64x2*>(pmean + at + j);
        const f64x2 f = *reinterpret_cast<const f64x2*>(pmean + at + j + 2);
        w[j] = a.x, w[j + 1] = a.y, w[j + 2] = a.z, w[j + 3] = a.w;
        m2[j] = b.x, m2[j + 1] = b.y, m2[j + 2] = b.z, m2[j + 3] = b.w;
        cs[j] = c.x, cs[j + 1] = c.y, cs[j + 2] = c.z, cs[j + 3] = c.w;
        v[j] = d.x, v[j + 1] = d.y, v[j + 2] = d.z, v[j + 3] = d.w;
        mn[j] = e.x, mn[j + 1] = e.y, mn[j + 2] = f.x, mn[j + 3] = f.y;
      }
    } else {
#pragma unroll
      for (uint32_t j = 0; j < kStdFrames; j++) {
        const bool in = j < nf;
        w[j] = in ? pw[at + j] : 0.f;
        m2[j] = in ? pm2[at + j] : 0.f;
        cs[j] = in ? pcs[at + j] : 0.f;
        v[j] = in && pv ? pv[at + j] : 0.f;
        mn[j] = in ? pmean[at + j] : 0.0;
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < kStdFrames; j++) {
      chan_merge(W[j], M[j], M2[j], (double)w[j], mn[j], (double)m2[j]);
      C[j] = C[j] + (double)cs[j];
      V[j] = V[j] + (double)v[j];
    }
  }
#pragma unroll
  for (uint32_t off = 1; off < kStdFinalLanes; off <<= 1) {
#pragma unroll
    for (uint32_t j = 0; j < kStdFrames; j++) {
      const double wb = __shfl_xor(W[j], (int)off, kStdFinalLanes);
      const double mb = __shfl_xor(M[j], (int)off, kStdFinalLanes);
      const double m2b = __shfl_xor(M2[j], (int)off, kStdFinalLanes);
      const double cb = __shfl_xor(C[j], (int)off, kStdFinalLanes);
      const double vb = __shfl_xor(V[j], (int)off, kStdFinalLanes);
      // The lower lane of each pair is the first operand, so both lanes of a
      // pair hold the same bits.
      if (l & off) {
        double w2 = wb, mm = mb, mm2 = m2b;
        chan_merge(w2, mm, mm2, W[j], M[j], M2[j]);
        W[j] = w2;
        M[j] = mm;
        M2[j] = mm2;
        C[j] = cb + C[j];
        V[j] = vb + V[j];
      } else {
        chan_merge(W[j], M[j], M2[j], wb, mb, m2b);
        C[j] = C[j] + cb;
        V[j] = V[j] + vb;
      }
    }
  }
  if (l < nf) {
    double w = 0, c = 0, v = 0, m2 = 0;
#pragma unroll
    for (uint32_t j = 0; j < kStdFrames; j++)
      if (j == l) {
        w = W[j];
        c = C[j];
        v = V[j];
        m2 = M2[j];
      }
    const uint64_t o = (uint64_t)l * nbins + b;
    if (w > 0.0) {
      I[o] = c / w;
      if (S) S[o] = sqrt(fmax(v, 0.0)) / w;
      SD[o] = sqrt(fmax(m2 / w, 0.0));
    } else {
      I[o] = 0.0;
      if (S) S[o] = 0.0;
      SD[o] = NAN;
    }
  }
}

}  // namespace

size_t std_pass_bytes(const TiledTable& t, uint32_t frames, bool errors) {
  return (size_t)t.nseg * frames * ((errors ? 4 : 3) * sizeof(float) + sizeof(double)) + 256;
}

uint32_t std_frames_per_pass() { return kStdFrames; }

void azimuthal_std_tiled(const TiledTable& t, const double* norm, void* scratch,
                         size_t scratch_bytes, const void* frames, int dtype, uint64_t n_frames,
                         bool errors, double* intensity, double* sigma, double* std_out,
                         hipStream_t s) {
  if (n_frames == 0) return;
  const size_t esz = dtype_bytes(dtype);
  const uint64_t npix = (uint64_t)t.dim1 * t.dim2;
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
    const void* ks[] = {reinterpret_cast<const void*>(&k_tstd<uint16_t>),
                        reinterpret_cast<const void*>(&k_tstd<uint32_t>),
                        reinterpret_cast<const void*>(&k_tstd<int32_t>),
                        reinterpret_cast<const void*>(&k_tstd<float>),
                        reinterpret_cast<const void*>(&k_tstd<double>)};
    for (const void* k : ks)
      DIO_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kStdLds));
    done.fetch_or(bit);
  }
  const uint32_t grid = std::min<uint32_t>(t.ntiles, t.max_wgs ? std::min(t.cus, t.max_wgs) : t.cus);
  for (uint64_t f0 = 0; f0 < n_frames; f0 += F) {
    const uint32_t nf = (uint32_t)std::min<uint64_t>(F, n_frames - f0);
    // Planes of nf-frame rows: mean (f64) first, then W, sum c s, M2 and
    // (errors) sum c^2 v as float32.
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
    a.pmean = reinterpret_cast<double*>(base);
    a.pw = reinterpret_cast<float*>(base + rows * 8);
    a.pcs = a.pw + rows;
    a.pm2 = a.pcs + rows;
    a.pv = errors ? a.pm2 + rows : nullptr;
    double* I = intensity + f0 * t.nbins;
    double* S = errors ? sigma + f0 * t.nbins : nullptr;
    double* D = std_out + f0 * t.nbins;
    if (t.nseg) {
      switch (dtype) {
        case DIO_U16: k_tstd<uint16_t><<<grid, kStdThreads, kStdLds, s>>>(a); break;
        case DIO_U32: k_tstd<uint32_t><<<grid, kStdThreads, kStdLds, s>>>(a); break;
        case DIO_I32: k_tstd<int32_t><<<grid, kStdThreads, kStdLds, s>>>(a); break;
        case DIO_F32T: k_tstd<float><<<grid, kStdThreads, kStdLds, s>>>(a); break;
        case DIO_F64T: k_tstd<double><<<grid, kStdThreads, kStdLds, s>>>(a); break;
        default: fail(DIO_ERR_INVALID, "unknown frame dtype");
      }
    }
    k_tstd_final<<<(t.nbins + kStdFinalWaves - 1) / kStdFinalWaves,
                   kStdFinalLanes * kStdFinalWaves, 0, s>>>(a.pw, a.pcs, a.pm2, a.pv, a.pmean, nf,
                                                            t.bin_ptr.p, t.nbins, I, S, D);
    DIO_HIP(hipGetLastError());
  }
}

}  // namespace dio
