// Shared device code of the tiled gather kernels (internal to the library).
#pragma once
#include <cfloat>
#include <cstdint>

#include "tile.hpp"

namespace dio {
namespace tk {

constexpr unsigned kXcds = 8;  // accelerator complex dies of an MI355X

// Native vector types for table words held in registers and LDS: their loads
// and stores are single instructions, where copies of HIP's vector structs
// can become byte copies through scratch memory.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// ---------------------------------------------------------------- accumulators

// Float32 fused multiply-add (CudaPrecision::F32).
struct TPlain {
  float s;
  __device__ void init() { s = 0.0f; }
  __device__ void add(float w, double, float x) { s = fmaf(w, x, s); }
  __device__ void add_var(float w, double, float x) { s = fmaf(w * w, fmaxf(x, 0.0f), s); }
  __device__ double value() const { return (double)s; }
};

// Compensated float32 (CudaPrecision::F32Compensated): the running sum hi is
// updated with Knuth's TwoSum; its rounding errors and the FMA residuals of
// the products (exact: w*x = p + pe) are accumulated in lo (Ogita-Rump-Oishi
// Sum2 with error-free products, error ~ u|s| + n^2 u^2 sum|p|). A non-finite
// hi is returned as is.
struct TComp {
  float hi, lo;
  __device__ void init() {
    hi = 0.0f;
    lo = 0.0f;
  }
  __device__ void two_sum(float p, float pe) {
    float s = hi + p;
    float bb = s - hi;
    float err = (hi - (s - bb)) + (p - bb);
    lo = lo + (err + pe);
    hi = s;
  }
  __device__ void add(float w, double, float x) {
    float p = w * x;
    two_sum(p, fmaf(w, x, -p));
  }
  __device__ void add_var(float w, double, float x) {
    x = fmaxf(x, 0.0f);
    float w2 = w * w;
    float w2l = fmaf(w, w, -w2);
    float p = w2 * x;
    float pe = fmaf(w2l, x, fmaf(w2, x, -p));
    two_sum(p, pe);
  }
  __device__ double value() const {
    if (!(fabsf(hi) <= FLT_MAX)) return (double)hi;
    return (double)hi + (double)lo;
  }
};

// Float32 weights and signal, float64 register accumulation: one full-rate
// DFMA per term on CDNA.
struct TWide {
  double s;
  __device__ void init() { s = 0.0; }
  __device__ void add(float w, double, float x) { s = fma((double)w, (double)x, s); }
  __device__ void add_var(float w, double, float x) {
    double wd = (double)w;
    s = fma(wd * wd, (double)fmaxf(x, 0.0f), s);
  }
  __device__ double value() const { return s; }
};

// Float64 weights, signal and accumulation (CudaPrecision::F64).
struct TF64 {
  double s;
  __device__ void init() { s = 0.0; }
  __device__ void add(float, double w, double x) { s = fma(w, x, s); }
  __device__ void add_var(float, double w, double x) { s = fma(w * w, fmax(x, 0.0), s); }
  __device__ double value() const { return s; }
};

template <int MODE> struct TAcc;
// T: accumulator; L: frame value type in LDS; P: stored segment partial
// (float32 partials add a relative error of at most 2^-24 per partial).
template <> struct TAcc<0> { using T = TPlain; using L = float; using P = float; };
template <> struct TAcc<1> { using T = TComp; using L = float; using P = float; };
template <> struct TAcc<2> { using T = TF64; using L = double; using P = double; };
template <> struct TAcc<3> { using T = TWide; using L = float; using P = float; };

struct TileArgs {
  const void* frames;
  int dtype;
  const uint16_t* lidx;
  const float* w32;
  const double* w64;
  const uint32_t* seg_group;
  const uint32_t* tile_meta;  // per tile: first/end segment, first/end group
  uint32_t nseg, ntx, ntiles, dim1, dim2;
  uint64_t npix;
  uint32_t nf;
  uint32_t aligned8;  // pixel pairs are aligned to their size (8 bytes, 4 for 16-bit frames)
  void* part;
  uint32_t flip = 0;  // frames hold the detector's rows bottom-up (row r is frame row dim1-1-r)
  // Double-buffered kernel only: npass passes over the same table in one
  // launch. Pass h reads frames h * F onwards and writes its partials pblock
  // bytes after pass h - 1's; passes before the last hold F frames, the last
  // nf_last (a single pass holds nf).
  uint32_t npass = 1;
  uint32_t nf_last = 0;
  uint64_t pblock = 0;
  // Elements per segment row of partials, 0 for F (a row per pass). With
  // prow = npass * F and pblock the bytes of F elements, the passes of a
  // launch share one row per segment: columns h * F .. h * F + F - 1 hold
  // pass h.
  uint32_t prow = 0;
  // Per-frame masks folded into the staging (double-buffered kernel, masked
  // rows): a pass of 16 rows holds 8 frames (rows 0-7, masked pixels read as
  // 0) and their masked float32 normalization (rows 8-15). mbits holds one
  // byte per pixel and pass (bit i: frame i of the pass is masked there),
  // input order, the passes mbits_stride bytes apart; norm32 the plan's
  // normalization as float32, detector order; nfr the frames of the last
  // pass (8 for the others).
  const uint8_t* mbits = nullptr;
  const float* norm32 = nullptr;
  uint64_t mbits_stride = 0;
  uint32_t nfr = 0;
};

// Segment (relative to the chunk's first) a slot takes in round r. Segments
// are stored longest first; odd rounds walk the slots backwards, so a slot
// with a long segment in one round gets a short one in the next.
template <uint32_t SLOTS>
__device__ __forceinline__ uint32_t seg_of_slot(uint32_t slot, uint32_t r) {
  return (r & 1) ? (r + 1) * SLOTS - 1 - slot : slot + r * SLOTS;
}

// Rounds whose segment bounds are prefetched into registers: enough for 128
// segments per tile.
template <uint32_t SLOTS> struct SegRounds {
  static constexpr int R = SLOTS >= 64 ? 2 : 4;
};

// Group bounds of the segments a slot takes in the first R rounds of a tile.
template <int R> struct SegRegs {
  uint32_t lo[R], hi[R];
};

// One segment's partials of this lane's G frames (and their variance terms).
template <int G, int MODE, bool VAR> struct SegVals {
  using PT = typename TAcc<MODE>::P;
  PT v[G];
  PT e[VAR ? G : 1];
};

// Writes segment s's partials of this lane's frames.
template <int F, int G, int MODE, bool VAR>
__device__ __forceinline__ void seg_write(const TileArgs& a, uint32_t s, uint32_t q,
                                          const SegVals<G, MODE, VAR>& o) {
  using PT = typename TAcc<MODE>::P;
  const size_t RS = a.prow ? a.prow : (uint32_t)F;
  PT* out = reinterpret_cast<PT*>(a.part) + (size_t)s * RS + q * G;
  if constexpr (sizeof(PT) == 4 && (G == 2 || G == 4)) {
    // All G frames present: one vector store per kind (the row offset is a
    // multiple of G floats).
    using V = typename std::conditional<G == 2, f32x2, f32x4>::type;
    if (q * G + G <= a.nf) {
      V v, e;
#pragma unroll
      for (int j = 0; j < G; j++) {
        v[j] = o.v[j];
        if constexpr (VAR) e[j] = o.e[j];
      }
      *reinterpret_cast<V*>(out) = v;
      if constexpr (VAR) *reinterpret_cast<V*>(out + (size_t)a.nseg * RS) = e;
      return;
    }
  }
#pragma unroll
  for (int j = 0; j < G; j++) {
    if (q * G + j < a.nf) {
      out[j] = o.v[j];
      if constexpr (VAR) out[(size_t)a.nseg * RS + j] = o.e[j];
    }
  }
}

// A lane's frames in a staged tile. Frame-major (FmView): frame j of the
// lane at base[j * kLdsStride + pixel]. Quad-major (PmView, 16-frame tiles of
// float32 values): frames 4k .. 4k + 3 of pixel p at word k * kPmQuad + 4p,
// so that a lane's 4 frames are one 16-byte read, and the quads a segment's
// 4 lanes read start at bank 4 (p + k) mod 32 (8 positions, not 2).
template <class LT> struct FmView {
  using T = LT;
  const LT* base;
  template <int G> __device__ __forceinline__ void load(uint32_t l, T (&x)[G]) const {
#pragma unroll
    for (int j = 0; j < G; j++) x[j] = base[j * kLdsStride + l];
  }
};

constexpr uint32_t kPmFrames = 16;
constexpr uint32_t kPmQuad = 4 * (kPadSlot + 1);  // words per quad plane (pixels and zero pixel)
__device__ __forceinline__ uint32_t pm_word(uint32_t p, uint32_t f) {
  return (f >> 2) * kPmQuad + 4u * p + (f & 3u);
}

struct PmView {
  using T = float;
  const float* fl;
  uint32_t f0;  // the lane's first frame
  template <int G> __device__ __forceinline__ void load(uint32_t l, T (&x)[G]) const {
    static_assert(G == 1 || G == 2 || G == 4, "frames per lane");
    const float* p = fl + pm_word(l, f0);
    if constexpr (G == 4) {
      const f32x4 v = *reinterpret_cast<const f32x4*>(p);
      x[0] = v.x, x[1] = v.y, x[2] = v.z, x[3] = v.w;
    } else if constexpr (G == 2) {
      const f32x2 v = *reinterpret_cast<const f32x2*>(p);
      x[0] = v.x, x[1] = v.y;
    } else {
      x[0] = *p;
    }
  }
};

// Applies a segment (table groups [ga, gb) of the LDS chunk) to this lane's G
// frames.
template <int G, int MODE, bool VAR, class FV>
__device__ __forceinline__ SegVals<G, MODE, VAR> seg_eval(const FV& base, const u32x2* tl,
                                                          const unsigned char* twb, uint32_t ga,
                                                          uint32_t gb) {
  using Acc = typename TAcc<MODE>::T;
  using PT = typename TAcc<MODE>::P;
  Acc acc[G], var[G];
#pragma unroll
  for (int j = 0; j < G; j++) {
    acc[j].init();
    if constexpr (VAR) var[j].init();
  }
  // Software pipeline over the segment's groups: the table words (pixel
  // indices and weights) of group g + 1 are read while group g's frame values
  // are in flight and its products accumulate, so a group waits for one LDS
  // round trip (its frame values), not two. The loop is unrolled twice over
  // two register sets (A, B), so the words move into place without copies.
  // The read after the segment's last group repeats that group (in the
  // chunk, unused). The products are added in the same order as a plain
  // loop over the groups, so the bits do not depend on the pipeline.
  const f64x2* twd = reinterpret_cast<const f64x2*>(twb);
  const f32x4* twf = reinterpret_cast<const f32x4*>(twb);
  struct Words {
    u32x2 l;
    f32x4 wf;
    f64x2 wd0, wd1;
  };
  auto read_words = [&](uint32_t g, Words& w) {
    w.l = tl[g];
    if constexpr (MODE == 2) {
      w.wd0 = twd[2 * g];
      w.wd1 = twd[2 * g + 1];
    } else {
      w.wf = twf[g];
    }
  };
  // Reads group g's frame values (indices in cur), then the next group's
  // words into nxt, then adds group g's products.
  auto step = [&](uint32_t g, const Words& cur, Words& nxt) {
    const uint32_t l[4] = {cur.l.x & 0xffffu, cur.l.x >> 16, cur.l.y & 0xffffu,
                           cur.l.y >> 16};
    typename FV::T x[4][G];
#pragma unroll
    for (int e = 0; e < 4; e++) base.template load<G>(l[e], x[e]);
    read_words(g + 1 < gb ? g + 1 : g, nxt);
    // Keep the reads above the products: the scheduler would otherwise sink
    // the next group's reads to just before their use.
    __builtin_amdgcn_sched_barrier(0);
    float wf[4];
    double wd[4];
    if constexpr (MODE == 2) {
      wd[0] = cur.wd0.x;
      wd[1] = cur.wd0.y;
      wd[2] = cur.wd1.x;
      wd[3] = cur.wd1.y;
      wf[0] = wf[1] = wf[2] = wf[3] = 0.f;
    } else {
      wf[0] = cur.wf.x;
      wf[1] = cur.wf.y;
      wf[2] = cur.wf.z;
      wf[3] = cur.wf.w;
      wd[0] = wd[1] = wd[2] = wd[3] = 0.0;
    }
#pragma unroll
    for (int e = 0; e < 4; e++)
#pragma unroll
      for (int j = 0; j < G; j++) {
        acc[j].add(wf[e], wd[e], x[e][j]);
        if constexpr (VAR) var[j].add_var(wf[e], wd[e], x[e][j]);
      }
  };
  if (ga < gb) {
    Words wa, wb;
    read_words(ga, wa);
    for (uint32_t g = ga;;) {
      step(g, wa, wb);
      if (++g >= gb) break;
      step(g, wb, wa);
      if (++g >= gb) break;
    }
  }
  SegVals<G, MODE, VAR> o;
#pragma unroll
  for (int j = 0; j < G; j++) {
    o.v[j] = static_cast<PT>(acc[j].value());
    if constexpr (VAR) o.e[j] = static_cast<PT>(var[j].value());
  }
  return o;
}

// Applies segment s to this lane's G frames and writes its partials.
template <int F, int G, int MODE, bool VAR, class FV>
__device__ __forceinline__ void seg_apply(const TileArgs& a, const FV& base, const u32x2* tl,
                                          const unsigned char* twb, uint32_t ga, uint32_t gb,
                                          uint32_t s, uint32_t q) {
  seg_write<F, G, MODE, VAR>(a, s, q, seg_eval<G, MODE, VAR>(base, tl, twb, ga, gb));
}

// Persistent-grid tile order: workgroups are dealt to the XCDs round-robin
// by id, so the workgroups of one XCD get consecutive tiles of a round; the
// row runs neighbouring tiles share (a 256-byte run rarely starts on a
// 128-byte line) are then fetched into the same L2 at about the same time.
__device__ __forceinline__ uint32_t first_tile() {
  const uint32_t grid = gridDim.x;
  if (grid % kXcds) return blockIdx.x;
  return (blockIdx.x % kXcds) * (grid / kXcds) + blockIdx.x / kXcds;
}

}  // namespace tk
}  // namespace dio
