// Detector-tiled gather (gfx950): the batched integration hot path.
//
// HBM layout of a tiled plan: lidx u16 + w32 f32 per entry (6 bytes), in
// 4-entry groups; segment -> first group u32[nseg + 1]; tile -> first
// segment u32[ntiles + 1]; bin -> segment list u32[nseg] + u32[nbins + 1];
// denominators f64[nbins]. Partials of one pass: [1 + errors][nseg][F]
// float32 (float64 for the F64 precision); a launch of the double-buffered
// kernel writes its passes side by side in one row per segment.
//
// k_tile: persistent workgroups of 512 threads, each walking detector tiles
// of kTileRows x kTileCols pixels (tile b, b + grid, ...).
//  1. Staging: every thread owns 2 consecutive pixels of one tile row (32
//     lanes cover a 256-byte row run) for each of the pass's F frames. For
//     32-bit frames the next tile's words and its table chunk (4-entry
//     groups: 8-byte indices + 16-byte weights) are loaded into registers
//     while the current tile is computed, then converted once and stored to
//     LDS as frames [frame][pixel] and the table chunk. Every frame byte is
//     read from HBM exactly once per pass, the table once per pass.
//  2. Gather: L = F / G lanes share a segment, each holding G frames in
//     registers, and apply the segment's groups from LDS. Segments of a tile
//     are stored longest first so the lanes of a wavefront run similar trip
//     counts.
//  3. One partial per (segment, frame), written frame-contiguous.
// k_tfinal adds each bin's partials in a fixed order (coalesced over frames)
// and divides by the FP64 denominator.
//
// The arithmetic of a frame (entry order inside a segment, partial order
// inside a bin) depends only on the table, never on F, G or the batch, so a
// frame gives the same bits alone, in any batch and in any pass size.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <type_traits>
#include <vector>

#include "tile_impl.hpp"

namespace dio {

using namespace tk;

namespace {

constexpr unsigned kThreads = 256;      // layout and reduction kernels
// k_tile workgroup size: 512 threads (8 wavefronts) for passes of 8 or more
// frames, 256 below, where each workgroup needs less LDS and more of them
// share a CU.
constexpr unsigned tile_threads(int F) { return F >= 8 ? 512u : 256u; }
constexpr unsigned kRunLanes = kTileCols / 2;  // lanes per 2-pixel row run
template <unsigned NT> struct Staging {
  static constexpr unsigned kRowsPerPass = NT / kRunLanes;
  static constexpr int kPasses = (int)(kTileRows / kRowsPerPass);  // staging loads per frame
  static_assert(kPasses >= 1 && kPasses * kRowsPerPass == kTileRows,
                "staging map covers the tile");
};

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

// Pieces: a (tile, bin) run longer than kMaxSegGroups groups is cut into
// pieces of that many groups (the last one shorter), each a segment of its
// own. Entry i starts a piece if it starts a run or sits a whole number of
// pieces into its run (run_first: first entry of each run).
__global__ void k_piece_flags(const uint32_t* flag, const uint32_t* runid,
                              const uint32_t* run_first, uint64_t nnz, uint32_t* pflag) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nnz) return;
  if (i == nnz) {
    pflag[i] = 0;
    return;
  }
  const uint32_t run = runid[i] + flag[i] - 1;
  const uint64_t pos = i - run_first[run];
  pflag[i] = (pos % ((uint64_t)kMaxSegGroups * kGroup) == 0) ? 1u : 0u;
}

__global__ void k_seg_fill(const uint32_t* flag, const uint32_t* segid, const uint32_t* tile,
                           const uint32_t* bin, uint64_t nnz, uint32_t* seg_first,
                           uint32_t* seg_tile, uint32_t* seg_bin) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz || !flag[i]) return;
  uint32_t s = segid[i];
  seg_first[s] = (uint32_t)i;
  seg_tile[s] = tile[i];
  seg_bin[s] = bin[i];
}

// Sort key placing a tile's segments longest first (ties keep bin order).
__global__ void k_seg_keys(const uint32_t* seg_first, const uint32_t* seg_tile, uint32_t nseg,
                           uint64_t* key) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  uint32_t ng = (seg_first[s + 1] - seg_first[s] + kGroup - 1) / kGroup;
  key[s] = ((uint64_t)seg_tile[s] << 32) | (0xFFFFFFFFu - ng);
}

__global__ void k_seg_perm(const uint32_t* perm, const uint32_t* seg_first,
                           const uint32_t* seg_tile, const uint32_t* seg_bin, uint32_t nseg,
                           uint32_t* inv, uint32_t* ngroups, uint32_t* tile_n, uint32_t* bin_n) {
  uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n > nseg) return;
  if (n == nseg) {
    ngroups[n] = 0;
    return;
  }
  uint32_t o = perm[n];
  inv[o] = n;
  ngroups[n] = (seg_first[o + 1] - seg_first[o] + kGroup - 1) / kGroup;
  tile_n[n] = seg_tile[o];
  bin_n[n] = seg_bin[o];
}

__global__ void k_pad_fill(uint64_t n, uint16_t* lidx, float* w32, double* w64) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  lidx[i] = (uint16_t)kPadSlot;
  w32[i] = 0.0f;
  if (w64) w64[i] = 0.0;
}

__global__ void k_place(const uint32_t* flag, const uint32_t* segid, const uint32_t* seg_first,
                        const uint32_t* inv, const uint32_t* seg_group, uint64_t nnz,
                        const uint16_t* lidx_in, const float* w32_in, const double* w64_in,
                        uint16_t* lidx, float* w32, double* w64) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  uint32_t so = segid[i] + flag[i] - 1;  // segment holding entry i
  uint64_t dst = (uint64_t)seg_group[inv[so]] * kGroup + (i - seg_first[so]);
  lidx[dst] = lidx_in[i];
  w32[dst] = w32_in[i];
  if (w64) w64[dst] = w64_in[i];
}

__global__ void k_tile_meta(const uint32_t* tile_seg, const uint32_t* seg_group,
                            uint32_t ntiles, uint32_t* meta) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const uint32_t s0 = tile_seg[t], s1 = tile_seg[t + 1];
  meta[4 * t] = s0;
  meta[4 * t + 1] = s1;
  meta[4 * t + 2] = seg_group[s0];
  meta[4 * t + 3] = seg_group[s1];
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

// rank[perm[k]] = k: the position of each segment in bin order.
__global__ void k_invert32(const uint32_t* perm, uint32_t n, uint32_t* rank) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) rank[perm[k]] = k;
}

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

// ---------------------------------------------------------------- LDS plan

// One workgroup's LDS: frames [F][kLdsStride] of LT, then a table chunk of at
// most kCap groups (kCap uint2 indices, then kCap float4 weights, or 2 x
// double2 per group for float64 weights). kWgs workgroups share a CU (160 KiB
// of LDS), so one stages while the other computes: two if that still leaves
// room for a chunk of kWantGroups groups, enough for a whole tile of the usual
// tables (entries per pixel about 2).
template <int F, int MODE> struct Plan {
  using LT = typename TAcc<MODE>::L;
  static constexpr unsigned kThreads = tile_threads(F);
  static constexpr size_t kFrameBytes = ((size_t)F * kLdsStride * sizeof(LT) + 15) / 16 * 16;
  static constexpr size_t kWBytes = MODE == 2 ? 32 : 16;
  static constexpr size_t kWantGroups = kTilePixels * 5 / (2 * kGroup);
  static constexpr size_t budget(uint32_t w) { return (160 * 1024 / w) / 256 * 256; }
  static constexpr bool fits(uint32_t w) {
    return budget(w) > kFrameBytes && (budget(w) - kFrameBytes) / (8 + kWBytes) >= kWantGroups;
  }
  // At most 16 wavefronts per CU (128 registers per lane); float64 lanes
  // need more registers, so at most 3 of their 256-thread workgroups.
  static constexpr uint32_t kWgs =
      kThreads == 512 ? (fits(2) ? 2 : 1)
                      : (MODE != 2 && fits(4)) ? 4 : fits(3) ? 3 : fits(2) ? 2 : 1;
  static constexpr size_t kBudget = budget(kWgs);
  static constexpr size_t kRoom = (kBudget - kFrameBytes) / (8 + kWBytes);
  static constexpr uint32_t kCap = kRoom >= 1024 ? 1024 : (uint32_t)(kRoom / 8 * 8);
  static constexpr size_t kBytes = kFrameBytes + (size_t)kCap * (8 + kWBytes);
  static constexpr int kKT = (kCap + kThreads - 1) / kThreads;  // table loads per thread
  static_assert(kBudget > kFrameBytes, "frames exceed the LDS budget");
  // A (tile, bin) segment holds at most one entry per tile pixel.
  static_assert(kCap >= kTilePixels / kGroup, "a whole segment must fit one table chunk");
};

// ---------------------------------------------------------------- staging

template <class P, size_t A> struct alignas(A) Run2 { P v[2]; };
template <class LT> struct alignas(2 * sizeof(LT)) Lds2 { LT v[2]; };

// Staging map of NT threads: thread tid owns pixel pairs (row h *
// kRowsPerPass + tid / 32, columns 2 * (tid % 32) + {0, 1}) of the tile for
// h < kPasses, so 32 lanes read one 256-byte row run and a wavefront two rows
// per load.
template <unsigned NT> struct TileGeom {
  static constexpr int kPasses = Staging<NT>::kPasses;
  uint32_t r0, cp;
  uint32_t nvalid[kPasses];  // valid pixels of the pair in each of its rows
  uint32_t off[kPasses];     // pixel offset in the frame (frames hold < 2^29 pixels)
};

template <unsigned NT>
__device__ __forceinline__ TileGeom<NT> tile_geom(const TileArgs& a, uint32_t t) {
  TileGeom<NT> g;
  g.r0 = threadIdx.x / kRunLanes;
  g.cp = threadIdx.x % kRunLanes;
  const uint32_t c = (t % a.ntx) * kTileCols + g.cp * 2;
#pragma unroll
  for (int h = 0; h < Staging<NT>::kPasses; h++) {
    const uint32_t r = (t / a.ntx) * kTileRows + g.r0 + Staging<NT>::kRowsPerPass * h;
    g.nvalid[h] = (r < a.dim1 && c < a.dim2) ? min(2u, a.dim2 - c) : 0u;
    g.off[h] = (a.flip && g.nvalid[h] ? a.dim1 - 1 - r : r) * a.dim2 + c;
  }
  return g;
}

template <unsigned NT>
__device__ __forceinline__ uint32_t lds_pixel(uint32_t r0, uint32_t cp, int h) {
  return (r0 + Staging<NT>::kRowsPerPass * h) * kTileCols + cp * 2;
}

// Raw 32-bit frame words of one tile held in registers (u32, i32, f32
// frames), loaded one tile ahead of their use.
template <int F, unsigned NT> struct FrameRegs {
  uint32_t v[F][Staging<NT>::kPasses][2];
};

// Load at a workgroup-uniform base plus a 32-bit per-lane byte offset. This
// is the addressing form of global loads with a scalar base register: each
// load needs one 32-bit offset register, not a 64-bit address pair.
template <class T>
__device__ __forceinline__ T load_at(const void* base, uint32_t byte_off) {
  return *reinterpret_cast<const T*>(static_cast<const unsigned char*>(base) + byte_off);
}

// Pair loads are 8-byte loads (A = 8) when every pair starts 8-byte aligned
// (even row length and an 8-byte aligned frame block), else two 4-byte loads
// (A = 4). The loads are unconditional, so all of a tile's loads issue back to
// back: lanes outside the detector, or of frames past the pass, read a valid
// address (pixel 0 of the pass's frame 0 or last frame) and their words are
// never used (no table entry points at such a pixel; partials of frames >= nf
// are not written). Byte offsets fit 32 bits: a frame holds < 2^29 pixels.
template <int F, int A, unsigned NT>
__device__ __forceinline__ void frames_load32(const TileArgs& a, uint32_t t,
                                              FrameRegs<F, NT>& fr) {
  static_assert(A == 4 || A == 8, "pair loads are 4 or 8 bytes");
  constexpr int kPasses = Staging<NT>::kPasses;
  const TileGeom<NT> g = tile_geom<NT>(a, t);
  const unsigned char* src = static_cast<const unsigned char*>(a.frames);
  const uint32_t flast = a.nf - 1;
  uint32_t o0[kPasses], o1[kPasses];
#pragma unroll
  for (int h = 0; h < kPasses; h++) {
    o0[h] = g.nvalid[h] > 0 ? g.off[h] * 4u : 0u;
    o1[h] = g.nvalid[h] > 1 ? g.off[h] * 4u + 4u : 0u;
  }
#pragma unroll
  for (int f = 0; f < F; f++) {
    const unsigned char* pf = src + (uint64_t)min((uint32_t)f, flast) * a.npix * 4u;
#pragma unroll
    for (int h = 0; h < kPasses; h++) {
      if constexpr (A == 8) {
        const u32x2 r = load_at<u32x2>(pf, o0[h]);
        fr.v[f][h][0] = r.x;
        fr.v[f][h][1] = r.y;
      } else {
        fr.v[f][h][0] = load_at<uint32_t>(pf, o0[h]);
        fr.v[f][h][1] = load_at<uint32_t>(pf, o1[h]);
      }
    }
  }
}

template <class W, class LT>
__device__ __forceinline__ LT word_to(uint32_t w) {
  if constexpr (std::is_same<W, float>::value) return static_cast<LT>(__uint_as_float(w));
  else return static_cast<LT>(static_cast<W>(w));
}

template <int F, class LT>
__device__ __forceinline__ void pad_slots(LT* fl) {
  static_assert(kPadSlot + 2 <= kLdsStride, "zero slot and spare word inside the row");
  if (threadIdx.x < (unsigned)F) {
    fl[threadIdx.x * kLdsStride + kPadSlot] = LT(0);
    fl[threadIdx.x * kLdsStride + kPadSlot + 1] = LT(0);
  }
}

template <int F, class W, class LT, unsigned NT>
__device__ __forceinline__ void frames_store_as(const FrameRegs<F, NT>& fr, LT* fl) {
  const uint32_t r0 = threadIdx.x / kRunLanes, cp = threadIdx.x % kRunLanes;
#pragma unroll
  for (int f = 0; f < F; f++) {
#pragma unroll
    for (int h = 0; h < Staging<NT>::kPasses; h++) {
      Lds2<LT> o;
      o.v[0] = word_to<W, LT>(fr.v[f][h][0]);
      o.v[1] = word_to<W, LT>(fr.v[f][h][1]);
      *reinterpret_cast<Lds2<LT>*>(fl + f * kLdsStride + lds_pixel<NT>(r0, cp, h)) = o;
    }
  }
}

template <int F, class LT, unsigned NT>
__device__ __forceinline__ void frames_store32(const TileArgs& a, const FrameRegs<F, NT>& fr,
                                               LT* fl) {
  if (a.dtype == DIO_U32) frames_store_as<F, uint32_t, LT>(fr, fl);
  else if (a.dtype == DIO_I32) frames_store_as<F, int32_t, LT>(fr, fl);
  else frames_store_as<F, float, LT>(fr, fl);
  pad_slots<F>(fl);
}

// Direct staging (u16 and f64 frames): load and store, eight frames of loads
// in flight per thread.
template <int F, class P, class LT, unsigned NT>
__device__ __forceinline__ void stage_direct(const TileArgs& a, uint32_t t, LT* fl) {
  constexpr int kPasses = Staging<NT>::kPasses;
  const P* fr = static_cast<const P*>(a.frames);
  const TileGeom<NT> g = tile_geom<NT>(a, t);
  constexpr int B = F < 8 ? F : 8;
#pragma unroll
  for (int k0 = 0; k0 < F; k0 += B) {
    P v[B][kPasses][2];
#pragma unroll
    for (int k = 0; k < B; k++) {
      // Unconditional loads from valid addresses (see frames_load32).
      const P* pf = fr + min((uint32_t)(k0 + k), a.nf - 1) * a.npix;
#pragma unroll
      for (int h = 0; h < kPasses; h++) {
        v[k][h][0] = pf[g.nvalid[h] > 0 ? g.off[h] : 0u];
        v[k][h][1] = pf[g.nvalid[h] > 1 ? g.off[h] + 1 : 0u];
      }
    }
#pragma unroll
    for (int k = 0; k < B; k++) {
#pragma unroll
      for (int h = 0; h < kPasses; h++) {
        Lds2<LT> o;
        o.v[0] = static_cast<LT>(v[k][h][0]);
        o.v[1] = static_cast<LT>(v[k][h][1]);
        *reinterpret_cast<Lds2<LT>*>(fl + (k0 + k) * kLdsStride + lds_pixel<NT>(g.r0, g.cp, h)) =
            o;
      }
    }
  }
  pad_slots<F>(fl);
}

// Table groups [gb, ge) of one chunk held in registers.
template <int KT, int MODE> struct TableRegs {
  u32x2 li[KT];
  f32x4 w[KT];
  f64x2 wa[MODE == 2 ? KT : 1], wb[MODE == 2 ? KT : 1];
};

template <int KT, int MODE, unsigned NT>
__device__ __forceinline__ void table_load(const TileArgs& a, uint32_t gb, uint32_t ge,
                                           TableRegs<KT, MODE>& tr) {
  // Unconditional loads (indices past the chunk repeat its last group and
  // are not stored): no divergent exec masks, no partially written registers.
  // The chunk's base is uniform; lanes add 32-bit byte offsets.
  const uint32_t last = ge - gb - 1;
  const void* lb = reinterpret_cast<const unsigned char*>(a.lidx) + (uint64_t)gb * 8;
  const void* wb32 = reinterpret_cast<const unsigned char*>(a.w32) + (uint64_t)gb * 16;
  const void* wb64 = reinterpret_cast<const unsigned char*>(a.w64) + (uint64_t)gb * 32;
#pragma unroll
  for (int k = 0; k < KT; k++) {
    const uint32_t i = min(threadIdx.x + k * NT, last);
    tr.li[k] = load_at<u32x2>(lb, i * 8u);
    if constexpr (MODE == 2) {
      tr.wa[k] = load_at<f64x2>(wb64, i * 32u);
      tr.wb[k] = load_at<f64x2>(wb64, i * 32u + 16u);
    } else {
      tr.w[k] = load_at<f32x4>(wb32, i * 16u);
    }
  }
}

template <int KT, int MODE, unsigned NT>
__device__ __forceinline__ void table_store(const TableRegs<KT, MODE>& tr, uint32_t n,
                                            u32x2* tl, unsigned char* twb) {
#pragma unroll
  for (int k = 0; k < KT; k++) {
    const uint32_t i = threadIdx.x + k * NT;
    if (i < n) {
      tl[i] = tr.li[k];
      if constexpr (MODE == 2) {
        f64x2* tw = reinterpret_cast<f64x2*>(twb);
        tw[2 * i] = tr.wa[k];
        tw[2 * i + 1] = tr.wb[k];
      } else {
        reinterpret_cast<f32x4*>(twb)[i] = tr.w[k];
      }
    }
  }
}

// Group bounds of the segments this slot takes in the first R rounds of the
// tile with metadata m (clamped to the tile's last segment).
template <uint32_t SLOTS, int R>
__device__ __forceinline__ void seg_bounds(const TileArgs& a, uint4 m, uint32_t slot,
                                           SegRegs<R>& sr) {
  const void* sb = reinterpret_cast<const unsigned char*>(a.seg_group) + (uint64_t)m.x * 4;
  const uint32_t last = m.y - m.x - 1;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t i = min(seg_of_slot<SLOTS>(slot, r), last);
    sr.lo[r] = load_at<uint32_t>(sb, i * 4u);
    sr.hi[r] = load_at<uint32_t>(sb, i * 4u + 4u);
  }
}

// Loads tile tt's frames, first table chunk and the slot's first segment
// bounds into registers (m = tile_meta[tt]); returns the groups loaded.
template <int F, int A, int KT, int MODE, uint32_t C, uint32_t SLOTS, unsigned NT>
__device__ __forceinline__ uint32_t prefetch(const TileArgs& a, uint32_t tt, uint4 m,
                                             uint32_t slot, FrameRegs<F, NT>& fr,
                                             TableRegs<KT, MODE>& tr,
                                             SegRegs<SegRounds<SLOTS>::R>& sr) {
  if (m.x == m.y) return 0;
  const uint32_t n = min(m.w - m.z, C);
  // Segment bounds first: loads complete in issue order, so the copy of the
  // bounds at the top of the next iteration (before its barrier) waits only
  // for these, not for the frames and table behind them.
  seg_bounds<SLOTS>(a, m, slot, sr);
  frames_load32<F, A>(a, tt, fr);
  table_load<KT, MODE, NT>(a, m.z, m.z + n, tr);
  return n;
}

// ---------------------------------------------------------------- tile kernel

// s_waitcnt immediate for vmcnt(0) (gfx9 encoding: vmcnt in bits 3:0 and
// 15:14, expcnt 6:4 and lgkmcnt 11:8 left at their no-wait maxima).
constexpr int kWaitVm0 = 0x0f70;

// Tile metadata (s0, s1, g0, g1) of a workgroup's tiles t0 + k * grid, 64
// iterations at a time: lane i of every wavefront holds iteration k0 + i. The
// tile loop reads it with v_readlane into scalar registers, so it waits on no
// memory there (a batch load waits once per 64 tiles). Iterations past the
// last tile read as empty tiles.
constexpr uint32_t kMetaLanes = 64;
struct MetaLanes {
  u32x4 v;
  uint32_t k0;
  __device__ __forceinline__ void load(const TileArgs& a, uint32_t t0, uint32_t grid,
                                       uint32_t k) {
    k0 = k;
    const uint32_t i = t0 + (k + threadIdx.x % kMetaLanes) * grid;
    const u32x4 z = {0u, 0u, 0u, 0u};
    v = i < a.ntiles ? reinterpret_cast<const u32x4*>(a.tile_meta)[i] : z;
    // Wait here, on this rare path: a wait where get() reads the lanes would
    // also wait for every store and load issued before it.
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
  }
  __device__ __forceinline__ uint4 get(uint32_t k) const {  // k0 <= k < k0 + 64
    const int l = (int)(k - k0);
    return make_uint4((uint32_t)__builtin_amdgcn_readlane((int)v.x, l),
                      (uint32_t)__builtin_amdgcn_readlane((int)v.y, l),
                      (uint32_t)__builtin_amdgcn_readlane((int)v.z, l),
                      (uint32_t)__builtin_amdgcn_readlane((int)v.w, l));
  }
};

// Metadata of one tile, read by every lane (rare path: it waits for its load
// at once, for the reason MetaLanes::load does).
__device__ __forceinline__ uint4 tile_meta_of(const TileArgs& a, uint32_t t) {
  const u32x4 v = reinterpret_cast<const u32x4*>(a.tile_meta)[t];
  __builtin_amdgcn_s_waitcnt(kWaitVm0);
  return make_uint4((uint32_t)__builtin_amdgcn_readfirstlane((int)v.x),
                    (uint32_t)__builtin_amdgcn_readfirstlane((int)v.y),
                    (uint32_t)__builtin_amdgcn_readfirstlane((int)v.z),
                    (uint32_t)__builtin_amdgcn_readfirstlane((int)v.w));
}

// Partials of the segments a slot takes in the first R rounds of a tile,
// held in registers (seg = ~0u: none) and written later.
template <int R, int G, int MODE, bool VAR> struct Pending {
  uint32_t seg[R];
  SegVals<G, MODE, VAR> vals[R];
  __device__ void clear() {
#pragma unroll
    for (int k = 0; k < R; k++) seg[k] = ~0u;
  }
};

template <int F, int R, int G, int MODE, bool VAR>
__device__ __forceinline__ void pending_write(const TileArgs& a, uint32_t q,
                                              Pending<R, G, MODE, VAR>& pd) {
#pragma unroll
  for (int k = 0; k < R; k++)
    if (pd.seg[k] != ~0u) seg_write<F, G, MODE, VAR>(a, pd.seg[k], q, pd.vals[k]);
  pd.clear();
}

// Computes one staged tile (m = its tile_meta): segments [s0, s1) in chunks
// of at most C table groups; the first chunk is already in LDS (unless
// first_in_lds is false), later chunks of a large tile are loaded here. cur
// holds the group bounds of this slot's first R segments. With pd, the
// partials of those R segments are left in pd (the caller writes them)
// instead of being stored here.
template <int F, int G, int MODE, bool VAR, uint32_t C, int KT, uint32_t SLOTS, unsigned NT,
          class FV>
__device__ __forceinline__ void tile_apply(
    const TileArgs& a, uint4 m, const SegRegs<SegRounds<SLOTS>::R>& cur, uint32_t slot,
    uint32_t q, const FV& base, u32x2* tl, unsigned char* twb,
    Pending<SegRounds<SLOTS>::R, G, MODE, VAR>* pd = nullptr, bool first_in_lds = true) {
  // Chunk bounds are workgroup-uniform: read them through the scalar cache so
  // that waiting for them never waits on the vector loads in flight.
  typedef const __attribute__((address_space(4))) uint32_t CWord;
  CWord* sg = (CWord*)(a.seg_group);
  const uint32_t s0 = m.x, s1 = m.y, g0 = m.z, g1 = m.w;
  uint32_t sa = s0, gc = g0;
  while (sa < s1) {
    uint32_t sb = s1;
    if (g1 - gc > C) {  // largest sb with seg_group[sb] - gc <= C
      uint32_t lo = sa + 1, hi = s1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (sg[mid] - gc <= C) lo = mid;
        else hi = mid - 1;
      }
      sb = lo;
    }
    if (sa != s0 || !first_in_lds) {  // a chunk not in LDS (a later one of a large tile)
      __syncthreads();
      TableRegs<KT, MODE> td;
      const uint32_t n = sg[sb] - gc;
      table_load<KT, MODE, NT>(a, gc, gc + n, td);
      table_store<KT, MODE, NT>(td, n, tl, twb);
      // Lanes that store nothing skip the store's wait; wait here on every
      // lane so no load of this rare path is still pending where it rejoins
      // the common path (the compiler would otherwise wait there, on the
      // next tile's prefetch as well).
      __builtin_amdgcn_s_waitcnt(kWaitVm0);
      __syncthreads();
    }
    const uint32_t rounds = (sb - sa + SLOTS - 1) / SLOTS;
    uint32_t r0 = 0;
    if (sa == s0) {
      // The first R rounds take their bounds from registers: no global load,
      // hence no wait on the next tile's prefetch, in the common path.
      constexpr int R = SegRounds<SLOTS>::R;
      const uint32_t rr = min(rounds, (uint32_t)R);
      for (uint32_t r = 0; r < rr; r++) {
        uint32_t lo = cur.lo[0], hi = cur.hi[0];
#pragma unroll
        for (int k = 1; k < R; k++) {
          lo = r == (uint32_t)k ? cur.lo[k] : lo;
          hi = r == (uint32_t)k ? cur.hi[k] : hi;
        }
        const uint32_t s_r = sa + seg_of_slot<SLOTS>(slot, r);
        if (s_r < sb) {
          if (pd) {
            const SegVals<G, MODE, VAR> v = seg_eval<G, MODE, VAR>(base, tl, twb, lo - gc, hi - gc);
#pragma unroll
            for (int k = 0; k < R; k++)
              if (r == (uint32_t)k) {
                pd->seg[k] = s_r;
                pd->vals[k] = v;
              }
          } else {
            seg_apply<F, G, MODE, VAR>(a, base, tl, twb, lo - gc, hi - gc, s_r, q);
          }
        }
      }
      r0 = R;
    }
    for (uint32_t r = r0; r < rounds; r++) {
      const uint32_t s = sa + seg_of_slot<SLOTS>(slot, r);
      if (s < sb)
        seg_apply<F, G, MODE, VAR>(a, base, tl, twb, a.seg_group[s] - gc,
                                   a.seg_group[s + 1] - gc, s, q);
    }
    sa = sb;
    if (sa < s1) gc = sg[sa];
  }
}

// Persistent: workgroup b processes tiles b, b + grid, ... For 32-bit frames
// (FA = 4 or 8: the byte width of a pair load; FA = 0 stages other frame types
// directly) the loop is software-pipelined: each iteration stages the tile whose
// frames and first table chunk are in registers, loads the next tile into
// the same registers, then computes the staged tile while those loads are in
// flight. There is one prefetch site, so the prefetched registers are not
// duplicated across branches (which would push the kernel past its register
// budget into spills that wait on the prefetch).
template <int F, int G, int MODE, bool VAR, int FA>
__global__ __launch_bounds__(tile_threads(F))
__attribute__((amdgpu_waves_per_eu(Plan<F, MODE>::kWgs * tile_threads(F) / 256))) void k_tile(
    TileArgs a) {
  using LT = typename TAcc<MODE>::L;
  using PL = Plan<F, MODE>;
  constexpr uint32_t C = PL::kCap;
  constexpr int KT = PL::kKT;
  constexpr uint32_t L = F / G;
  constexpr unsigned NT = PL::kThreads;
  constexpr uint32_t SLOTS = NT / L;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  LT* fl = reinterpret_cast<LT*>(lds_raw);
  u32x2* tl = reinterpret_cast<u32x2*>(lds_raw + PL::kFrameBytes);
  unsigned char* twb = lds_raw + PL::kFrameBytes + (size_t)C * 8;

  const uint32_t slot = threadIdx.x / L, q = threadIdx.x % L;
  const FmView<LT> base{fl + q * G * kLdsStride};

  const uint32_t grid = gridDim.x;
  const uint4 none = make_uint4(0u, 0u, 0u, 0u);
  const uint32_t t0 = first_tile();
  if (t0 >= a.ntiles) return;
  MetaLanes ml;
  ml.load(a, t0, grid, 0);

  if constexpr (FA != 0) {
    FrameRegs<F, NT> fr;
    TableRegs<KT, MODE> tr;
    SegRegs<SegRounds<SLOTS>::R> sr = {};
    uint32_t pn = 0;    // groups of the chunk in tr
    uint4 m = none;     // tile held in registers (none before the first load)
    uint4 mn = ml.get(0);
    // Iteration k prefetches tile t0 + k * grid and computes the one before.
    for (uint32_t k = 0;; k++) {
      const uint32_t tn = t0 + k * grid;
      const bool more = tn < a.ntiles;
      const bool work = m.x != m.y;
      const SegRegs<SegRounds<SLOTS>::R> cur = sr;
      if (work) {
        __syncthreads();  // the previous tile's compute is done with the LDS
        frames_store32<F, LT>(a, fr, fl);
        table_store<KT, MODE, NT>(tr, pn, tl, twb);
        __syncthreads();
      }
      if (k + 1 - ml.k0 >= kMetaLanes) ml.load(a, t0, grid, k + 1);
      const uint4 mnn = ml.get(k + 1);
      if (more) pn = prefetch<F, FA, KT, MODE, C, SLOTS, NT>(a, tn, mn, slot, fr, tr, sr);
      if (work) tile_apply<F, G, MODE, VAR, C, KT, SLOTS, NT>(a, m, cur, slot, q, base, tl, twb);
      if (!more) break;
      m = mn;
      mn = mnn;
    }
  } else {
    for (uint32_t k = 0;; k++) {
      const uint32_t t = t0 + k * grid;
      if (t >= a.ntiles) break;
      if (k - ml.k0 >= kMetaLanes) ml.load(a, t0, grid, k);
      const uint4 m = ml.get(k);
      if (m.x == m.y) continue;  // nothing of the table in this tile
      __syncthreads();
      switch (a.dtype) {
        case DIO_U16: stage_direct<F, uint16_t, LT, NT>(a, t, fl); break;
        default: stage_direct<F, double, LT, NT>(a, t, fl); break;
      }
      TableRegs<KT, MODE> td;
      const uint32_t n = min(m.w - m.z, C);
      table_load<KT, MODE, NT>(a, m.z, m.z + n, td);
      table_store<KT, MODE, NT>(td, n, tl, twb);
      SegRegs<SegRounds<SLOTS>::R> cur;
      seg_bounds<SLOTS>(a, m, slot, cur);
      __syncthreads();
      tile_apply<F, G, MODE, VAR, C, KT, SLOTS, NT>(a, m, cur, slot, q, base, tl, twb);
    }
  }
}

// ---------------------------------------------------------------- double-buffered tile kernel

// Double-buffered variant for 16-frame passes of 32-bit and 16-bit frames
// (float32 or float64 sums of float32 weights): one workgroup
// of 1024 threads per CU holds two LDS frame tiles and two table chunks.
// The work of a workgroup is a sequence of items, an item being one tile and
// one pass: with several passes per launch (npass, up to four), the passes of
// a tile are consecutive items that share one table chunk in LDS, so the table
// is read from HBM once per launch. While item i is computed from one frame
// buffer, item i + 1 (loaded into registers during the previous iteration) is
// staged into the other, and item i + 2 is loaded into the registers, so the
// frame loads stay in flight across the whole iteration and an iteration
// needs one barrier. Threads 0-511 stage the first half of the frames,
// threads 512-1023 the second half, each one pixel pair of a row.
// An item's partials (those of its first two rounds of segments) are written
// in the next iteration, after the wait for the staged frames and before the
// next loads: the memory counter also counts stores, so stores issued
// between a load and its wait would be waited for too.
constexpr unsigned kDbThreads = 1024;

template <int F, int MODE> struct PlanDb {
  using LT = typename TAcc<MODE>::L;
  static constexpr size_t kFrameBytes = (Plan<F, MODE>::kFrameBytes + 255) / 256 * 256;
  static constexpr size_t kWBytes = MODE == 2 ? 32 : 16;
  static constexpr size_t kRoom = (80 * 1024 - kFrameBytes) / (8 + kWBytes);
  static constexpr uint32_t kCap = kRoom >= 1024 ? 1024 : (uint32_t)(kRoom / 8 * 8);
  static constexpr size_t kTableBytes = ((size_t)kCap * (8 + kWBytes) + 255) / 256 * 256;
  // [frames 0][frames 1][table 0][table 1]
  static constexpr size_t kTable0 = 2 * kFrameBytes;
  static constexpr size_t kBytes = 2 * kFrameBytes + 2 * kTableBytes;
  static_assert(kBytes <= 160 * 1024, "two LDS tiles exceed one CU");
  static_assert(kCap >= kTilePixels / kGroup, "a whole segment must fit one table chunk");
  static_assert(kCap <= kDbThreads, "one table group per thread");
};

template <int F> struct DbFrames {
  uint32_t v[F / 2][2];
  uint32_t mk0, mk1;  // masked rows: the pair's mask bit bytes (pixel 0, pixel 1)
};

// Bytes per pixel of the frames a pair load of A bytes reads (16-bit frames:
// A = 2, one 4-byte load per pixel pair, or A = 1, two 2-byte loads).
template <int A> constexpr uint32_t db_esz() { return A <= 2 ? 2u : 4u; }

// Arguments of pass h < npass of a launch.
template <int F, int A>
__device__ __forceinline__ TileArgs db_pass(const TileArgs& a, uint32_t h) {
  TileArgs r = a;
  if (a.npass > 1) {
    r.frames = static_cast<const unsigned char*>(a.frames) +
               (uint64_t)h * F * a.npix * db_esz<A>();
    r.nf = h + 1 == a.npass ? a.nf_last : (uint32_t)F;
    r.part = static_cast<unsigned char*>(a.part) + h * a.pblock;
  }
  return r;
}

// Pair loads: 32-bit frames one 8-byte load (A = 8) or two 4-byte loads
// (A = 4); 16-bit frames one 4-byte load holding both pixels (A = 2, kept
// packed until the store) or two 2-byte loads (A = 1).
template <int F, int A>
__device__ __forceinline__ void db_load(const TileArgs& a, uint32_t t, DbFrames<F>& fr) {
  static_assert(A == 1 || A == 2 || A == 4 || A == 8, "pair loads of 1, 2, 4 or 8 bytes");
  constexpr uint32_t E = db_esz<A>();
  const uint32_t half = threadIdx.x / 512, lt = threadIdx.x % 512;
  const uint32_t r0 = lt / kRunLanes, cp = lt % kRunLanes;
  const uint32_t c = (t % a.ntx) * kTileCols + cp * 2;
  const uint32_t r = (t / a.ntx) * kTileRows + r0;
  const uint32_t nvalid = (r < a.dim1 && c < a.dim2) ? min(2u, a.dim2 - c) : 0u;
  const uint32_t off = (a.flip && nvalid ? a.dim1 - 1 - r : r) * a.dim2 + c;
  // Unconditional loads from valid addresses, as in frames_load32.
  const uint32_t o0 = nvalid > 0 ? off * E : 0u, o1 = nvalid > 1 ? off * E + E : 0u;
  const unsigned char* src = static_cast<const unsigned char*>(a.frames);
  const uint32_t flast = a.nf - 1;
#pragma unroll
  for (int i = 0; i < F / 2; i++) {
    const uint32_t f = half * (F / 2) + i;
    const unsigned char* pf = src + (uint64_t)min(f, flast) * a.npix * E;
    if constexpr (A == 8) {
      const u32x2 v = load_at<u32x2>(pf, o0);
      fr.v[i][0] = v.x;
      fr.v[i][1] = v.y;
    } else if constexpr (A == 4) {
      fr.v[i][0] = load_at<uint32_t>(pf, o0);
      fr.v[i][1] = load_at<uint32_t>(pf, o1);
    } else if constexpr (A == 2) {
      fr.v[i][0] = load_at<uint32_t>(pf, o0);
      fr.v[i][1] = 0u;
    } else {
      fr.v[i][0] = load_at<uint16_t>(pf, o0);
      fr.v[i][1] = load_at<uint16_t>(pf, o1);
    }
  }
}

// Stores the pair's values (frames or rows 8 h .. 8 h + 7 of both pixels,
// v[i][k]: row 8 h + i of pixel k) in the quad-major layout (PmView): two
// quads per pixel; neighbouring lanes store neighbouring pixel pairs, 32
// bytes apart, so each store instruction covers every bank.
__device__ __forceinline__ void pm_store_pair(float* fl, uint32_t px, uint32_t half,
                                              const float (&v)[8][2]) {
#pragma unroll
  for (uint32_t k = 0; k < 2; k++) {
#pragma unroll
    for (uint32_t qd = 0; qd < 2; qd++) {
      const f32x4 o = {v[4 * qd][k], v[4 * qd + 1][k], v[4 * qd + 2][k], v[4 * qd + 3][k]};
      *reinterpret_cast<f32x4*>(fl + pm_word(px + k, 8u * half + 4u * qd)) = o;
    }
  }
}

// PACKED: both pixels of a 16-bit pair in v[i][0] (low half first).
template <int F, class W, class LT, bool PACKED = false>
__device__ __forceinline__ void db_store_as(const DbFrames<F>& fr, LT* fl) {
  static_assert(F == 16 && std::is_same<LT, float>::value, "pixel-major 16-frame tiles");
  const uint32_t half = threadIdx.x / 512, lt = threadIdx.x % 512;
  const uint32_t r0 = lt / kRunLanes, cp = lt % kRunLanes;
  float v[8][2];
#pragma unroll
  for (int i = 0; i < F / 2; i++) {
    if constexpr (PACKED) {
      v[i][0] = word_to<W, LT>(fr.v[i][0] & 0xffffu);
      v[i][1] = word_to<W, LT>(fr.v[i][0] >> 16);
    } else {
      v[i][0] = word_to<W, LT>(fr.v[i][0]);
      v[i][1] = word_to<W, LT>(fr.v[i][1]);
    }
  }
  pm_store_pair(fl, r0 * kTileCols + cp * 2, half, v);
}

// The quad-major layout's zero pixel (kPadSlot).
__device__ __forceinline__ void pm_pad(float* fl) {
  if (threadIdx.x < kPmFrames) fl[pm_word(kPadSlot, threadIdx.x)] = 0.0f;
}

template <int F, int A, class LT>
__device__ __forceinline__ void db_store(const TileArgs& a, const DbFrames<F>& fr, LT* fl) {
  if constexpr (A <= 2) {
    db_store_as<F, uint16_t, LT, A == 2>(fr, fl);
  } else {
    if (a.dtype == DIO_U32) db_store_as<F, uint32_t, LT>(fr, fl);
    else if (a.dtype == DIO_I32) db_store_as<F, int32_t, LT>(fr, fl);
    else db_store_as<F, float, LT>(fr, fl);
  }
}

// Masked rows (see TileArgs::mbits): pass h of a launch holds frames 8 h ..
// 8 h + 7 (the last, nfr of them) and their masks' bit byte plane h.
constexpr uint32_t kMaskedFrames = 8;

template <int F, int A>
__device__ __forceinline__ TileArgs db_pass_masked(const TileArgs& a, uint32_t h) {
  TileArgs r = a;
  r.frames = static_cast<const unsigned char*>(a.frames) +
             (uint64_t)h * kMaskedFrames * a.npix * db_esz<A>();
  r.mbits = a.mbits + (uint64_t)h * a.mbits_stride;
  r.nfr = h + 1 == a.npass ? a.nfr : kMaskedFrames;
  r.nf = (uint32_t)F;
  r.part = static_cast<unsigned char*>(a.part) + h * a.pblock;
  return r;
}

// Threads 0-511 load the pair's 8 frames (frames past nfr read the pass's
// last one), threads 512-1023 the pair's float32 normalization (detector
// order; the same words for all 8 slots, so both halves run the same loads
// and fill the same registers); both the pair's mask bits. Unconditional
// loads from valid addresses, as in db_load. 16-bit frames (A = 2, rows of
// even length, pairs 4-byte aligned): each load is one 4-byte word, the
// frame pair packed, or normalization pixel i & 1 of the pair.
template <int F, int A>
__device__ __forceinline__ void db_load_masked(const TileArgs& a, uint32_t t, DbFrames<F>& fr) {
  static_assert(A == 2 || A == 4 || A == 8, "pair loads of 2, 4 or 8 bytes");
  constexpr uint32_t E = db_esz<A>();
  static_assert(F == 2 * kMaskedFrames, "8 frames and their normalization");
  const uint32_t half = threadIdx.x / 512, lt = threadIdx.x % 512;
  const uint32_t r0 = lt / kRunLanes, cp = lt % kRunLanes;
  const uint32_t c = (t % a.ntx) * kTileCols + cp * 2;
  const uint32_t r = (t / a.ntx) * kTileRows + r0;
  const uint32_t nvalid = (r < a.dim1 && c < a.dim2) ? min(2u, a.dim2 - c) : 0u;
  const uint32_t off = (a.flip && nvalid ? a.dim1 - 1 - r : r) * a.dim2 + c;
  const uint32_t o0 = nvalid > 0 ? off : 0u, o1 = nvalid > 1 ? off + 1 : 0u;
  const uint32_t d0 = nvalid > 0 ? r * a.dim2 + c : 0u, d1 = nvalid > 1 ? d0 + 1 : 0u;
  const uint32_t p0 = half ? d0 : o0, p1 = half ? d1 : o1;
  const uint64_t step = half ? 0u : (uint64_t)a.npix * E;
  const unsigned char* src =
      half ? reinterpret_cast<const unsigned char*>(a.norm32)
           : static_cast<const unsigned char*>(a.frames);
#pragma unroll
  for (uint32_t i = 0; i < kMaskedFrames; i++) {
    const unsigned char* pf = src + (uint64_t)min(i, a.nfr - 1) * step;
    if constexpr (A == 8) {
      const u32x2 v = load_at<u32x2>(pf, p0 * 4u);
      fr.v[i][0] = v.x;
      fr.v[i][1] = v.y;
    } else if constexpr (A == 4) {
      fr.v[i][0] = load_at<uint32_t>(pf, p0 * 4u);
      fr.v[i][1] = load_at<uint32_t>(pf, p1 * 4u);
    } else {
      fr.v[i][0] = load_at<uint32_t>(pf, half ? (p0 + (i & 1u)) * 4u : p0 * 2u);
      fr.v[i][1] = 0u;
    }
  }
  // Kept apart until the store: combining them here would wait for these
  // loads before the next item's loads are issued.
  fr.mk0 = a.mbits[o0];
  fr.mk1 = a.mbits[o1];
}

// Rows 0-7: the frames with masked pixels set to 0; rows 8-15: the
// normalization with masked pixels set to 0. PACKED: 16-bit frame pairs in
// v[i][0], the normalization pair in v[0][0] and v[1][0].
template <int F, class W, class LT, bool PACKED = false>
__device__ __forceinline__ void db_store_masked_as(const DbFrames<F>& fr, LT* fl) {
  const uint32_t half = threadIdx.x / 512, lt = threadIdx.x % 512;
  const uint32_t r0 = lt / kRunLanes, cp = lt % kRunLanes;
  const uint32_t px = r0 * kTileCols + cp * 2;
  const uint32_t m = fr.mk0 | (fr.mk1 << 8);
  static_assert(F == 16 && std::is_same<LT, float>::value, "pixel-major 16-frame tiles");
  float v[8][2];
  if (half == 0) {
#pragma unroll
    for (uint32_t i = 0; i < kMaskedFrames; i++) {
      const uint32_t w0 = PACKED ? fr.v[i][0] & 0xffffu : fr.v[i][0];
      const uint32_t w1 = PACKED ? fr.v[i][0] >> 16 : fr.v[i][1];
      v[i][0] = (m >> i) & 1u ? 0.0f : word_to<W, LT>(w0);
      v[i][1] = (m >> (8 + i)) & 1u ? 0.0f : word_to<W, LT>(w1);
    }
  } else {
    const float n0 = __uint_as_float(fr.v[0][0]);
    const float n1 = __uint_as_float(PACKED ? fr.v[1][0] : fr.v[0][1]);
#pragma unroll
    for (uint32_t i = 0; i < kMaskedFrames; i++) {
      v[i][0] = (m >> i) & 1u ? 0.0f : n0;
      v[i][1] = (m >> (8 + i)) & 1u ? 0.0f : n1;
    }
  }
  pm_store_pair(fl, px, half, v);
}

template <int F, int A, class LT>
__device__ __forceinline__ void db_store_masked(const TileArgs& a, const DbFrames<F>& fr,
                                                LT* fl) {
  if constexpr (A == 2) db_store_masked_as<F, uint16_t, LT, true>(fr, fl);
  else if (a.dtype == DIO_U32) db_store_masked_as<F, uint32_t, LT>(fr, fl);
  else if (a.dtype == DIO_I32) db_store_masked_as<F, int32_t, LT>(fr, fl);
  else db_store_masked_as<F, float, LT>(fr, fl);
}

// Loads item (tile tt, pass arguments ap): frames and, for the first pass of
// a tile, its segment bounds and first table chunk (returns its groups). The
// later passes of the tile keep the bounds already in sr.
template <int F, int A, int MODE, uint32_t C, uint32_t SLOTS, bool MASKED = false>
__device__ __forceinline__ uint32_t db_prefetch(const TileArgs& ap, uint32_t tt, uint4 m,
                                                bool first, uint32_t slot, DbFrames<F>& fr,
                                                TableRegs<1, MODE>& tr,
                                                SegRegs<SegRounds<SLOTS>::R>& sr) {
  if (m.x == m.y) return 0;
  const uint32_t n = min(m.w - m.z, C);
  if (first) seg_bounds<SLOTS>(ap, m, slot, sr);
  if constexpr (MASKED) db_load_masked<F, A>(ap, tt, fr);
  else db_load<F, A>(ap, tt, fr);
  if (first) table_load<1, MODE, kDbThreads>(ap, m.z, m.z + n, tr);
  return n;
}

template <int F, int G, int MODE, bool VAR, int FA, bool MASKED = false>
__global__ __launch_bounds__(kDbThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_tile_db(
    TileArgs a) {
  static_assert(FA == 1 || FA == 2 || FA == 4 || FA == 8, "16- and 32-bit frames only");
  static_assert(!MASKED || FA >= 2, "masked rows need pair-aligned 16-bit or 32-bit frames");
  static_assert(F % 2 == 0, "frames split over two thread halves");
  using LT = typename TAcc<MODE>::L;
  using PL = PlanDb<F, MODE>;
  constexpr uint32_t C = PL::kCap;
  constexpr uint32_t L = F / G;
  constexpr uint32_t SLOTS = kDbThreads / L;
  constexpr int R = SegRounds<SLOTS>::R;
  static_assert((size_t)(kPmFrames / 4) * kPmQuad * sizeof(float) <= PL::kFrameBytes,
                "a quad-major tile fits its frame buffer");
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const uint32_t slot = threadIdx.x / L, q = threadIdx.x % L;
  // Items: K = ntiles / grid full rounds give the workgroup's tiles
  // t0 + k * grid, each for H passes in a row (item k * H + h). The H passes
  // of the rem = ntiles - K * grid remaining tiles are dealt one item per
  // workgroup (item K * H + e is item j = blockIdx.x + e * grid of the
  // remainder), so the last round costs about one item, not H.
  const uint32_t H = a.npass;
  const uint32_t grid = gridDim.x;
  const uint32_t t0 = first_tile();
  if (t0 >= a.ntiles) return;
  const uint32_t K = a.ntiles / grid, rem = a.ntiles - K * grid;
  struct Item {
    uint32_t tile, pass, tbuf;
    bool valid, first, round;  // round: a full-round item (metadata in lanes)
  };
  auto item = [&](uint32_t i) {
    Item it;
    if (i < K * H) {
      const uint32_t k = i / H;
      it.tile = t0 + k * grid;
      it.pass = i % H;
      it.tbuf = k & 1;
      it.valid = true;
      it.first = it.pass == 0;
      it.round = true;
    } else {
      const uint32_t e = i - K * H, j = blockIdx.x + e * grid;
      it.tile = K * grid + j / H;
      it.pass = j % H;
      it.tbuf = (K + e) & 1;
      it.valid = j < rem * H;
      it.first = true;
      it.round = false;
    }
    return it;
  };
  auto frames_buf = [&](uint32_t i) { return lds_raw + (i & 1) * PL::kFrameBytes; };
  auto table_buf = [&](const Item& it) {
    return lds_raw + PL::kTable0 + it.tbuf * PL::kTableBytes;
  };
  // Arguments of pass h of the launch.
  auto pass_args = [&](uint32_t h) {
    if constexpr (MASKED) return db_pass_masked<F, FA>(a, h);
    else return db_pass<F, FA>(a, h);
  };
  MetaLanes ml;
  ml.load(a, t0, grid, 0);
  // Metadata of an item (zero past the last one).
  auto meta = [&](uint32_t i, const Item& it) {
    if (!it.valid) return make_uint4(0u, 0u, 0u, 0u);
    if (it.round) {
      const uint32_t k = i / H;
      if (k - ml.k0 >= kMetaLanes) ml.load(a, t0, grid, k);
      return ml.get(k);
    }
    return tile_meta_of(a, it.tile);
  };
  pm_pad(reinterpret_cast<float*>(frames_buf(0)));
  pm_pad(reinterpret_cast<float*>(frames_buf(1)));

  DbFrames<F> fr;
  TableRegs<1, MODE> tr;
  SegRegs<R> sr = {};
  Pending<R, G, MODE, VAR> pd;
  pd.clear();
  uint32_t ph = 0;  // pass of the item whose partials are pending
  // Prologue: item 0 staged, item 1 loaded into the registers.
  Item iA = item(0);
  if (!iA.valid) return;
  uint4 mA = meta(0, iA);
  uint32_t pn = db_prefetch<F, FA, MODE, C, SLOTS, MASKED>(pass_args(iA.pass), iA.tile, mA,
                                                           true, slot, fr, tr, sr);
  if (mA.x != mA.y) {
    LT* fb0 = reinterpret_cast<LT*>(frames_buf(0));
    if constexpr (MASKED) db_store_masked<F, FA, LT>(a, fr, fb0);
    else db_store<F, FA, LT>(a, fr, fb0);
    table_store<1, MODE, kDbThreads>(tr, pn, reinterpret_cast<u32x2*>(table_buf(iA)),
                                     table_buf(iA) + (size_t)C * 8);
  }
  SegRegs<R> cA = sr;
  Item iB = item(1);
  uint4 mB = meta(1, iB);
  if (iB.valid)
    pn = db_prefetch<F, FA, MODE, C, SLOTS, MASKED>(pass_args(iB.pass), iB.tile, mB,
                                                    iB.first, slot, fr, tr, sr);
  __syncthreads();

  for (uint32_t i = 0;; i++) {
    if (!iA.valid) break;
    // 1. Item i + 1 (registers) into the other frame buffer, free since the
    //    last barrier, and its table chunk if it opens a tile (the other
    //    table buffer, last read by the previous tile).
    if (mB.x != mB.y) {
      LT* fb1 = reinterpret_cast<LT*>(frames_buf(i + 1));
      if constexpr (MASKED) db_store_masked<F, FA, LT>(a, fr, fb1);
      else db_store<F, FA, LT>(a, fr, fb1);
      if (iB.first)
        table_store<1, MODE, kDbThreads>(tr, pn, reinterpret_cast<u32x2*>(table_buf(iB)),
                                         table_buf(iB) + (size_t)C * 8);
    }
    const SegRegs<R> cB = sr;
    // Every load of the last prefetch has landed on every path (the store
    // above waited for them): said explicitly, so that no later wait in this
    // iteration also waits for the partial stores issued next.
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    // 2. Partials of item i - 1.
    pending_write<F>(pass_args(ph), q, pd);
    // 3. Item i + 2 into the registers.
    const Item iC = item(i + 2);
    const uint4 mC = meta(i + 2, iC);
    if (iC.valid)
      pn = db_prefetch<F, FA, MODE, C, SLOTS, MASKED>(pass_args(iC.pass), iC.tile, mC,
                                                      iC.first, slot, fr, tr, sr);
    // 4. Item i from its buffers. A later pass of a tile whose table takes
    //    several chunks reloads the first (the last chunk is in LDS).
    if (mA.x != mA.y) {
      const PmView base{reinterpret_cast<const float*>(frames_buf(i)), q * G};
      unsigned char* tb = table_buf(iA);
      const bool first_in_lds = iA.first || mA.w - mA.z <= C;
      // The first slots hold the tile's longest segments, which set the
      // iteration's length: their wavefronts issue first while they compute.
      if (threadIdx.x < kDbThreads / 2) __builtin_amdgcn_s_setprio(2);
      tile_apply<F, G, MODE, VAR, C, 1, SLOTS, kDbThreads>(
          pass_args(iA.pass), mA, cA, slot, q, base, reinterpret_cast<u32x2*>(tb),
          tb + (size_t)C * 8, &pd, first_in_lds);
      __builtin_amdgcn_s_setprio(0);
    }
    ph = iA.pass;
    __syncthreads();
    iA = iB;
    mA = mB;
    cA = cB;
    iB = iC;
    mB = mC;
  }
  pending_write<F>(pass_args(ph), q, pd);
}

// prepared.rs:985-997 (normalize_bin) after adding a bin's partials: I =
// s / den, sigma = sqrt(max(v / den^2, 0)), 0 where den == 0. One launch
// reduces the frames of consecutive passes: frame f of segment g is column
// f % FS of row g of the block at part + (f / FS) * BS (a block per pass
// with FS = F, or one block whose rows hold every pass of a launch side by
// side with FS = passes * F). A workgroup serves one bin and one group of at
// most 32 frames (blockIdx.y): 64 segment lanes times Q frame quads of FPT
// frames (Q = ceil(frames / FPT) rounded up to a power of two, so a
// wavefront holds whole segment lanes; quads past the group's frames read
// columns that are not stored, and no quad straddles two rows as FPT
// divides FS). Segment lane l adds the bin's partials l, l + 64, l + 128, ...
// (tile order), FPT frames at a time with one vector load; the 64 lane sums
// are then combined by a fixed pairwise tree (l with l ^ 1, then l ^ 2, ...
// l ^ 32), inside the wavefront by lane exchange and across wavefronts
// through LDS. The order depends only on the table, never on the pass or
// batch shape.
constexpr unsigned kFinalLanes = 64;
constexpr unsigned kFinalMaxQuads = 8;  // at most 32 frames per workgroup

template <class PT, int FPT> struct PartVec { PT v[FPT]; };

template <class PT, int FPT>
__device__ __forceinline__ PartVec<PT, FPT> load_part(const PT* p) {
  PartVec<PT, FPT> r;
  if constexpr (FPT == 4 && sizeof(PT) == 4) {
    const f32x4 x = *reinterpret_cast<const f32x4*>(p);
    r.v[0] = x.x; r.v[1] = x.y; r.v[2] = x.z; r.v[3] = x.w;
  } else if constexpr (FPT == 4 && sizeof(PT) == 8) {
    const f64x2 x = reinterpret_cast<const f64x2*>(p)[0], y = reinterpret_cast<const f64x2*>(p)[1];
    r.v[0] = x.x; r.v[1] = x.y; r.v[2] = y.x; r.v[3] = y.y;
  } else {
#pragma unroll
    for (int k = 0; k < FPT; k++) r.v[k] = p[k];
  }
  return r;
}

template <bool PLAIN, bool VAR, class PT, int FPT>
__global__ __launch_bounds__(kFinalLanes * kFinalMaxQuads) void k_tfinal(
    const PT* __restrict__ part, uint32_t nseg, uint32_t FS, uint64_t BS,
    const uint32_t* __restrict__ bin_ptr,
    const uint32_t* __restrict__ bin_seg, const double* __restrict__ den, uint32_t nbins,
    uint32_t ntot, uint32_t RF, int raw, double* __restrict__ I, double* __restrict__ S) {
  using AT = typename std::conditional<PLAIN, float, double>::type;
  // Frame group blockIdx.y: frames [f0, f0 + nf) of the ntot, RF per group
  // (a multiple or a divisor of FS, so a group's rows lie in whole blocks).
  const uint32_t f0 = blockIdx.y * RF, nf = min(RF, ntot - f0);
  I += (uint64_t)f0 * nbins;
  if constexpr (VAR) S += (uint64_t)f0 * nbins;
  constexpr int NV = VAR ? 2 : 1;
  __shared__ AT red[kFinalMaxQuads][kFinalMaxQuads][NV][FPT];  // [wavefront][quad][kind][frame]
  const uint32_t quads = blockDim.x / kFinalLanes;
  const uint32_t tid = threadIdx.x;
  const uint32_t q = tid % quads, l = tid / quads;
  const uint32_t b = blockIdx.x;
  const uint32_t q0 = bin_ptr[b], q1 = bin_ptr[b + 1];
  AT acc[NV][FPT];
#pragma unroll
  for (int k = 0; k < FPT; k++) {
    acc[0][k] = 0;
    if constexpr (VAR) acc[NV - 1][k] = 0;
  }
  const uint32_t fc = f0 + q * FPT;
  const PT* pr = part + (uint64_t)(fc / FS) * BS + fc % FS;
  const PT* pv = pr + (size_t)nseg * FS;
  // Partials of this lane (i = q0 + l, q0 + l + 64, ...), added in order:
  // U segment indices are read at once, then the U partial rows they name,
  // so a bin of up to 64 U segments costs two memory round trips.
  constexpr int U = sizeof(PT) == 8 ? 4 : 8;
  for (uint32_t i0 = q0 + l; i0 < q1; i0 += U * kFinalLanes) {
    uint32_t g[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t i = i0 + u * kFinalLanes;
      g[u] = i < q1 ? bin_seg[i] : 0u;
    }
    PartVec<PT, FPT> x[U], y[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (i0 + u * kFinalLanes < q1) {
        x[u] = load_part<PT, FPT>(pr + (size_t)g[u] * FS);
        if constexpr (VAR) y[u] = load_part<PT, FPT>(pv + (size_t)g[u] * FS);
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (i0 + u * kFinalLanes < q1) {
#pragma unroll
        for (int k = 0; k < FPT; k++) {
          acc[0][k] = acc[0][k] + static_cast<AT>(x[u].v[k]);
          if constexpr (VAR) acc[NV - 1][k] = acc[NV - 1][k] + static_cast<AT>(y[u].v[k]);
        }
      }
  }
  // Tree over segment lanes: level j pairs l with l ^ 2^j, i.e. thread tid
  // with tid ^ (quads << j); levels inside the wavefront by lane exchange.
  for (uint32_t j = 0; j < 6 && (quads << j) < 64; j++) {
    const int m = (int)(quads << j);
#pragma unroll
    for (int k = 0; k < FPT; k++) {
      acc[0][k] = acc[0][k] + __shfl_xor(acc[0][k], m);
      if constexpr (VAR) acc[NV - 1][k] = acc[NV - 1][k] + __shfl_xor(acc[NV - 1][k], m);
    }
  }
  const uint32_t waves = quads;  // 64 * quads threads
  if (waves > 1) {
    const uint32_t w = tid / 64;
    if ((tid & 63) < quads) {
#pragma unroll
      for (int k = 0; k < FPT; k++) {
        red[w][q][0][k] = acc[0][k];
        if constexpr (VAR) red[w][q][NV - 1][k] = acc[NV - 1][k];
      }
    }
    __syncthreads();
    if (tid >= quads) return;
    // Remaining levels pair wavefront w with w ^ 1, then w ^ 2, then w ^ 4.
#pragma unroll
    for (int k = 0; k < FPT; k++)
#pragma unroll
      for (int n = 0; n < NV; n++) {
        AT x[kFinalMaxQuads];
        for (uint32_t w2 = 0; w2 < waves; w2++) x[w2] = red[w2][q][n][k];
        for (uint32_t h = 1; h < waves; h *= 2)
          for (uint32_t w2 = 0; w2 < waves; w2 += 2 * h) x[w2] = x[w2] + x[w2 + h];
        acc[n][k] = x[0];
      }
  } else if (tid >= quads) {
    return;
  }
  const double d = den[b];
#pragma unroll
  for (int k = 0; k < FPT; k++) {
    const uint32_t f = q * FPT + k;
    if (f >= nf) continue;
    if (raw) {  // sums before division (lut/integrate.rs:372-500)
      I[(uint64_t)f * nbins + b] = (double)acc[0][k];
      if constexpr (VAR) S[(uint64_t)f * nbins + b] = (double)acc[NV - 1][k];
      continue;
    }
    I[(uint64_t)f * nbins + b] = d == 0.0 ? 0.0 : (double)acc[0][k] / d;
    if constexpr (VAR) {
      const double vv = d == 0.0 ? 0.0 : (double)acc[NV - 1][k] / (d * d);
      S[(uint64_t)f * nbins + b] = sqrt(fmax(vv, 0.0));
    }
  }
}

// ------------------------------------------------------------- dispatch

template <int F, int G, int MODE, bool VAR, int FA>
void launch_tile(const TiledTable& t, const TileArgs& a, hipStream_t s) {
  using PL = Plan<F, MODE>;
  constexpr size_t lds = PL::kBytes;
  static_assert(lds <= 160 * 1024, "LDS plan exceeds one CU");
  // The attribute is per device; set it once on each.
  static std::atomic<uint64_t> done{0};
  int dev = 0;
  DIO_HIP(hipGetDevice(&dev));
  const uint64_t bit = 1ull << (dev & 63);
  if (!(done.load() & bit)) {
    DIO_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile<F, G, MODE, VAR, FA>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    done.fetch_or(bit);
  }
  uint32_t grid = std::min<uint32_t>(t.ntiles, t.cus * PL::kWgs);
  if (t.max_wgs) grid = std::min(grid, t.max_wgs);
  k_tile<F, G, MODE, VAR, FA><<<grid, PL::kThreads, lds, s>>>(a);
}

template <int F, int G, int MODE, bool VAR, int FA, bool MASKED = false>
void launch_tile_db(const TiledTable& t, const TileArgs& a, hipStream_t s) {
  using PL = PlanDb<F, MODE>;
  constexpr size_t lds = PL::kBytes;
  static std::atomic<uint64_t> done{0};
  int dev = 0;
  DIO_HIP(hipGetDevice(&dev));
  const uint64_t bit = 1ull << (dev & 63);
  if (!(done.load() & bit)) {
    DIO_HIP(hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k_tile_db<F, G, MODE, VAR, FA, MASKED>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    done.fetch_or(bit);
  }
  uint32_t grid = std::min<uint32_t>(t.ntiles, t.cus);
  if (t.max_wgs) grid = std::min(grid, t.max_wgs);
  k_tile_db<F, G, MODE, VAR, FA, MASKED><<<grid, kDbThreads, lds, s>>>(a);
}

// Masked rows (TileArgs::mbits): 16-row passes, frames-per-lane G, float32
// weights, with or without the variance sums.
template <int G, bool VAR>
void launch_masked(const TiledTable& t, const TileArgs& a, int precision, hipStream_t s) {
  auto go = [&](auto mode) {
    constexpr int M = decltype(mode)::value;
    if (a.dtype == DIO_U16) launch_tile_db<16, G, M, VAR, 2, true>(t, a, s);
    else if (a.aligned8) launch_tile_db<16, G, M, VAR, 8, true>(t, a, s);
    else launch_tile_db<16, G, M, VAR, 4, true>(t, a, s);
  };
  switch (precision) {
    case DIO_F32: go(std::integral_constant<int, 0>{}); break;
    case DIO_F32_COMPENSATED: go(std::integral_constant<int, 1>{}); break;
    case DIO_F32_WIDE: go(std::integral_constant<int, 3>{}); break;
    default: fail(DIO_ERR_INVALID, "masked rows need float32 weights");
  }
}

template <int F, int G, int MODE>
void launch_v(const TiledTable& t, const TileArgs& a, bool var, hipStream_t s) {
  const bool w32 = a.dtype == DIO_U32 || a.dtype == DIO_I32 || a.dtype == DIO_F32T;
  const int fa = !w32 ? 0 : a.aligned8 ? 8 : 4;
  if constexpr (F == 16 && MODE != 2) {
    // Pair loads of the double-buffered kernel: 8 or 4 bytes for 32-bit
    // frames, 2 or 1 for 16-bit ones.
    const int da = a.dtype == DIO_U16 ? (a.aligned8 ? 2 : 1) : fa;
    if (t.db && da != 0) {
      auto go = [&](auto v) {
        constexpr bool V = decltype(v)::value;
        switch (da) {
          case 8: launch_tile_db<F, G, MODE, V, 8>(t, a, s); break;
          case 4: launch_tile_db<F, G, MODE, V, 4>(t, a, s); break;
          case 2: launch_tile_db<F, G, MODE, V, 2>(t, a, s); break;
          default: launch_tile_db<F, G, MODE, V, 1>(t, a, s); break;
        }
      };
      if (var) go(std::true_type{});
      else go(std::false_type{});
      return;
    }
  }
  if (var) {
    if (fa == 8) launch_tile<F, G, MODE, true, 8>(t, a, s);
    else if (fa == 4) launch_tile<F, G, MODE, true, 4>(t, a, s);
    else launch_tile<F, G, MODE, true, 0>(t, a, s);
  } else {
    if (fa == 8) launch_tile<F, G, MODE, false, 8>(t, a, s);
    else if (fa == 4) launch_tile<F, G, MODE, false, 4>(t, a, s);
    else launch_tile<F, G, MODE, false, 0>(t, a, s);
  }
}

template <int F, int G>
void launch_m(const TiledTable& t, const TileArgs& a, bool var, int precision, hipStream_t s) {
  switch (precision) {
    case DIO_F32: launch_v<F, G, 0>(t, a, var, s); break;
    case DIO_F32_COMPENSATED: launch_v<F, G, 1>(t, a, var, s); break;
    case DIO_F64:
      if constexpr (F <= 8) launch_v<F, G, 2>(t, a, var, s);
      else fail(DIO_ERR_INVALID, "float64 precision allows at most 8 frames per pass");
      break;
    case DIO_F32_WIDE: launch_v<F, G, 3>(t, a, var, s); break;
    default: fail(DIO_ERR_INVALID, "unknown precision");
  }
}

// npass > 1 (double-buffered kernel only): npass passes in one launch, the
// last of nf frames and the others of F; pass h reads frames h * F onwards
// and writes its partials at part + h * pblock bytes, in segment rows of prow
// elements (0: F).
void launch_pass(const TiledTable& t, uint32_t F, uint32_t G, const void* fr, int dtype,
                 uint32_t nf, bool var, int precision, void* part, hipStream_t s,
                 uint32_t npass = 1, uint64_t pblock = 0, uint32_t prow = 0) {
  // Pixel pairs aligned to their size (8 bytes for 32-bit frames, 4 for
  // 16-bit ones) in every frame.
  const uintptr_t pair = dtype == DIO_U16 ? 4 : 8;
  const uint32_t al8 =
      (t.dim2 % 2 == 0 && reinterpret_cast<uintptr_t>(fr) % pair == 0) ? 1u : 0u;
  TileArgs a{fr,     dtype,   t.lidx.p, t.w32.p, t.w64.p, t.seg_group.p,
             t.tile_meta.p, t.nseg, t.ntx, t.ntiles, t.dim1, t.dim2,
             (uint64_t)t.dim1 * t.dim2, nf, al8, part};
  if (npass < 1 || npass > kMaxPassesPerLaunch || nf < 1 || nf > F)
    fail(DIO_ERR_INVALID, "invalid passes per launch");
  a.npass = npass;
  a.nf_last = nf;
  a.flip = t.flip ? 1u : 0u;
  a.nf = npass > 1 ? F : nf;
  a.pblock = pblock;
  a.prow = prow;
  if (npass > 1 && !double_buffered(t, F, dtype, precision))
    fail(DIO_ERR_INVALID, "several passes per launch need the double-buffered kernel");
  if (precision == DIO_F64 && !t.w64.p) fail(DIO_ERR_INVALID, "plan has no float64 weights");
  switch (F * 8 + G) {
    case 1 * 8 + 1: launch_m<1, 1>(t, a, var, precision, s); break;
    case 2 * 8 + 2: launch_m<2, 2>(t, a, var, precision, s); break;
    case 4 * 8 + 4: launch_m<4, 4>(t, a, var, precision, s); break;
    case 8 * 8 + 2: launch_m<8, 2>(t, a, var, precision, s); break;
    case 8 * 8 + 4: launch_m<8, 4>(t, a, var, precision, s); break;
    case 16 * 8 + 1: launch_m<16, 1>(t, a, var, precision, s); break;
    case 16 * 8 + 2: launch_m<16, 2>(t, a, var, precision, s); break;
    case 16 * 8 + 4: launch_m<16, 4>(t, a, var, precision, s); break;
    default: fail(DIO_ERR_INVALID, "unsupported (frames per pass, frames per lane) pair");
  }
  DIO_HIP(hipGetLastError());
}

// Reduces ntot frames whose partials are rows of F frames (frame f in column
// f % F of the block at part + (f / F) * BS elements), in groups of RF frames
// (at most 32), one launch for all groups. raw: the sums before division
// instead of I and sigma.
template <bool PLAIN, bool VAR, class PT>
void launch_final(const TiledTable& t, const void* part, uint32_t F, uint64_t BS, uint32_t ntot,
                  uint32_t RF, bool raw, double* I, double* S, hipStream_t s) {
  const PT* p = static_cast<const PT*>(part);
  // F: frames per row (a pass's F, or a launch's passes times F). A quad of
  // fpt frames stays inside one row when fpt divides F and RF.
  const uint32_t fpt = std::min(F, 4u);
  const uint32_t nf = std::min(ntot, RF);
  unsigned quads = 1;
  while (quads * fpt < nf) quads *= 2;
  if (quads > kFinalMaxQuads || F % fpt || RF % fpt)
    fail(DIO_ERR_INVALID, "invalid reduction shape");
  const unsigned threads = kFinalLanes * quads;
  const dim3 grid(t.nbins, (ntot + RF - 1) / RF);
  if (fpt == 4)
    k_tfinal<PLAIN, VAR, PT, 4><<<grid, threads, 0, s>>>(p, t.nseg, F, BS, t.bin_ptr.p,
                                                          t.bin_seg.p, t.den, t.nbins, ntot, RF,
                                                          raw ? 1 : 0, I, S);
  else if (fpt == 2)
    k_tfinal<PLAIN, VAR, PT, 2><<<grid, threads, 0, s>>>(p, t.nseg, F, BS, t.bin_ptr.p,
                                                          t.bin_seg.p, t.den, t.nbins, ntot, RF,
                                                          raw ? 1 : 0, I, S);
  else
    k_tfinal<PLAIN, VAR, PT, 1><<<grid, threads, 0, s>>>(p, t.nseg, F, BS, t.bin_ptr.p,
                                                          t.bin_seg.p, t.den, t.nbins, ntot, RF,
                                                          raw ? 1 : 0, I, S);
}

}  // namespace

uint32_t max_frames_for(int precision) { return precision == DIO_F64 ? 8 : kMaxFrames; }

bool valid_tile(int precision, uint32_t F, uint32_t G) {
  if (F > max_frames_for(precision)) return false;
  switch (F * 8 + G) {
    case 1 * 8 + 1: case 2 * 8 + 2: case 4 * 8 + 4: case 8 * 8 + 2: case 8 * 8 + 4:
    case 16 * 8 + 1: case 16 * 8 + 2: case 16 * 8 + 4:
      return true;
    default:
      return false;
  }
}
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
  t.tile_seg.alloc(t.ntiles + 1);
  t.bin_ptr.alloc(nbins + 1);
  if (nnz == 0) {
    t.nseg = 0;
    t.ngroups = 0;
    t.lidx.alloc(kGroup);
    t.w32.alloc(kGroup);
    if (keep_w64) t.w64.alloc(kGroup);
    t.seg_group.alloc(1);
    t.bin_seg.alloc(1);
    t.seg_rank.alloc(1);
    t.tile_meta.alloc(4 * (size_t)t.ntiles);
    DIO_HIP(hipMemsetAsync(t.tile_meta.p, 0, 16 * (size_t)t.ntiles, s));
    DIO_HIP(hipMemsetAsync(t.seg_group.p, 0, 4, s));
    DIO_HIP(hipMemsetAsync(t.tile_seg.p, 0, (t.ntiles + 1) * 4, s));
    DIO_HIP(hipMemsetAsync(t.bin_ptr.p, 0, (nbins + 1) * 4, s));
    DIO_HIP(hipStreamSynchronize(s));
    return;
  }
  // 1. Entries keyed by (tile, bin, pixel): stable sort of the CSR order by tile.
  DevBuf<uint32_t> tile0(nnz), bin0(nnz), order0(nnz), order(nnz), tkey(nnz), tile1(nnz),
      bin1(nnz);
  DevBuf<uint16_t> lidx0(nnz), lidx1(nnz);
  DevBuf<float> w32_1(nnz);
  DevBuf<double> w64_1(keep_w64 ? nnz : 0);
  k_entry_keys<<<blocks_for(nnz, kThreads), kThreads, 0, s>>>(row_ptr, nbins, col, nnz, dim2,
                                                             t.ntx, tile0.p, bin0.p, lidx0.p);
  k_iota32<<<blocks_for(nnz, kThreads), kThreads, 0, s>>>(order0.p, (uint32_t)nnz);
  sort_pairs(tile0.p, tkey.p, order0.p, order.p, nnz, bits_for(t.ntiles), s, tmp);
  k_tile_gather<<<blocks_for(nnz, kThreads), kThreads, 0, s>>>(
      order.p, tile0.p, bin0.p, lidx0.p, w32, keep_w64 ? w64 : nullptr, nnz, tile1.p, bin1.p,
      lidx1.p, w32_1.p, keep_w64 ? w64_1.p : nullptr);
  // 2. Runs of equal (tile, bin), cut into segments of at most kMaxSegGroups
  // groups.
  DevBuf<uint32_t> flag(nnz + 1), segid(nnz + 1);
  uint32_t nseg = 0;
  {
    DevBuf<uint32_t> rflag(nnz + 1), runid(nnz + 1);
    k_seg_flags<<<blocks_for(nnz + 1, kThreads), kThreads, 0, s>>>(tile1.p, bin1.p, nnz,
                                                                  rflag.p);
    scan_u32(rflag.p, runid.p, nnz + 1, s, tmp);
    uint32_t nrun = 0;
    DIO_HIP(hipMemcpyAsync(&nrun, runid.p + nnz, 4, hipMemcpyDeviceToHost, s));
    DIO_HIP(hipStreamSynchronize(s));
    DevBuf<uint32_t> run_first(nrun), run_tile(nrun), run_bin(nrun);
    k_seg_fill<<<blocks_for(nnz, kThreads), kThreads, 0, s>>>(rflag.p, runid.p, tile1.p,
                                                             bin1.p, nnz, run_first.p,
                                                             run_tile.p, run_bin.p);
    k_piece_flags<<<blocks_for(nnz + 1, kThreads), kThreads, 0, s>>>(rflag.p, runid.p,
                                                                    run_first.p, nnz, flag.p);
    scan_u32(flag.p, segid.p, nnz + 1, s, tmp);
    DIO_HIP(hipMemcpyAsync(&nseg, segid.p + nnz, 4, hipMemcpyDeviceToHost, s));
    DIO_HIP(hipStreamSynchronize(s));
  }
  t.nseg = nseg;
  DevBuf<uint32_t> seg_first(nseg + 1), seg_tile(nseg), seg_bin(nseg);
  k_seg_fill<<<blocks_for(nnz, kThreads), kThreads, 0, s>>>(flag.p, segid.p, tile1.p, bin1.p,
                                                           nnz, seg_first.p, seg_tile.p,
                                                           seg_bin.p);
  k_set1<<<1, 1, 0, s>>>(seg_first.p + nseg, (uint32_t)nnz);
  // 3. Inside each tile, longest segment first.
  DevBuf<uint64_t> skey(nseg), skey2(nseg);
  DevBuf<uint32_t> sid(nseg), perm(nseg), inv(nseg), ngr(nseg + 1), tile_n(nseg), bin_n(nseg);
  k_seg_keys<<<blocks_for(nseg, kThreads), kThreads, 0, s>>>(seg_first.p, seg_tile.p, nseg,
                                                            skey.p);
  k_iota32<<<blocks_for(nseg, kThreads), kThreads, 0, s>>>(sid.p, nseg);
  sort_pairs(skey.p, skey2.p, sid.p, perm.p, nseg, 32 + bits_for(t.ntiles), s, tmp);
  k_seg_perm<<<blocks_for(nseg + 1, kThreads), kThreads, 0, s>>>(
      perm.p, seg_first.p, seg_tile.p, seg_bin.p, nseg, inv.p, ngr.p, tile_n.p, bin_n.p);
  t.seg_group.alloc(nseg + 1);
  scan_u32(ngr.p, t.seg_group.p, nseg + 1, s, tmp);
  uint32_t ngroups = 0;
  DIO_HIP(hipMemcpyAsync(&ngroups, t.seg_group.p + nseg, 4, hipMemcpyDeviceToHost, s));
  DIO_HIP(hipStreamSynchronize(s));
  t.ngroups = ngroups;
  // 4. Entries into their padded groups.
  uint64_t nent = (uint64_t)ngroups * kGroup;
  t.lidx.alloc(nent);
  t.w32.alloc(nent);
  if (keep_w64) t.w64.alloc(nent);
  k_pad_fill<<<blocks_for(nent, kThreads), kThreads, 0, s>>>(nent, t.lidx.p, t.w32.p,
                                                            keep_w64 ? t.w64.p : nullptr);
  k_place<<<blocks_for(nnz, kThreads), kThreads, 0, s>>>(
      flag.p, segid.p, seg_first.p, inv.p, t.seg_group.p, nnz, lidx1.p, w32_1.p,
      keep_w64 ? w64_1.p : nullptr, t.lidx.p, t.w32.p, keep_w64 ? t.w64.p : nullptr);
  // 5. Tile -> first segment.
  DevBuf<uint32_t> hist(std::max(t.ntiles, nbins) + 1);
  DIO_HIP(hipMemsetAsync(hist.p, 0, (t.ntiles + 1) * 4, s));
  k_count_u32<<<blocks_for(nseg, kThreads), kThreads, 0, s>>>(tile_n.p, nseg, hist.p);
  scan_u32(hist.p, t.tile_seg.p, t.ntiles + 1, s, tmp);
  t.tile_meta.alloc(4 * (size_t)t.ntiles);
  k_tile_meta<<<blocks_for(t.ntiles, kThreads), kThreads, 0, s>>>(t.tile_seg.p, t.seg_group.p,
                                                                 t.ntiles, t.tile_meta.p);
  // 6. Bin -> segments in tile order (stable sort of the tile-major ids by bin).
  DevBuf<uint32_t> bkey(nseg);
  t.bin_seg.alloc(nseg);
  sort_pairs(bin_n.p, bkey.p, sid.p, t.bin_seg.p, nseg, bits_for(nbins), s, tmp);
  DIO_HIP(hipMemsetAsync(hist.p, 0, (nbins + 1) * 4, s));
  k_count_u32<<<blocks_for(nseg, kThreads), kThreads, 0, s>>>(bin_n.p, nseg, hist.p);
  scan_u32(hist.p, t.bin_ptr.p, nbins + 1, s, tmp);
  t.seg_rank.alloc(nseg);
  k_invert32<<<blocks_for(nseg, kThreads), kThreads, 0, s>>>(t.bin_seg.p, nseg, t.seg_rank.p);
  DIO_HIP(hipGetLastError());
  std::vector<uint32_t> meta(4 * (size_t)t.ntiles);
  DIO_HIP(hipMemcpyAsync(meta.data(), t.tile_meta.p, meta.size() * 4, hipMemcpyDeviceToHost, s));
  DIO_HIP(hipStreamSynchronize(s));
  t.max_tile_groups = 0;
  for (uint32_t k = 0; k < t.ntiles; k++)
    t.max_tile_groups = std::max(t.max_tile_groups, meta[4 * k + 3] - meta[4 * k + 2]);
}

void tiled_pass_only(const TiledTable& t, TileScratch& sc, const void* frames, int dtype,
                     uint32_t nf, bool var, int precision, uint32_t F, uint32_t G,
                     hipStream_t s) {
  if (!valid_tile(precision, F, G) || nf == 0 || nf > F)
    fail(DIO_ERR_INVALID, "invalid frames per pass / frames per lane");
  (void)dtype_bytes(dtype);
  sc.ensure(t, F, var);
  if (t.nseg) launch_pass(t, F, G, frames, dtype, nf, var, precision, sc.ptr(), s);
}

bool double_buffered(const TiledTable& t, uint32_t F, int dtype, int precision) {
  const bool w = dtype == DIO_U32 || dtype == DIO_I32 || dtype == DIO_F32T || dtype == DIO_U16;
  return F == 16 && precision != DIO_F64 && t.db && w;
}

size_t pass_bytes(const TiledTable& t, uint32_t F, bool var, int precision) {
  return (size_t)t.nseg * F * (var ? 2 : 1) * (precision == DIO_F64 ? sizeof(double)
                                                                     : sizeof(float));
}

uint32_t passes_per_launch(const TiledTable& t, uint32_t F, int dtype, int precision) {
  if (precision == DIO_F64) return 1;
  return double_buffered(t, F, dtype, precision) ? kMaxPassesPerLaunch : 2;
}

void tiled_gather(const TiledTable& t, TileScratch& sc, const void* frames, int dtype,
                  uint64_t n_frames, bool var, int precision, uint32_t F, uint32_t G,
                  double* intensity, double* sigma, hipStream_t s, LaunchEvents* le,
                  bool raw, uint32_t max_passes) {
  if (!valid_tile(precision, F, G))
    fail(DIO_ERR_INVALID, "invalid frames per pass / frames per lane");
  const size_t esz = dtype_bytes(dtype);
  const char* fb = static_cast<const char*>(frames);
  const uint64_t npix = (uint64_t)t.dim1 * t.dim2;
  const bool plain = precision == DIO_F32, f64 = precision == DIO_F64;
  // The scratch holds the float32 partials of kMaxPassesPerLaunch passes (or
  // the float64 ones of one). The double-buffered kernel runs that many
  // passes in one launch over one read of the table, each segment's partials
  // of all of them in one row (P * F columns, so the reduction reads whole
  // 128-byte pieces of a row); otherwise a launch is a pass, with a block of
  // rows per pass. A reduction launch serves up to 32 frames (two 16-frame
  // passes). The frames' arithmetic does not depend on the passes per launch
  // or on the layout.
  const bool db = !f64 && double_buffered(t, F, dtype, precision);
  uint32_t P = passes_per_launch(t, F, dtype, precision);
  const size_t pbytes = pass_bytes(t, F, var, precision);
  if (max_passes) {
    P = std::min(P, max_passes);
    if (!sc.ext || sc.ext_n * sizeof(double) < (size_t)P * pbytes)
      fail(DIO_ERR_INVALID, "external tile scratch is too small");
  } else {
    sc.ensure(t, F, var);
  }
  const uint32_t FR = F * P;
  const uint32_t RF = std::min(FR, 32u);  // frames per reduction workgroup
  const uint64_t BS = (uint64_t)t.nseg * F * (var ? 2 : 1);  // elements per pass block
  auto timed = [&](uint32_t frames_in_launch, auto&& launch) {
    const bool rec = le && le->n < le->cap;
    if (rec) DIO_HIP(hipEventRecord(le->ev[2 * le->n], s));
    if (t.nseg) launch();
    if (rec) {
      DIO_HIP(hipEventRecord(le->ev[2 * le->n + 1], s));
      if (le->frames) le->frames[le->n] = frames_in_launch;
      le->n++;
    }
  };
  for (uint64_t g0 = 0; g0 < n_frames; g0 += FR) {
    const uint32_t ng = (uint32_t)std::min<uint64_t>(FR, n_frames - g0);
    if (db) {
      const uint32_t np = (ng + F - 1) / F;
      timed(ng, [&] {
        launch_pass(t, F, G, fb + g0 * npix * esz, dtype, ng - (np - 1) * F, var, precision,
                    sc.ptr(), s, np, (uint64_t)F * sizeof(float), FR);
      });
    } else {
      for (uint32_t j = 0; j * F < ng; j++) {
        const uint64_t f0 = g0 + (uint64_t)j * F;
        const uint32_t nf = std::min(F, ng - j * F);
        timed(nf, [&] {
          launch_pass(t, F, G, fb + f0 * npix * esz, dtype, nf, var, precision,
                      reinterpret_cast<char*>(sc.ptr()) + j * pbytes, s);
        });
      }
    }
    {
      const void* pr = sc.ptr();
      const uint32_t FS = db ? FR : F;
      double* I = intensity + g0 * t.nbins;
      double* S = var ? sigma + g0 * t.nbins : nullptr;
      if (f64) {
        if (var) launch_final<false, true, double>(t, pr, FS, BS, ng, RF, raw, I, S, s);
        else launch_final<false, false, double>(t, pr, FS, BS, ng, RF, raw, I, S, s);
      } else if (plain) {
        if (var) launch_final<true, true, float>(t, pr, FS, BS, ng, RF, raw, I, S, s);
        else launch_final<true, false, float>(t, pr, FS, BS, ng, RF, raw, I, S, s);
      } else {
        if (var) launch_final<false, true, float>(t, pr, FS, BS, ng, RF, raw, I, S, s);
        else launch_final<false, false, float>(t, pr, FS, BS, ng, RF, raw, I, S, s);
      }
    }
    DIO_HIP(hipGetLastError());
  }
}

bool masked_rows_supported(const TiledTable& t, const void* frames, int dtype, int precision,
                           uint32_t G) {
  // 16-bit frames: pixel pairs read as 4-byte words (even rows, 4-byte
  // aligned frames; passes of 8 frames keep the alignment).
  if (dtype == DIO_U16 && (t.dim2 % 2 || reinterpret_cast<uintptr_t>(frames) % 4)) return false;
  return double_buffered(t, 16, dtype, precision) && valid_tile(precision, 16, G);
}

uint32_t masked_rows_frames_per_pass() { return kMaskedFrames; }

void tiled_gather_masked(const TiledTable& t, TileScratch& sc, const void* frames, int dtype,
                         const uint8_t* mbits, uint64_t mbits_stride, const float* norm32,
                         uint64_t n_frames, bool var, int precision, uint32_t G,
                         double* rows_sum, double* rows_var, hipStream_t s) {
  if (!masked_rows_supported(t, frames, dtype, precision, G))
    fail(DIO_ERR_INVALID, "masked rows need 32-bit or pair-aligned 16-bit frames and 16-row passes");
  const uint32_t esz = (uint32_t)dtype_bytes(dtype);
  constexpr uint32_t F = 16;
  const uint64_t npix = (uint64_t)t.dim1 * t.dim2;
  const char* fb = static_cast<const char*>(frames);
  const uint32_t P = kMaxPassesPerLaunch;
  sc.ensure(t, F, var);
  // Each segment's partials of the launch's passes in one row of P * F
  // columns (one block; BS is not used).
  const uint32_t FS = P * F;
  const uint64_t BS = 0;
  const bool plain = precision == DIO_F32;
  for (uint64_t g0 = 0; g0 < n_frames; g0 += (uint64_t)P * kMaskedFrames) {
    const uint32_t ng = (uint32_t)std::min<uint64_t>((uint64_t)P * kMaskedFrames, n_frames - g0);
    const uint32_t np = (ng + kMaskedFrames - 1) / kMaskedFrames;
    const uint64_t pass0 = g0 / kMaskedFrames;
    const void* fr = fb + g0 * npix * esz;
    const uint32_t al8 =
        (t.dim2 % 2 == 0 && reinterpret_cast<uintptr_t>(fr) % (2 * esz) == 0) ? 1u : 0u;
    TileArgs a{fr,     dtype,   t.lidx.p, t.w32.p, t.w64.p, t.seg_group.p,
               t.tile_meta.p, t.nseg, t.ntx, t.ntiles, t.dim1, t.dim2,
               npix,   F,       al8,      sc.ptr()};
    a.flip = t.flip ? 1u : 0u;
    a.npass = np;
    a.nf_last = F;
    a.pblock = F * sizeof(float);
    a.prow = FS;
    a.mbits = mbits + pass0 * mbits_stride;
    a.mbits_stride = mbits_stride;
    a.norm32 = norm32;
    a.nfr = ng - (np - 1) * kMaskedFrames;
    if (t.nseg) {
      switch (G * 2 + (var ? 1 : 0)) {
        case 2: launch_masked<1, false>(t, a, precision, s); break;
        case 3: launch_masked<1, true>(t, a, precision, s); break;
        case 4: launch_masked<2, false>(t, a, precision, s); break;
        case 5: launch_masked<2, true>(t, a, precision, s); break;
        case 8: launch_masked<4, false>(t, a, precision, s); break;
        case 9: launch_masked<4, true>(t, a, precision, s); break;
        default: fail(DIO_ERR_INVALID, "unsupported frames per lane");
      }
    }
    double* I = rows_sum + pass0 * F * t.nbins;
    double* V = var ? rows_var + pass0 * F * t.nbins : nullptr;
    if (plain) {
      if (var) launch_final<true, true, float>(t, sc.ptr(), FS, BS, np * F, 32, true, I, V, s);
      else launch_final<true, false, float>(t, sc.ptr(), FS, BS, np * F, 32, true, I, V, s);
    } else {
      if (var) launch_final<false, true, float>(t, sc.ptr(), FS, BS, np * F, 32, true, I, V, s);
      else launch_final<false, false, float>(t, sc.ptr(), FS, BS, np * F, 32, true, I, V, s);
    }
    DIO_HIP(hipGetLastError());
  }
}

}  // namespace dio
