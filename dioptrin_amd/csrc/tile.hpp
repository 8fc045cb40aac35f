// Detector-tiled execution table and its gather (gfx950).
#pragma once
#include "common.hpp"

namespace dio {

// The execution table re-grouped by detector tile: tile = kTileRows x
// kTileCols pixels; inside a tile, entries are grouped by bin (a segment per
// (tile, bin)), pixels ascending as in the CSR rows, each segment padded to a
// whole number of 4-entry groups with zero-weight entries that point at an
// all-zero LDS slot. A frame tile is read from HBM once, coalesced, into LDS,
// and every entry of the table tile is applied to it; segment partials are
// added per bin in tile order.
constexpr uint32_t kTileCols = 64;
constexpr uint32_t kTileRows = 16;
constexpr uint32_t kTilePixels = kTileCols * kTileRows;
constexpr uint32_t kPadSlot = kTilePixels;          // LDS index of the zero pixel
// Per-frame LDS row: the pixels, the zero slot and one spare word. The stride
// is 2 mod 32 words, so the frames a segment's lanes read at one pixel fall in
// distinct banks.
constexpr uint32_t kLdsStride = kTilePixels + 2;
constexpr uint32_t kGroup = 4;                      // entries per table group
constexpr uint32_t kMaxFrames = 16;
// Longest segment, in groups: a (tile, bin) run beyond it is cut into several
// segments, so no lane of a tile walks far more groups than the others.
constexpr uint32_t kMaxSegGroups = 12;
// Passes over one read of the table in a launch of the double-buffered kernel.
constexpr uint32_t kMaxPassesPerLaunch = 4;

struct TiledTable {
  uint32_t dim1 = 0, dim2 = 0, nbins = 0, ntx = 0, nty = 0, ntiles = 0;
  uint32_t cus = 256;          // compute units of the device (persistent grid size)
  uint32_t max_wgs = 0;        // cap on the persistent grid, 0 for none
  bool db = true;              // double-buffered kernel for 16-frame passes of 16/32-bit frames
  bool flip = false;           // frames hold the detector's rows bottom-up
  uint64_t nnz = 0;
  uint32_t nseg = 0, ngroups = 0;
  uint32_t max_tile_groups = 0;  // largest table of one tile, in groups
  DevBuf<uint16_t> lidx;       // 4 * ngroups: pixel index inside its tile (kPadSlot = padding)
  DevBuf<float> w32;           // 4 * ngroups
  DevBuf<double> w64;          // 4 * ngroups (F64 precision only)
  DevBuf<uint32_t> seg_group;  // nseg + 1: first group of each segment
  DevBuf<uint32_t> tile_seg;   // ntiles + 1: first segment of each tile
  DevBuf<uint32_t> tile_meta;  // 4 * ntiles: first, end segment; first, end group
  DevBuf<uint32_t> bin_seg;    // nseg: segments sorted by (bin, tile)
  DevBuf<uint32_t> bin_ptr;    // nbins + 1
  DevBuf<uint32_t> seg_rank;   // nseg: position of each segment in bin_seg
  const double* den = nullptr;
  size_t bytes() const {
    return lidx.bytes() + w32.bytes() + w64.bytes() + seg_group.bytes() + tile_seg.bytes() +
           tile_meta.bytes() + bin_seg.bytes() + bin_ptr.bytes() + seg_rank.bytes();
  }
};

// Builds the tiled layout from a CSR execution table resident on the device.
void build_tiled(TiledTable& t, uint32_t dim1, uint32_t dim2, uint32_t nbins, uint64_t nnz,
                 const uint32_t* row_ptr, const uint32_t* col, const float* w32,
                 const double* w64, bool keep_w64, hipStream_t s);

// Segment partials, float32 (float64 for the F64 precision), for up to
// kMaxPassesPerLaunch passes of float32 partials (or one pass of float64
// ones): per pass a block [1 + errors][nseg][frames per pass], or, for the
// passes of one launch of the double-buffered kernel, one block
// [1 + errors][nseg][passes x frames per pass] (the same size).
// Either owned (grown on demand) or external memory of a stated size (the
// serving loop's stream-ordered slot buffers).
struct TileScratch {
  DevBuf<double> part;
  double* ext = nullptr;
  size_t ext_n = 0;
  static size_t need(const TiledTable& t, uint32_t frames, bool var) {
    return (size_t)t.nseg * frames * (var ? 2 : 1) * (kMaxPassesPerLaunch / 2);
  }
  double* ptr() const { return ext ? ext : part.p; }
  size_t bytes() const { return (ext ? ext_n : part.n) * sizeof(double); }
  // At least b bytes of owned scratch (external memory is used as it is).
  void ensure_bytes(size_t b) {
    if (ext) return;
    const size_t n = (b + sizeof(double) - 1) / sizeof(double);
    if (part.n < n) part.alloc(n);
  }
  void ensure(const TiledTable& t, uint32_t frames, bool var) {
    const size_t n = need(t, frames, var);
    if (ext) {
      if (ext_n < n) fail(DIO_ERR_DEVICE, "external tile scratch is too small");
      return;
    }
    if (part.n < n) part.alloc(n);
  }
};

// Largest frames per pass the LDS holds for this precision.
uint32_t max_frames_for(int precision);
// Valid (frames per pass, frames per lane) pair for this precision.
bool valid_tile(int precision, uint32_t frames, uint32_t lanes);

// Events recorded around each launch of the tile kernel (measurement): 2 per
// launch in ev (room for cap launches); n counts the launches, frames[k] (if
// given) receives the frames of launch k.
struct LaunchEvents {
  const hipEvent_t* ev = nullptr;
  uint64_t cap = 0;
  uint64_t n = 0;
  uint32_t* frames = nullptr;
};

// True when 16-frame passes of this frame type and precision run the
// double-buffered kernel, which takes two passes per launch.
bool double_buffered(const TiledTable& t, uint32_t frames_per_pass, int dtype, int precision);

// Bytes of segment partials one pass of frames_per_pass frames writes.
size_t pass_bytes(const TiledTable& t, uint32_t frames_per_pass, bool var, int precision);

// Passes per launch of the gather (several share one read of the table in the
// double-buffered kernel; partials of that many passes are held at once).
uint32_t passes_per_launch(const TiledTable& t, uint32_t frames_per_pass, int dtype,
                           int precision);

// max_passes > 0 caps the passes per launch and uses sc's external memory
// only (it must hold max_passes pass blocks): the gather then allocates
// nothing.
void tiled_gather(const TiledTable& t, TileScratch& sc, const void* frames, int dtype,
                  uint64_t n_frames, bool var, int precision, uint32_t frames_per_pass,
                  uint32_t frames_per_lane, double* intensity, double* sigma, hipStream_t s,
                  LaunchEvents* launches = nullptr, bool raw = false, uint32_t max_passes = 0);

// Per-frame masks folded into the double-buffered kernel's staging: passes
// of 16 rows, rows 0-7 the pass's 8 frames with masked pixels read as 0 and
// rows 8-15 the plan's float32 normalization (norm32, detector order) with
// the same pixels set to 0. mbits: one byte per pixel and pass (bit i:
// frame i of the pass masked), input order, passes mbits_stride bytes apart.
// rows_sum / rows_var receive, per pass, the 16 rows' raw sums (sum w x and
// sum w^2 max(x, 0), the latter only with var; the normalization rows' sums
// are the frames' denominators) as [pass][16][nbins]. Needs 32-bit frames, or
// 16-bit ones with even rows starting 4-byte aligned, and float32 weights
// (masked_rows_supported).
bool masked_rows_supported(const TiledTable& t, const void* frames, int dtype, int precision,
                           uint32_t lanes);
uint32_t masked_rows_frames_per_pass();
void tiled_gather_masked(const TiledTable& t, TileScratch& sc, const void* frames, int dtype,
                         const uint8_t* mbits, uint64_t mbits_stride, const float* norm32,
                         uint64_t n_frames, bool var, int precision, uint32_t lanes,
                         double* rows_sum, double* rows_var, hipStream_t s);

// Only the tile kernel for one pass of nf <= frames_per_pass frames (timing).
void tiled_pass_only(const TiledTable& t, TileScratch& sc, const void* frames, int dtype,
                     uint32_t nf, bool var, int precision, uint32_t frames_per_pass,
                     uint32_t frames_per_lane, hipStream_t s);

}  // namespace dio
