// Per-frame and batched CSR gather on gfx950.
#pragma once
#include "common.hpp"

namespace dio {

// Gather work units: every row of the execution table is cut into segments
// of at most kSegLen consecutive entries; one wavefront reduces one segment
// for a tile of frames. Segment boundaries depend only on the table, so the
// per-frame arithmetic (and its bits) is the same for any tile size.
constexpr uint32_t kSegLen = 1024;
constexpr uint32_t kMaxTile = 32;

struct GatherTable {
  const uint32_t* col = nullptr;
  const float* w32 = nullptr;
  const double* w64 = nullptr;
  const double* den = nullptr;
  uint32_t nbins = 0;
  uint64_t npix = 0;
  uint64_t nnz = 0;
  DevBuf<uint32_t> seg_begin, seg_end, row_seg;  // row_seg: nbins + 1
  uint32_t nseg = 0;
};

struct StdTable {
  const uint32_t* row_ptr = nullptr;  // unmasked table
  const uint32_t* col = nullptr;
  const float* w32 = nullptr;
  const uint8_t* mask = nullptr;      // may be null
  const double* norm = nullptr;
  uint32_t nbins = 0;
  uint64_t npix = 0;
};

void make_segments(GatherTable& t, const uint32_t* d_row_ptr, hipStream_t s);

// Partial-sum scratch for `tile` frames (signal and variance).
struct GatherScratch {
  DevBuf<double> part;
  void ensure(const GatherTable& t, uint32_t tile, bool var) {
    size_t need = (size_t)t.nseg * tile * (var ? 2 : 1);
    if (part.n < need) part.alloc(need);
  }
};

// Integrates n_frames device frames in tiles of `tile` frames.
void gather_frames(const GatherTable& t, GatherScratch& sc, const void* frames, int dtype,
                   uint64_t n_frames, bool var, int precision, uint32_t tile,
                   double* intensity, double* sigma, hipStream_t s);

// Only the segment kernel of one tile (for kernel timing).
void gather_segments_only(const GatherTable& t, GatherScratch& sc, const void* frames, int dtype,
                          uint32_t nf, bool var, int precision, hipStream_t s);

void azimuthal_std_frames(const StdTable& t, const void* frames, int dtype, uint64_t n_frames,
                          bool errors, double* intensity, double* sigma, double* std_out,
                          hipStream_t s);

}  // namespace dio
