// Plan table views and the azimuthal-spread kernel on gfx950.
#pragma once
#include "common.hpp"
#include "tile.hpp"

namespace dio {

// Device pointers and sizes of a plan's execution CSR table.
struct GatherTable {
  const uint32_t* col = nullptr;
  const float* w32 = nullptr;
  const double* w64 = nullptr;
  const double* den = nullptr;
  uint32_t nbins = 0;
  uint64_t npix = 0;
  uint64_t nnz = 0;
};

struct StdTable {
  const uint32_t* row_ptr = nullptr;  // unmasked table
  const uint32_t* col = nullptr;
  const float* w32 = nullptr;
  const double* w64 = nullptr;        // FP64 weights
  int weighted = 1;                   // 0: unit weights (counting tables)
  const uint8_t* mask = nullptr;      // may be null
  const double* norm = nullptr;
  uint32_t nbins = 0;
  uint64_t npix = 0;
  uint32_t dim1 = 0, dim2 = 0;
  bool flip = false;  // frames hold the detector's rows bottom-up
};

// azimuthal.rs:48-118 for n_frames device frames on the detector-tiled
// table (std_tile.hip): intensity, sigma (when errors) and the weighted
// population spread of every bin. norm is the plan's pixel normalization
// (detector order). The segment partials of a pass go to scratch
// (scratch_bytes): passes hold std_frames_per_pass() frames, fewer when the
// scratch holds fewer; less than one frame's partials is refused. A frame's
// result does not depend on the pass.
size_t std_pass_bytes(const TiledTable& t, uint32_t frames, bool errors);
uint32_t std_frames_per_pass();
void azimuthal_std_tiled(const TiledTable& t, const double* norm, void* scratch,
                         size_t scratch_bytes, const void* frames, int dtype, uint64_t n_frames,
                         bool errors, double* intensity, double* sigma, double* std_out,
                         hipStream_t s);

// Dynamic masks (Dynamic1dPlan::integrate, integration/dynamic.rs:143-186;
// prepared.rs:1011-1121 with a per-frame mask): n_frames device frames, each
// with its own device mask (nonzero = masked), on the unmasked table; the
// plan's own mask is not applied. Per bin: I = sum w*x / sum w*n and
// sigma = sqrt(max(sum w^2*max(x, 0) / (sum w*n)^2, 0)) over unmasked
// entries, 0 where the denominator is 0.
void masked_frames(const StdTable& t, const void* frames, int dtype, const uint8_t* masks,
                   uint64_t n_frames, bool errors, double* intensity, double* sigma,
                   hipStream_t s);

// Dynamic masks on a plan whose execution table is the unmasked table: the
// masked frames are turned into two frames each for the tile kernel's raw
// sums, the signal with masked pixels set to 0 (xo, the frames' dtype) and
// the pixel normalization with masked pixels set to 0 (no, float64 when
// norm_f64, else float32), both in input order; masked_finish then divides:
// I = S_x / S_n, sigma = sqrt(max(S_v / S_n^2, 0)), 0 where S_n is 0.
void masked_prepare(const StdTable& t, const void* frames, int dtype, const uint8_t* masks,
                    uint64_t n_frames, bool norm_f64, void* xo, void* no, hipStream_t s);
void masked_finish(double* intensity, double* sigma, const double* den, uint64_t n, bool errors,
                   hipStream_t s);

// Masks folded into the tile kernel's staging (tiled_gather_masked): the
// masks' bit bytes per 8-frame pass ([pass][npix]), the normalization as
// float32, and the division of the passes' row sums: I = sum w x / D,
// sigma = sqrt(max(V / D^2, 0)), 0 where D is 0 (as masked_finish).
// Planes of mask_bits are mask_bits_stride(npix) bytes apart (a multiple of
// 16, out must be 16-byte aligned).
uint64_t mask_bits_stride(uint64_t npix);
void mask_bits(const uint8_t* masks, uint64_t npix, uint32_t n_frames, uint8_t* out,
               hipStream_t s);
void norm_to_f32(const double* norm, uint64_t n, float* out, hipStream_t s);
void masked_rows_finish(const double* rows_sum, const double* rows_var, uint64_t n_frames,
                        uint32_t nbins, double* intensity, double* sigma, hipStream_t s);

}  // namespace dio
