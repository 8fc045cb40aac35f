// Plan table views and the azimuthal-spread kernel on gfx950.
#pragma once
#include "common.hpp"

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
  const uint8_t* mask = nullptr;      // may be null
  const double* norm = nullptr;
  uint32_t nbins = 0;
  uint64_t npix = 0;
  uint32_t dim1 = 0, dim2 = 0;
  bool flip = false;  // frames hold the detector's rows bottom-up
};

// azimuthal.rs:48-118 for n_frames device frames: intensity, sigma (when
// errors) and the weighted population spread of every bin.
void azimuthal_std_frames(const StdTable& t, const void* frames, int dtype, uint64_t n_frames,
                          bool errors, double* intensity, double* sigma, double* std_out,
                          hipStream_t s);

}  // namespace dio
