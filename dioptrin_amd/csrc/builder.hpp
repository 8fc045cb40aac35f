// Device lookup-table builder interface.
#pragma once
#include <vector>

#include "common.hpp"
#include "geometry.hpp"

namespace dio {

// A CSR table resident in device memory. Rows list their pixels in
// ascending pixel order, like the reference builders.
struct Table {
  DevBuf<uint32_t> row_ptr;  // nbins + 1
  DevBuf<uint32_t> col;      // nnz
  DevBuf<double> w64;        // nnz (builder weights, FP64)
  DevBuf<float> w32;         // nnz (RN(w64), execution storage)
  uint64_t nnz = 0;
  uint32_t nbins = 0;
};

struct Built {
  Table full;        // unmasked table (Lut1d / Lut2d)
  Table exec;        // masked execution table (empty when no pixel is masked)
  bool has_exec = false;
  DevBuf<double> norm;   // per pixel polarization * solid angle
  DevBuf<double> den;    // per bin strict FP64 denominators of the execution table
  DevBuf<uint8_t> mask;  // per pixel, only when a mask was given
  bool weighted = true;
  uint32_t n_radial = 0, n_azimuthal = 1;
  double hist_min = 0, hist_step = 0;              // radial axis
  double az_min = 0, az_step = 0;                  // azimuthal axis (2D)
  uint64_t uncertain = 0;
  uint32_t singular = 0;
  const Table& exec_table() const { return has_exec ? exec : full; }
};

Geo make_geo(const dio_poni& p, const dio_plan_desc& d);
void build_tables(const Geo& geo, const dio_plan_desc& d, hipStream_t s, Built& out);
// Per bin of the execution table: sum of its weights in row order (entry
// count for unweighted tables), enqueued on s.
void bin_counts(const Built& b, double* cnt, hipStream_t s);

}  // namespace dio
