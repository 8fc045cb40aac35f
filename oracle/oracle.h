/* FP64 CPU oracle of the Dioptrin integration hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests, the smoke
 * check and the CPU-baseline leg of the benchmark link it to compare the HIP
 * path against. Nothing in the product (dioptrin_amd/) may call it.
 *
 * It restates, in plain C with IEEE double arithmetic, no FMA contraction and
 * the host libm, the reference algorithms (paths relative to the reference
 * repository root):
 *   geometry        src/geometry.rs:229-320, 371-636, 639-718
 *   bin rules       src/math/histogram.rs:120-203, src/math/mod.rs:8-32,
 *                   src/lut/common.rs:33-65
 *   units           src/integration.rs:115-208
 *   corrections     src/integration/corrections.rs:57-93
 *   counting 1D     src/lut/counting_csr.rs:54-134
 *   pixel split 1D  src/lut/pixel_split.rs:14-455
 *   beam pixels     src/lut/pyfai_1d.rs:180-252, integration/grid.rs:48-160
 *   2D tables       src/lut/shared_grid.rs:13-232, integration/grid.rs:26-222,
 *                   579-675, 978-1028, integration/methods.rs:365-388
 *   supersampling   src/lut/supersampled.rs:21-595, 1381-1477, lut/common.rs:33-65
 *   pyFAI split     src/lut/pyfai_1d.rs:36-354, src/lut/pyfai_split.rs:35-167
 *   masked table    src/lut/prepared.rs:163-262
 *   gather          src/lut/prepared.rs:985-1121
 *   azimuthal std   src/lut/azimuthal.rs:48-118
 * Each function's comment cites the lines it follows.
 *
 * Pinning (tests/test_oracle_golden.py, tests/test_oracle_pins.py,
 * tests/test_cpu_baseline.py): geometry by pyFAI vectors and closed forms;
 * bin rules, pixel split, supersampling, 2D grids, gather and std by the
 * reference's own known answers, brute force and invariants; pyFAI 1D by
 * pyFAI golden weights and the reference's sector, renormalisation, seam and
 * wrap_tau cases; 2D placement along the azimuth by brute force per pixel
 * (exact polygon-rectangle areas for pixel split, pyFAI's raster rule for
 * pyfai_full, subpixel counts for supersampling) and the reference's sector
 * and index-rounding cases; the CSC baseline against the strict gather.
 * Parity unpinned: the ulp-level bits of the reference's builders on large
 * geometries, which depend on Rust's platform libm and fold order; no
 * reference test pins them below 1e-10 / 1e-6, so bit-exactness is claimed
 * against this oracle with the host glibc.
 */
#ifndef DIOPTRIN_ORACLE_H
#define DIOPTRIN_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double pixel1, pixel2, distance, poni1, poni2, rot1, rot2, rot3, wavelength;
} or_poni;

enum { OR_UNIT_TTH_RAD = 0, OR_UNIT_TTH_DEG = 1, OR_UNIT_Q_A = 2, OR_UNIT_Q_NM = 3 };

typedef struct {
  size_t nbins;
  size_t nnz;
  uint64_t *row_ptr; /* nbins + 1 */
  uint64_t *col;     /* nnz */
  double *w;         /* nnz, NULL for unweighted (counting) tables */
  double hist_min, hist_step;
} or_lut;

/* Sample-frame angles of detector points (d1 = row, d2 = column, in pixel
 * units): tth = atan2(sqrt(y^2 + x^2), z), chi = atan2(x, y). */
void or_tth_chi(const or_poni *p, const double *d1, const double *d2, size_t n,
                double *tth, double *chi);
/* Sample-frame coordinates (x, y, z) of detector points. */
void or_positions(const or_poni *p, const double *d1, const double *d2, size_t n,
                  double *x, double *y, double *z);
/* Corner grid angles, (dim1+1)*(dim2+1) values each. */
void or_corners(const or_poni *p, size_t dim1, size_t dim2, double *tth, double *chi);
/* Pixel-centre angles, dim1*dim2 values each. */
void or_centres(const or_poni *p, size_t dim1, size_t dim2, double *tth, double *chi);
void or_solid_angle(const or_poni *p, size_t dim1, size_t dim2, double *out);
void or_polarization(double factor, const double *tth, const double *chi, size_t n, double *out);
/* normalization = polarization * solid angle (polarization 1 if !has_pol). */
void or_normalization(const or_poni *p, size_t dim1, size_t dim2, int has_pol, double factor,
                      double *out);
void or_convert_unit(int unit, double wavelength, const double *tth, size_t n, double *out);

/* Table builders. Return 0 on success, nonzero on an invalid request. */
int or_build_lut1d_ps(const or_poni *p, size_t dim1, size_t dim2, size_t n_points, int unit,
                      int has_range, double rmin, double rmax, or_lut *out);
int or_build_lut1d_counting(const or_poni *p, size_t dim1, size_t dim2, size_t n_points,
                            int unit, int has_range, double rmin, double rmax, or_lut *out);
int or_counting_from_radial(const double *rad, size_t npix, size_t n_points, int has_range,
                            double rmin, double rmax, or_lut *out);
int or_ps_fractions(const or_poni *p, size_t dim1, size_t dim2, size_t n_points, int unit,
                    double min, double step, int clamp_low, or_lut *out);
/* Default (or explicit-range) 2D shared-grid table of Counting (method 0) or
 * PixelSplit (method 1); rows are az * n_radial + r. edges (may be NULL)
 * receives r0, r1, a0, a1 of the grid. */
int or_build_lut2d(const or_poni *p, size_t dim1, size_t dim2, int method, int unit,
                   size_t n_radial, size_t n_azimuthal, int has_rr, double rlo, double rhi,
                   int has_ar, double alo, double ahi, or_lut *out, double *edges);
/* Supersampled tables (n_ss x n_ss subpixels, weight count / n_ss^2). */
int or_build_lut1d_ss(const or_poni *p, size_t dim1, size_t dim2, size_t n_points, size_t n_ss,
                      int unit, int has_range, double rmin, double rmax, or_lut *out);
int or_build_lut2d_ss(const or_poni *p, size_t dim1, size_t dim2, size_t n_radial,
                      size_t n_azimuthal, size_t n_ss, int unit, int has_rr, double rlo,
                      double rhi, int has_ar, double alo, double ahi, or_lut *out,
                      double *edges);
/* pyFAI-compatible full splitting, 1D (pyfai_1d.rs:55-354) and 2D
 * (pyfai_split.rs:35-167 in shared_grid.rs:13-150). */
int or_build_lut1d_pyfai(const or_poni *p, size_t dim1, size_t dim2, size_t n_points, int unit,
                         int has_range, double rmin, double rmax, or_lut *out);
int or_build_lut2d_pyfai(const or_poni *p, size_t dim1, size_t dim2, int unit, size_t n_radial,
                         size_t n_azimuthal, int has_rr, double rlo, double rhi, int has_ar,
                         double alo, double ahi, or_lut *out, double *edges);
/* pyfai_1d.rs:259-354 split1d on one polygon with the grid's azimuthal range
 * [a0, a1], sector exclusion and corner sign; pyfai_1d.rs:180-186 wrap_tau. */
int or_pyfai_split1d_sector(const double *poly8, double r0, double r1, size_t nbins, double a0,
                            double a1, int sector, double sign, uint64_t *bins, double *weights,
                            size_t cap);
/* grid.rs:26-222 IntegrationGrid2d bin with clamp flags (bit 0 radial end,
 * 1 index rounding, 2 radial outside, 3 azimuthal outside); -1 none, -2
 * invalid grid. */
long long or_grid2d_bin(size_t nr, size_t na, double r0, double r1, double a0, double a1,
                        int flags, double radial, double az);
double or_wrap_tau(double x);
int or_pyfai_split1d(const double *poly8, double r0, double r1, size_t nbins, uint64_t *bins,
                     double *weights, size_t cap);
void or_lut_free(or_lut *lut);

/* Masked execution table: entries of masked pixels removed, order kept. */
void or_fixed_table(const or_lut *in, const uint8_t *mask, or_lut *out);
/* Strict per-bin denominators sum_k w_k * norm[col_k] in row order. */
void or_denominators(const or_lut *lut, const double *norm, double *den);
/* Strict gather: I = sum(w x)/den, sigma = sqrt(max(sum(w^2 max(x,0))/den^2, 0)). */
void or_gather(const or_lut *lut, const double *den, const double *signal, int with_var,
               double *intensity, double *sigma);
/* Weighted population spread of corrected pixels (unmasked table, mask per
 * entry). sigma may be NULL when errors are not requested. */
void or_azimuthal_std(const or_lut *lut, const uint8_t *mask, const double *norm,
                      const double *signal, int errors, double *intensity, double *sigma,
                      double *std_out);
/* Bin centres min + (i + 0.5) * step. */
void or_bins(size_t n, double min, double step, double *out);

/* Multithreaded restatement of the reference's default CPU batch1d path:
 * pixel-major CSC scatter into 4 lane histograms, 4-frame tiles, frames over
 * a thread pool (src/lut/csc.rs:198-700, src/python/integrator.rs:715-760).
 * Returns an opaque handle. */
void *or_csc_new(const or_lut *lut, size_t npix, const uint8_t *mask, const double *norm);
void or_csc_free(void *csc);
/* frames: n_frames x npix uint32 frames; out_i: n_frames x nbins. */
void or_csc_batch_u32(const void *csc, const uint32_t *frames, size_t n_frames, int threads,
                      double *out_i);

/* Bitshuffle + LZ4 chunks (bslz4.c; the codec of HDF5 filter 32008,
 * bitshuffle 0.5.2): LZ4 block decode (-1 when malformed), bit transposition
 * both ways, whole-chunk decode (0 ok, 1 header, 2 truncated, 3 bad LZ4
 * block), and the reference's two test encoders (mode 0 literal-only,
 * metal/tests.rs:297-357; mode 1 byte runs, metal/bench.rs:150-230). */
long or_lz4_decode_block(const uint8_t *in, size_t in_len, uint8_t *out, size_t out_cap);
void or_bitshuffle(const uint8_t *src, size_t n, size_t eb, uint8_t *dst);
void or_bitunshuffle(const uint8_t *src, size_t n, size_t eb, uint8_t *dst);
int or_bslz4_decode(const uint8_t *ch, size_t len, size_t pixels, size_t eb, uint8_t *out);
size_t or_bslz4_encode(const uint8_t *raw, size_t pixels, size_t eb, size_t block_bytes, int mode,
                       uint8_t *out, size_t out_cap);

#ifdef __cplusplus
}
#endif
#endif
