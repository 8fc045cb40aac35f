/* Multithreaded CPU baseline: a restatement of the reference's default CPU
 * batch1d path (TEST/BENCHMARK INFRASTRUCTURE ONLY; see oracle.h).
 *
 * Default Integrator settings are fast_math = true, SparseLayout::Csc and a
 * fixed mask (src/integration.rs:300-315). batch1d over arrays integrates
 * 4-frame tiles with Csc::fixed_results_batch4 when frames >= 4 * threads
 * (src/python/integrator.rs:715-760) and spreads tiles over a thread pool;
 * remaining frames go through the 1-frame path. Both accumulate pixel-major
 * into 4 lane histograms (lane = pixel & 3), in chunks of active pixels, and
 * merge lanes as (l0 + l1) + (l2 + l3) (src/lut/csc.rs:198-505, 631-690).
 * madd is a * b + c on x86-64 builds without the fma target feature.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
  size_t npix, nbins, nnz;
  uint32_t *ptr;    /* npix + 1 */
  uint32_t *bins;   /* nnz, pixel-major, bins ascending per pixel */
  double *weights;  /* nnz or NULL */
  uint32_t *active; /* pixels with entries that are not masked */
  size_t n_active;
  double *den;      /* nbins */
  size_t chunks;
} csc_t;

/* csc.rs:500-505. */
static size_t fixed_chunks(size_t nbins, size_t pixels) {
  size_t bytes_per_hist = nbins * (4 * 8 * 3);
  if (bytes_per_hist == 0) return 1;
  size_t limit = (32u * 1024u * 1024u) / bytes_per_hist;
  if (limit < 1) limit = 1;
  size_t c = (pixels + 262143) / 262144;
  if (c > limit) c = limit;
  return c < 1 ? 1 : c;
}

void *or_csc_new(const or_lut *lut, size_t npix, const uint8_t *mask, const double *norm) {
  csc_t *c = (csc_t *)calloc(1, sizeof(csc_t));
  c->npix = npix;
  c->nbins = lut->nbins;
  c->nnz = lut->nnz;
  c->ptr = (uint32_t *)calloc(npix + 1, sizeof(uint32_t));
  c->bins = (uint32_t *)malloc((lut->nnz ? lut->nnz : 1) * sizeof(uint32_t));
  c->weights = lut->w ? (double *)malloc((lut->nnz ? lut->nnz : 1) * sizeof(double)) : NULL;
  /* csc.rs:198-260: transpose in bin order. */
  for (size_t k = 0; k < lut->nnz; k++) c->ptr[lut->col[k] + 1]++;
  for (size_t p = 0; p < npix; p++) c->ptr[p + 1] += c->ptr[p];
  uint32_t *next = (uint32_t *)malloc((npix ? npix : 1) * sizeof(uint32_t));
  memcpy(next, c->ptr, npix * sizeof(uint32_t));
  for (size_t b = 0; b < lut->nbins; b++)
    for (uint64_t k = lut->row_ptr[b]; k < lut->row_ptr[b + 1]; k++) {
      uint64_t p = lut->col[k];
      uint32_t at = next[p]++;
      c->bins[at] = (uint32_t)b;
      if (c->weights) c->weights[at] = lut->w[k];
    }
  free(next);
  /* csc.rs:263-272. */
  c->active = (uint32_t *)malloc((npix ? npix : 1) * sizeof(uint32_t));
  for (size_t p = 0; p < npix; p++)
    if (c->ptr[p] != c->ptr[p + 1] && !(mask && mask[p])) c->active[c->n_active++] = (uint32_t)p;
  /* Denominators over the masked table, computed once per plan (not timed). */
  or_lut fixed;
  or_fixed_table(lut, mask, &fixed);
  c->den = (double *)malloc((lut->nbins ? lut->nbins : 1) * sizeof(double));
  or_denominators(&fixed, norm, c->den);
  or_lut_free(&fixed);
  c->chunks = fixed_chunks(c->nbins, c->n_active);
  return c;
}

void or_csc_free(void *h) {
  csc_t *c = (csc_t *)h;
  if (!c) return;
  free(c->ptr);
  free(c->bins);
  free(c->weights);
  free(c->active);
  free(c->den);
  free(c);
}

/* csc.rs:351-498 (fixed_results_batch4) for F frames at a time (F = 4 or 1,
 * the per-frame path being the same loop with one frame). */
static void run_tile(const csc_t *c, const uint32_t *const *sig, int F, double *const *out) {
  size_t n = c->nbins;
  double *acc = (double *)calloc(n * 4, sizeof(double));       /* [bin][frame] */
  double *chunk = (double *)malloc(4 * n * 4 * sizeof(double)); /* [lane*n+bin][frame] */
  for (size_t i = 0; i < c->chunks; i++) {
    memset(chunk, 0, 4 * n * 4 * sizeof(double));
    size_t lo = i * c->n_active / c->chunks, hi = (i + 1) * c->n_active / c->chunks;
    for (size_t q = lo; q < hi; q++) {
      uint32_t p = c->active[q];
      double x[4];
      for (int f = 0; f < F; f++) x[f] = (double)sig[f][p];
      size_t lane = (size_t)(p & 3) * n;
      for (uint32_t k = c->ptr[p]; k < c->ptr[p + 1]; k++) {
        double *dst = chunk + (c->bins[k] + lane) * 4;
        if (c->weights) {
          double w = c->weights[k];
          for (int f = 0; f < F; f++) dst[f] = w * x[f] + dst[f];
        } else {
          for (int f = 0; f < F; f++) dst[f] += x[f];
        }
      }
    }
    for (size_t b = 0; b < n; b++)
      for (int f = 0; f < F; f++)
        acc[b * 4 + f] += (chunk[b * 4 + f] + chunk[(n + b) * 4 + f]) +
                          (chunk[(2 * n + b) * 4 + f] + chunk[(3 * n + b) * 4 + f]);
  }
  for (int f = 0; f < F; f++)
    for (size_t b = 0; b < n; b++) {
      double d = c->den[b];
      out[f][b] = d == 0.0 ? 0.0 : acc[b * 4 + f] / d;
    }
  free(acc);
  free(chunk);
}

void or_csc_batch_u32(const void *h, const uint32_t *frames, size_t n_frames, int threads,
                      double *out_i) {
  const csc_t *c = (const csc_t *)h;
  if (threads < 1) threads = 1;
  size_t tiles = n_frames >= (size_t)(4 * threads) ? n_frames / 4 : 0;
  size_t rest_start = tiles * 4;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
  for (long long t = 0; t < (long long)(tiles + (n_frames - rest_start)); t++) {
    if ((size_t)t < tiles) {
      const uint32_t *sig[4];
      double *out[4];
      for (int f = 0; f < 4; f++) {
        sig[f] = frames + (size_t)(4 * t + f) * c->npix;
        out[f] = out_i + (size_t)(4 * t + f) * c->nbins;
      }
      run_tile(c, sig, 4, out);
    } else {
      size_t fr = rest_start + ((size_t)t - tiles);
      const uint32_t *sig[1] = {frames + fr * c->npix};
      double *out[1] = {out_i + fr * c->nbins};
      run_tile(c, sig, 1, out);
    }
  }
}
