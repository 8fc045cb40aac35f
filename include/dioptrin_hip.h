/* dioptrin_hip: C ABI of the MI355X-native Dioptrin integration hot path.
 *
 * Plain C types only: handles, pointers and sizes. Every entry returns a
 * status (0 = ok) and leaves a thread-local message for dio_last_error().
 * Status codes map one to one onto the reference's CudaError categories
 * (src/integration/cuda.rs:218-237).
 *
 * Reference interfaces each entry replaces (paths relative to the reference
 * repository root) are cited per declaration.
 */
#ifndef DIOPTRIN_HIP_H
#define DIOPTRIN_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (CudaError: NotCompiled/Unavailable, Unsupported,
 * InvalidInput, Device). */
enum {
  DIO_OK = 0,
  DIO_ERR_UNAVAILABLE = 1,
  DIO_ERR_UNSUPPORTED = 2,
  DIO_ERR_INVALID = 3,
  DIO_ERR_DEVICE = 4
};

/* IntegrationMethod (src/integration.rs:227-237). */
enum { DIO_COUNTING = 0, DIO_PIXEL_SPLIT = 1, DIO_PYFAI_FULL = 2, DIO_SUPERSAMPLED = 3 };
/* RadialUnit (src/integration.rs:115-126). */
enum { DIO_UNIT_TTH_RAD = 0, DIO_UNIT_TTH_DEG = 1, DIO_UNIT_Q_A = 2, DIO_UNIT_Q_NM = 3 };
/* Accumulation precision (CudaPrecision, src/integration/cuda.rs:64-73), plus
 * DIO_F32_WIDE: float32 storage with float64 register accumulation, which on
 * CDNA costs one full-rate DFMA per term. */
enum { DIO_F32 = 0, DIO_F32_COMPENSATED = 1, DIO_F64 = 2, DIO_F32_WIDE = 3 };
/* Frame element types (CudaPixel, src/integration/cuda.rs:354-367). */
enum { DIO_U16 = 0, DIO_U32 = 1, DIO_I32 = 2, DIO_F32T = 3, DIO_F64T = 4 };
/* Row order of incoming frames (InputOrder, src/integration/fixed.rs:20-24;
 * Cuda1dOptions::with_input_order, src/integration/cuda.rs:174-181):
 * DIO_ORDER_FLIP_VERTICAL frames hold the detector's rows bottom-up. The plan
 * reads them in that order; results are the bits of the upright frames with
 * a DIO_ORDER_DETECTOR plan. Exported tables stay in detector order. */
enum { DIO_ORDER_DETECTOR = 0, DIO_ORDER_FLIP_VERTICAL = 1 };
/* Output selection flags; DIO_DEVICE_FRAMES (dio_server_submit only): the
 * frames pointer is device memory of the server's context. DIO_RAW_SUMS:
 * the sums before division (IntermediateResult1d/2d, lut/integrate.rs:372-500):
 * intensity receives sum_signal = sum w*x per bin and sigma (with
 * DIO_WITH_ERRORS) sum_variance = sum w^2*max(x, 0) (Poisson variance);
 * sum_norm is dio_plan_denominators and count dio_plan_counts. */
enum { DIO_WITH_ERRORS = 1, DIO_WITH_AZIMUTHAL_STD = 2, DIO_DEVICE_FRAMES = 4, DIO_RAW_SUMS = 8 };

typedef struct dio_ctx dio_ctx;
typedef struct dio_plan dio_plan;
typedef struct dio_graph dio_graph;
typedef struct dio_server dio_server;

/* Poni (src/poni.rs:13-24): metres and radians. */
typedef struct {
  double pixel1, pixel2, distance, poni1, poni2, rot1, rot2, rot3, wavelength;
} dio_poni;

/* What a plan is built from: the Integrator settings that key the
 * reference's table caches (OneDKey/TwoDKey, src/integration/cache.rs:23-36)
 * plus the image shape, mask and corrections of its snapshot. */
typedef struct {
  uint32_t dim1, dim2;          /* image rows, columns */
  int32_t method;               /* DIO_COUNTING .. DIO_SUPERSAMPLED */
  uint32_t n_ss;                /* subpixels per side for DIO_SUPERSAMPLED */
  int32_t unit;                 /* DIO_UNIT_* */
  uint32_t ndim;                /* 1 or 2 */
  uint32_t n_radial;            /* bins (1D) or radial bins (2D) */
  uint32_t n_azimuthal;         /* 2D only */
  int32_t has_radial_range;
  double radial_min, radial_max;
  int32_t has_azimuthal_range;  /* 2D only */
  double azimuthal_min, azimuthal_max;
  int32_t has_polarization;
  double polarization_factor;
  const uint8_t *mask;          /* host, dim1*dim2, nonzero = masked; may be NULL */
  int32_t precision;            /* DIO_F32 .. DIO_F32_WIDE */
  int32_t input_order;          /* DIO_ORDER_DETECTOR (0) or DIO_ORDER_FLIP_VERTICAL */
} dio_plan_desc;

typedef struct {
  uint32_t nbins;               /* rows of the table (n_radial * n_azimuthal in 2D) */
  uint32_t n_radial, n_azimuthal;
  uint64_t nnz;                 /* entries of the unmasked table */
  uint64_t nnz_masked;          /* entries of the execution table */
  uint32_t weighted;            /* 0 for counting tables (unit weights) */
  uint32_t n_segments;          /* gather work units */
  uint64_t uncertain_angles;    /* angles recomputed with host libm */
  uint32_t singular_pixels;     /* beam pixels integrated by sampling */
  double build_ms;              /* plan build wall time */
  double hist_min, hist_step;   /* 1D radial axis */
  uint64_t device_bytes;        /* device memory held by the plan */
  uint32_t max_tile_groups;     /* largest table of one detector tile, in 4-entry groups */
} dio_plan_info;

typedef struct {
  char name[256];
  char arch[64];
  int32_t compute_units;
  int32_t xcds;
  uint64_t total_memory;
  int32_t runtime_version;
} dio_device_info;

/* Thread-local text of the last failure on this thread. */
const char *dio_last_error(void);
const char *dio_version(void);

/* Device context, cached per ordinal (device.rs:142-176 Cuda::open/device). */
int dio_open(int device, dio_ctx **out);
int dio_close(dio_ctx *ctx);
/* cuda_info (src/integration/cuda.rs:309-326). */
int dio_device_info_get(dio_ctx *ctx, dio_device_info *out);

/* Integrator::prepare_cuda1d / prepare_cuda2d (src/integration/cuda.rs:1089,
 * 1111): builds the lookup table ON THE DEVICE (corner geometry, splitting,
 * count-then-fill CSR, mask compaction, FP64 denominators) and keeps it
 * resident. The plan is a snapshot like Cuda1dPlan (cuda.rs:378-386). */
int dio_plan_create(dio_ctx *ctx, const dio_poni *poni, const dio_plan_desc *desc,
                    dio_plan **out);
int dio_plan_destroy(dio_plan *plan);
/* Cuda1dPlan::{n_points, table_entries, ...} (cuda.rs:411-470). */
int dio_plan_info_get(dio_plan *plan, dio_plan_info *out);
/* Radial bin centres (Cuda1dPlan::radial, cuda.rs:440; histogram::bins). */
int dio_plan_radial(dio_plan *plan, double *out);
/* Azimuthal bin centres of a 2D plan (Cuda2dPlan::azimuthal). */
int dio_plan_azimuthal(dio_plan *plan, double *out);
/* Copies the table for bit-exact checks: masked = 0 gives the builder's
 * Lut1d/Lut2d (counting_csr.rs:13-40), masked = 1 the execution table
 * (prepared.rs:163-201). Any output pointer may be NULL. Capacities are in
 * elements; row_ptr needs nbins + 1. */
int dio_plan_export_csr(dio_plan *plan, int masked, uint32_t *row_ptr, uint32_t *col_idx,
                        double *w64, float *w32, uint64_t capacity);
/* Strict FP64 denominators of the execution table (prepared.rs:229-262). */
int dio_plan_denominators(dio_plan *plan, double *out);
/* Per-pixel normalization (polarization * solid angle, corrections.rs:57-93). */
int dio_plan_normalization(dio_plan *plan, double *out);
/* Per bin: the sum of the execution table's weights in row order (entries
 * for unweighted tables), the count of IntermediateResult1d/2d
 * (lut/integrate.rs:52-110 with COUNT). nbins doubles. */
int dio_plan_counts(dio_plan *plan, double *out);

/* Cuda1dPlan::integrate_frames / integrate_stack and Cuda2dPlan
 * equivalents (cuda.rs:480-548, 770-863): host frames in, host results out.
 * frames: n_frames contiguous frames of dtype. intensity/sigma/std:
 * n_frames * nbins float64 each (sigma/std only with their flag).
 * Synchronous; frames are borrowed for the call. */
int dio_integrate(dio_plan *plan, const void *frames, int dtype, uint64_t n_frames,
                  uint32_t flags, double *intensity, double *sigma, double *std_out);

/* Device-resident batch: frames and outputs already in device memory
 * (the batch1d/batch2d hot path). Enqueued on `stream` (a hipStream_t,
 * NULL = the plan's stream); returns without synchronizing. */
int dio_integrate_device(dio_plan *plan, const void *d_frames, int dtype, uint64_t n_frames,
                         uint32_t flags, double *d_intensity, double *d_sigma, double *d_std,
                         void *stream);

/* dio_integrate_device that allocates nothing. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, flags and results (bit for bit) as
 * dio_integrate_device; the tile kernel's segment partials go to a fixed
 * 64 MiB device array that each device holds from dio_open until the
 * process ends. Passes are shortened to fit it; a plan whose partials of a
 * single frame exceed it is refused with DIO_ERR_INVALID. Every call on one
 * device shares that array: calls that overlap on one device (any plan,
 * any stream) race, so the caller serialises them. */
int dio_integrate_device_noalloc(dio_plan *plan, const void *d_frames, int dtype,
                                 uint64_t n_frames, uint32_t flags, double *d_intensity,
                                 double *d_sigma, double *d_std, void *stream);

/* Dynamic (per-frame) masks: Dynamic1dPlan::integrate(data, Some(mask),
 * errors, false) (src/integration/dynamic.rs:143-186; row sums
 * src/lut/prepared.rs:1011-1121, normalize_bin 985-996). Each frame comes
 * with its own mask (masks: n_frames * dim1 * dim2 bytes, nonzero = masked,
 * in the plan's input order) and is integrated on the plan's unmasked
 * table; the plan's own mask is not applied. Per bin: intensity =
 * sum w*x / sum w*norm, sigma = sqrt(max(sum w^2*max(x,0) / (sum w*norm)^2,
 * 0)), both 0 where the denominator is 0. flags: 0 or DIO_WITH_ERRORS;
 * 1D plans only. A plan without a mask computes at the plan's precision
 * (tile kernel); a plan with a mask in FP64 (CSR kernel). The device
 * variant uses plan-owned staging: calls on one plan that overlap on
 * different streams race. The host variant is synchronous; the device variant is
 * enqueued on `stream` (NULL = the plan's stream). */
int dio_integrate_masked(dio_plan *plan, const void *frames, int dtype, uint64_t n_frames,
                         const uint8_t *masks, uint32_t flags, double *intensity,
                         double *sigma);
int dio_integrate_masked_device(dio_plan *plan, const void *d_frames, int dtype,
                                uint64_t n_frames, const uint8_t *d_masks, uint32_t flags,
                                double *d_intensity, double *d_sigma, void *stream);

/* Supplied pixel variances (Prepared1d::integrate_with_variances,
 * src/lut/prepared.rs:737-766; Dynamic1dPlan::integrate_with_variances,
 * src/integration/dynamic.rs:128-139; on an explicit-grid 2D plan
 * Integrator::integrate2d_raw_on_grid_with_variances_from_slice,
 * src/integration/grid.rs:347-392). variances: n_frames planes of the image's
 * shape, of dtype vdtype, finite and nonnegative (grid.rs:1102-1107; checked,
 * DIO_ERR_INVALID otherwise). Per bin: sum_signal = sum w*x and
 * sum_variance = sum w^2*v over the execution table. flags DIO_RAW_SUMS:
 * intensity and sigma receive those sums; otherwise intensity =
 * sum_signal / den and sigma = sqrt(max(sum_variance / den^2, 0)), both 0
 * where den == 0 (normalize_bin, prepared.rs:985-996). Host buffers,
 * synchronous. */
int dio_integrate_variances(dio_plan *plan, const void *frames, int dtype,
                            const void *variances, int vdtype, uint64_t n_frames,
                            uint32_t flags, double *intensity, double *sigma);

/* Merged sums before division of dio_multi_integrate (the merge totals,
 * Multi2dResult::raw, src/multi.rs:265-297); any field may be NULL. */
typedef struct {
  double *sum_signal;   /* n_frames x nbins */
  double *sum_norm;     /* nbins: sum over detectors of den x factor */
  double *sum_variance; /* n_frames x nbins (with errors or variances) */
  double *count;        /* nbins */
} dio_raw_sums;

/* MultiIntegrator (src/multi.rs:198-316 for 2D, 346-510 for 1D): one plan per
 * detector, all on one device and one bin grid (same nbins and ndim; the
 * caller builds them with the shared radial range and, in 2D, the same
 * explicit azimuthal range). frames[d]: n_frames frames of detector d (host,
 * dtypes[d]); variances (may be NULL): per detector n_frames supplied
 * variance planes (vdtypes[d]), which imply errors; factors (may be NULL =
 * all 1): finite positive multipliers of each detector's denominators
 * (monitor x solid-angle scale, multi.rs:244-264). Per frame and bin the raw
 * sums of every detector are added on the device in detector order from 0:
 * S = sum sum_signal, N = sum den x factor, V = sum sum_variance; intensity =
 * S / N, and with DIO_WITH_ERRORS (or variances) sigma = sqrt(max(V / N^2,
 * 0)) for 1D plans and sqrt(max(V, 0)) / N for 2D plans, all 0 where
 * N == 0. raw (may be NULL) receives the totals. Synchronous. */
int dio_multi_integrate(uint32_t n_det, dio_plan *const *plans, const void *const *frames,
                        const int *dtypes, const void *const *variances, const int *vdtypes,
                        const double *factors, uint64_t n_frames, uint32_t flags,
                        double *intensity, double *sigma, dio_raw_sums *raw);

/* Bitshuffle + LZ4 chunks (raw HDF5 filter 32008 chunks, one per frame, as
 * h5py's read_direct_chunk returns them; bitshuffle 0.5.2 / LZ4 block
 * format). Cuda1dPlan::integrate_bslz4 / Cuda2dPlan::integrate_bslz4(_into)
 * (src/integration/cuda.rs:550-564, 758-767, 843-862; decode kernel
 * src/integration/cuda/kernels.cu:302-493; descriptor pre-scan
 * src/integration/cuda/device.rs:2029-2180). */
typedef struct {
  uint32_t src_off;      /* byte offset of the block in the compressed buffer */
  uint32_t comp_len;     /* compressed bytes; bit 31 set: verbatim tail bytes */
  uint32_t dst_off;      /* first element of the block in the decoded frames */
  uint32_t uncomp_bytes; /* decoded bytes of the block */
} dio_bslz4_block;

/* Decodes n_frames chunks (chunks[f], sizes[f] bytes, host memory) of dtype
 * DIO_U16, DIO_U32 or DIO_I32 on the device and integrates them with the
 * plan; flags DIO_WITH_ERRORS, DIO_RAW_SUMS. Results are the bits of
 * dio_integrate on the decoded frames. A malformed chunk (header, sizes,
 * block size, truncation: checked on the host; LZ4 stream: checked on the
 * device) fails with DIO_ERR_INVALID. Synchronous. */
int dio_integrate_bslz4(dio_plan *plan, const uint8_t *const *chunks, const uint64_t *sizes,
                        uint64_t n_frames, int dtype, uint32_t flags, double *intensity,
                        double *sigma);
/* Block descriptors of n_frames chunks of `pixels` elements each, their
 * offsets relative to the chunks concatenated in order (the reference's
 * pre-scan). blocks == NULL: count only. max_comp / max_block (may be NULL)
 * receive the largest compressed and decoded block sizes. */
int dio_bslz4_scan(const uint8_t *const *chunks, const uint64_t *sizes, uint64_t n_frames,
                   uint64_t pixels, int dtype, dio_bslz4_block *blocks, uint64_t cap,
                   uint64_t *n_blocks, uint32_t *max_comp, uint32_t *max_block);
/* Device-resident decode (no reference counterpart as an entry point: the
 * reference decodes inside integrate_bslz4): d_comp holds comp_bytes bytes
 * of concatenated chunks, d_blocks their descriptors; decoded elements go to
 * d_frames (n_elems elements of dtype). *d_err becomes nonzero when a block
 * is corrupt or inconsistent with the buffers (the caller zeroes it first);
 * nothing is written outside d_frames. Enqueued on `stream`. */
int dio_bslz4_decode_device(dio_ctx *ctx, const uint8_t *d_comp, uint64_t comp_bytes,
                            const dio_bslz4_block *d_blocks, uint64_t n_blocks, uint32_t max_comp,
                            uint32_t max_block, int dtype, void *d_frames, uint64_t n_elems,
                            uint32_t *d_err, void *stream);
/* Writes one frame as a Bitshuffle + LZ4 chunk (greedy LZ4 block
 * compressor; block_bytes 0 = 8192). dio_bslz4_bound gives a sufficient
 * capacity. No reference counterpart (the reference reads chunks only). */
int dio_bslz4_compress(const void *frame, uint64_t n_elems, int dtype, uint32_t block_bytes,
                       uint8_t *out, uint64_t cap, uint64_t *out_len);
uint64_t dio_bslz4_bound(uint64_t n_elems, int dtype, uint32_t block_bytes);

/* Launch shape of the fused batch kernel (no reference counterpart; the
 * reference launches one frame per kernel, src/integration/cuda/device.rs:
 * 1460-1584): frames per pass and frames per lane, one of the pairs (1, 1),
 * (2, 2), (4, 4), (8, 2), (8, 4), (16, 1), (16, 2), (16, 4) (at most 8 frames
 * per pass for DIO_F64); the lanes sharing one table segment are
 * frames_per_tile / frames_per_lane. dio_plan_autotune picks both from a
 * short device sweep and stores them in the plan; frames_per_lane may be
 * null. Results do not depend on either choice. */
int dio_plan_autotune(dio_plan *plan, int dtype, uint64_t n_frames_hint, uint32_t flags,
                      uint32_t *frames_per_tile, uint32_t *frames_per_lane);
/* frames_per_lane = 0 picks min(4, frames_per_tile). */
int dio_plan_set_tile(dio_plan *plan, uint32_t frames_per_tile, uint32_t frames_per_lane);

/* Steady-state replay: captures one dio_integrate_device call into a
 * hipGraph and replays it. */
int dio_graph_create(dio_plan *plan, const void *d_frames, int dtype, uint64_t n_frames,
                     uint32_t flags, double *d_intensity, double *d_sigma, double *d_std,
                     dio_graph **out);
int dio_graph_launch(dio_graph *graph, void *stream);
int dio_graph_destroy(dio_graph *graph);

/* Kernel-duration probe for roofline measurement: times `reps` launches of
 * the dominant gather kernel with HIP events on the plan's stream. */
int dio_plan_time_gather(dio_plan *plan, const void *d_frames, int dtype, uint64_t n_frames,
                         uint32_t flags, double *d_intensity, double *d_sigma, int reps,
                         double *avg_ms);
/* Runs one dio_integrate_device of the whole batch with HIP events recorded
 * on the plan's stream around each launch of the gather kernel (one pass, or
 * two passes sharing the table for the double-buffered kernel), waits, and
 * returns the kernel durations (ms) of the launches in order and, if frames
 * is not NULL, the frames each launch covered: min(cap, launches) values,
 * the launch count in *n_launches. Measurement only (no reference
 * counterpart). */
int dio_plan_time_passes(dio_plan *plan, const void *d_frames, int dtype, uint64_t n_frames,
                         uint32_t flags, double *d_intensity, double *d_sigma, float *ms,
                         uint32_t *frames, uint32_t cap, uint32_t *n_launches);

/* Device memory and synchronisation for callers without their own device
 * allocator (the reference's plans own their device buffers the same way,
 * src/integration/cuda/device.rs:1354-1460). */
int dio_device_alloc(dio_ctx *ctx, uint64_t bytes, void **d_ptr);
int dio_device_free(dio_ctx *ctx, void *d_ptr);
/* kind: 0 host to device, 1 device to host, 2 device to device. */
int dio_memcpy(dio_ctx *ctx, void *dst, const void *src, uint64_t bytes, int kind);
/* Page-locked host memory (the pinned staging of the reference's GPU path,
 * src/integration/cuda/device.rs:1063-1079): frames placed here reach the
 * device by DMA in dio_integrate, without a pageable staging copy. */
int dio_host_alloc(dio_ctx *ctx, uint64_t bytes, void **h_ptr);
int dio_host_free(dio_ctx *ctx, void *h_ptr);
int dio_synchronize(dio_ctx *ctx);

/* Seeded synthetic detector frames for benchmarks and tests (no reference
 * counterpart): d_out[f][p] = Poisson(rate[p] + background[f]) as uint32,
 * rate given per pixel on the device (float32), background per frame on the
 * host; stream (f, p) draws from a counter-based generator keyed by seed. */
int dio_synthetic_poisson_u32(dio_ctx *ctx, const float *d_rate, uint64_t npix,
                              const double *background, uint64_t n_frames, uint64_t seed,
                              uint32_t *d_out);

/* Continuous-batching server (no single reference counterpart: it extends the
 * reference's one-entry table caches, src/integration/cache.rs:42-96, to an
 * LRU of device-resident plans, and its one-stream-per-plan execution,
 * src/integration/cuda/device.rs:1354-1584, to a pool of stream slots).
 * Plans missing from the cache are built on a builder thread; a request
 * waits only for its own plan. Requests whose plan is ready are admitted in
 * arrival order when a slot is free and the server's device bytes (resident
 * plans + slot scratch + frames and outputs of requests in flight) fit
 * hbm_budget_bytes, evicting idle plans for room; a request that does not
 * fit with nothing in flight runs alone. Zeroed fields take the defaults
 * shown. */
typedef struct {
  uint32_t n_slots;            /* concurrent requests (streams), default 4, max 64 */
  uint32_t plan_cache_entries; /* device-resident plans kept, default 32 */
  uint64_t hbm_budget_bytes;   /* plans + scratch + requests in flight, default 32 GiB */
} dio_server_config;

typedef struct {
  uint64_t completed, failed, frames, queued;
  uint32_t inflight, peak_inflight;
  uint64_t bytes_in_use, peak_bytes; /* plan + scratch + request bytes, now and at peak */
  uint64_t plan_builds, cache_hits;
  double p50_ms, p99_ms, mean_ms, max_ms; /* submit-to-result latency */
  double span_s, frames_per_s;             /* first submit to last completion */
  uint64_t plan_bytes;    /* resident plans */
  uint64_t scratch_bytes; /* gather partials of the slots */
  uint64_t request_bytes; /* frames + outputs of requests in flight */
  uint32_t plans_resident, builds_pending;
} dio_server_stats;

int dio_server_create(dio_ctx *ctx, const dio_server_config *cfg, dio_server **out);
/* Queues one request; frames (n_frames x dim1 x dim2 of dtype, in host
 * memory, or device memory with DIO_DEVICE_FRAMES) and the host outputs
 * (n_frames x nbins doubles each) are borrowed until poll reports it done.
 * The plan is identified by every field of poni and desc that shapes it and
 * the full mask contents, like dio_plan_create's arguments (Integrator::integrate1d / integrate2d,
 * src/integration/methods.rs:94-100, 461). */
int dio_server_submit(dio_server *srv, const dio_poni *poni, const dio_plan_desc *desc,
                      const void *frames, int dtype, uint64_t n_frames, uint32_t flags,
                      double *intensity, double *sigma, double *std_out, uint64_t *ticket);
/* Waits up to timeout_s (0: no wait) for a ticket; *done = 1 once its outputs
 * are written, and the request's own status is returned then (the ticket is
 * released). */
int dio_server_poll(dio_server *srv, uint64_t ticket, double timeout_s, int *done,
                    double *latency_ms);
int dio_server_stats_get(dio_server *srv, dio_server_stats *out);
/* Drains queued and running requests, then frees every plan and stream. */
int dio_server_destroy(dio_server *srv);

#ifdef __cplusplus
}
#endif
#endif
