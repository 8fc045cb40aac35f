// C ABI of the MI355X integration path: contexts, plans (device-built
// tables), integration entry points, autotuner, graph replay.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "plan.hpp"

namespace {

thread_local std::string g_last_error;
std::mutex g_ctx_mu;
std::map<int, std::unique_ptr<dio_ctx>> g_ctx;

using dio::dtype_bytes;
using dio::guarded;
using dio::require;

// Configuration checks before any device work (cuda.rs:1001-1057).
void check_desc(const dio_plan_desc& d) {
  require(d.dim1 > 0 && d.dim2 > 0, DIO_ERR_INVALID, "image shape has a zero side");
  // The tiled gather addresses a frame with 32-bit byte offsets (up to 8
  // bytes per pixel), so an image holds fewer than 2^29 pixels.
  require((uint64_t)d.dim1 * d.dim2 < (1ull << 29), DIO_ERR_INVALID,
          "image must hold fewer than 2^29 pixels");
  require(d.ndim == 1 || d.ndim == 2, DIO_ERR_INVALID, "ndim must be 1 or 2");
  require(d.n_radial > 0 && d.n_radial <= 0x7fffffffu, DIO_ERR_INVALID,
          "n_points must be between 1 and 2^31-1");
  if (d.ndim == 2) {
    require(d.n_azimuthal > 0 && (uint64_t)d.n_radial * d.n_azimuthal <= 0x7fffffffull,
            DIO_ERR_INVALID, "2D bin count must be in 1..2^31-1");
  }
  require(d.method >= DIO_COUNTING && d.method <= DIO_SUPERSAMPLED, DIO_ERR_INVALID,
          "unknown method");
  require(d.method != DIO_SUPERSAMPLED || d.n_ss > 0, DIO_ERR_INVALID, "n_ss must be positive");
  require(d.unit >= DIO_UNIT_TTH_RAD && d.unit <= DIO_UNIT_Q_NM, DIO_ERR_INVALID, "unknown unit");
  require(d.precision >= DIO_F32 && d.precision <= DIO_F32_WIDE, DIO_ERR_INVALID,
          "unknown precision");
  require(d.input_order == DIO_ORDER_DETECTOR || d.input_order == DIO_ORDER_FLIP_VERTICAL,
          DIO_ERR_INVALID, "unknown input order");
  if (d.has_radial_range)
    require(std::isfinite(d.radial_min) && std::isfinite(d.radial_max) &&
                d.radial_min < d.radial_max,
            DIO_ERR_INVALID, "radial_range must be finite with min < max");
  if (d.has_polarization)
    require(d.polarization_factor >= -1.0 && d.polarization_factor <= 1.0, DIO_ERR_INVALID,
            "polarization factor must lie in [-1, 1]");
}

void ensure_out(dio::DevBuf<double>& b, size_t n) {
  if (b.n < n) b.alloc(n);
}

}  // namespace

namespace dio {
// Enqueue the whole integration of device frames on `s`.
void run_device(dio_plan* p, TileScratch& sc, const void* frames, int dtype,
                uint64_t n_frames, uint32_t flags, double* I, double* S, double* SD,
                hipStream_t s, LaunchEvents* launches) {
  bool var = flags & DIO_WITH_ERRORS;
  if (flags & DIO_WITH_AZIMUTHAL_STD) {
    require(!(flags & DIO_RAW_SUMS), DIO_ERR_UNSUPPORTED,
            "raw sums are not defined with the azimuthal std");
    require(p->desc.ndim == 1, DIO_ERR_UNSUPPORTED, "azimuthal std is defined for 1D plans");
    require(SD != nullptr, DIO_ERR_INVALID, "azimuthal std output missing");
    require(!var || S != nullptr, DIO_ERR_INVALID, "sigma output missing");
    sc.ensure_bytes(std_pass_bytes(p->tt, std_frames_per_pass(), var));
    azimuthal_std_tiled(p->tt, p->st.norm, sc.ptr(), sc.bytes(), frames, dtype, n_frames, var, I,
                        S, SD, s);
    return;
  }
  require(!var || S != nullptr, DIO_ERR_INVALID, "sigma output missing");
  tiled_gather(p->tt, sc, frames, dtype, n_frames, var, p->desc.precision, p->tile, p->lanes, I,
               S, s, launches, (flags & DIO_RAW_SUMS) != 0);
}

}  // namespace dio

namespace dio {
size_t plan_device_bytes(const dio_plan* p) {
  auto tb = [](const dio::Table& t) {
    return t.row_ptr.bytes() + t.col.bytes() + t.w64.bytes() + t.w32.bytes();
  };
  return tb(p->b.full) + (p->b.has_exec ? tb(p->b.exec) : 0) + p->b.norm.bytes() +
         p->b.den.bytes() + p->b.mask.bytes() + p->tt.bytes() + p->sc.part.bytes() +
         p->stage.bytes() + p->mstage.bytes() + p->dyn_x.bytes() + p->dyn_n.bytes() +
         p->dyn_d.bytes() + p->norm32.bytes() + p->dyn_m.bytes() + p->dyn_r.bytes() +
         p->out_i.bytes() + p->out_s.bytes() + p->out_d.bytes();
}

dio_plan* make_plan(dio_ctx* ctx, const dio_poni& poni, const dio_plan_desc& desc) {
  check_desc(desc);
  // RadialConverter::new (integration.rs:154-169) asserts a positive
  // wavelength for q units.
  if (desc.unit == DIO_UNIT_Q_A || desc.unit == DIO_UNIT_Q_NM)
    require(poni.wavelength > 0.0, DIO_ERR_INVALID,
            "wavelength must be positive for q units");
  auto t0 = std::chrono::steady_clock::now();
  DIO_HIP(hipSetDevice(ctx->device));
  auto p = std::make_unique<dio_plan>();
  p->ctx = ctx;
  p->device = ctx->device;
  p->desc = desc;
  p->desc.mask = nullptr;
  DIO_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
  p->geo = dio::make_geo(poni, desc);
  dio::build_tables(p->geo, desc, p->stream, p->b);
  const dio::Table& t = p->b.exec_table();
  p->gt.col = t.col.p;
  p->gt.w32 = t.w32.p;
  p->gt.w64 = t.w64.p;
  p->gt.den = p->b.den.p;
  p->gt.nbins = t.nbins;
  p->gt.npix = (uint64_t)desc.dim1 * desc.dim2;
  p->gt.nnz = t.nnz;
  dio::build_tiled(p->tt, desc.dim1, desc.dim2, t.nbins, t.nnz, t.row_ptr.p, t.col.p,
                   t.w32.p, t.w64.p, desc.precision == DIO_F64, p->stream);
  p->tt.den = p->b.den.p;
  p->tt.flip = desc.input_order == DIO_ORDER_FLIP_VERTICAL;
  p->tt.cus = (uint32_t)std::max(1, ctx->prop.multiProcessorCount);
  // DIOPTRIN_TILE_WORKGROUPS caps the gather's persistent grid (tuning, and
  // tests of the many-tiles-per-workgroup path on small detectors).
  if (const char* e = std::getenv("DIOPTRIN_TILE_WORKGROUPS")) {
    const long v = std::atol(e);
    if (v > 0) p->tt.max_wgs = (uint32_t)std::min<long>(v, 1L << 20);
  }
  // DIOPTRIN_TILE_DB=0 selects the single-buffered kernel for 16-frame passes
  // (comparison runs; results are the same bits either way).
  if (const char* e = std::getenv("DIOPTRIN_TILE_DB")) p->tt.db = std::atol(e) != 0;
  if (desc.precision == DIO_F64) p->tile = std::min(p->tile, dio::max_frames_for(DIO_F64));
  p->st.row_ptr = p->b.full.row_ptr.p;
  p->st.col = p->b.full.col.p;
  p->st.w32 = p->b.full.w32.p;
  p->st.w64 = p->b.full.w64.p;
  p->st.weighted = p->b.weighted ? 1 : 0;
  p->st.mask = p->b.mask.p;
  p->st.norm = p->b.norm.p;
  p->st.nbins = p->b.full.nbins;
  p->st.npix = p->gt.npix;
  p->st.dim1 = desc.dim1;
  p->st.dim2 = desc.dim2;
  p->st.flip = desc.input_order == DIO_ORDER_FLIP_VERTICAL;
  p->build_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return p.release();
}
}  // namespace dio

void dio_set_last_error(const char* msg) { g_last_error = msg; }

extern "C" {

int dio_ctx_device(dio_ctx* ctx) { return ctx->device; }

const char* dio_last_error(void) { return g_last_error.c_str(); }
const char* dio_version(void) { return "dioptrin_hip 0.1.0 (gfx950)"; }

int dio_open(int device, dio_ctx** out) {
  return guarded([&] {
    require(out != nullptr, DIO_ERR_INVALID, "null output handle");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
      dio::fail(DIO_ERR_UNAVAILABLE, "no HIP device available");
    require(device >= 0 && device < count, DIO_ERR_UNAVAILABLE, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    auto& slot = g_ctx[device];
    if (!slot) {
      auto c = std::make_unique<dio_ctx>();
      c->device = device;
      DIO_HIP(hipSetDevice(device));
      DIO_HIP(hipGetDeviceProperties(&c->prop, device));
      dio::noalloc_init(device);  // device address of the fixed scratch, this device current
      slot = std::move(c);
    }
    slot->refs++;
    *out = slot.get();
  });
}

int dio_close(dio_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    ctx->refs = std::max(0, ctx->refs - 1);
  });
}

int dio_device_info_get(dio_ctx* ctx, dio_device_info* out) {
  return guarded([&] {
    require(ctx && out, DIO_ERR_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    std::strncpy(out->name, ctx->prop.name, sizeof(out->name) - 1);
    std::strncpy(out->arch, ctx->prop.gcnArchName, sizeof(out->arch) - 1);
    out->compute_units = ctx->prop.multiProcessorCount;
    out->xcds = 8;
    out->total_memory = ctx->prop.totalGlobalMem;
    int v = 0;
    (void)hipRuntimeGetVersion(&v);
    out->runtime_version = v;
  });
}

int dio_plan_create(dio_ctx* ctx, const dio_poni* poni, const dio_plan_desc* desc,
                    dio_plan** out) {
  return guarded([&] {
    require(poni && desc && out, DIO_ERR_INVALID, "null argument");
    *out = nullptr;
    check_desc(*desc);
    require(ctx != nullptr, DIO_ERR_INVALID, "null context");
    *out = dio::make_plan(ctx, *poni, *desc);
  });
}

int dio_plan_destroy(dio_plan* plan) {
  return guarded([&] {
    if (!plan) return;
    (void)hipSetDevice(plan->device);
    (void)hipStreamSynchronize(plan->stream);
    delete plan;
  });
}

int dio_plan_info_get(dio_plan* p, dio_plan_info* out) {
  return guarded([&] {
    require(p && out, DIO_ERR_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->nbins = p->b.full.nbins;
    out->n_radial = p->b.n_radial;
    out->n_azimuthal = p->b.n_azimuthal;
    out->nnz = p->b.full.nnz;
    out->nnz_masked = p->b.exec_table().nnz;
    out->weighted = p->b.weighted ? 1 : 0;
    out->n_segments = p->tt.nseg;
    out->max_tile_groups = p->tt.max_tile_groups;
    out->uncertain_angles = p->b.uncertain;
    out->singular_pixels = p->b.singular;
    out->build_ms = p->build_ms;
    out->hist_min = p->b.hist_min;
    out->hist_step = p->b.hist_step;
    out->device_bytes = dio::plan_device_bytes(p);
  });
}

int dio_plan_radial(dio_plan* p, double* out) {
  return guarded([&] {
    require(p && out, DIO_ERR_INVALID, "null argument");
    // histogram.rs:199-203: min + (i + 0.5) * step.
    for (uint32_t i = 0; i < p->b.n_radial; i++)
      out[i] = p->b.hist_min + ((double)i + 0.5) * p->b.hist_step;
  });
}

int dio_plan_azimuthal(dio_plan* p, double* out) {
  return guarded([&] {
    require(p && out, DIO_ERR_INVALID, "null argument");
    require(p->desc.ndim == 2, DIO_ERR_INVALID, "not a 2D plan");
    for (uint32_t i = 0; i < p->b.n_azimuthal; i++)
      out[i] = p->b.az_min + ((double)i + 0.5) * p->b.az_step;
  });
}

int dio_plan_export_csr(dio_plan* p, int masked, uint32_t* row_ptr, uint32_t* col_idx,
                        double* w64, float* w32, uint64_t capacity) {
  return guarded([&] {
    require(p != nullptr, DIO_ERR_INVALID, "null plan");
    std::lock_guard<std::mutex> lk(p->mu);
    DIO_HIP(hipSetDevice(p->device));
    const dio::Table& t = masked ? p->b.exec_table() : p->b.full;
    if (row_ptr)
      DIO_HIP(hipMemcpy(row_ptr, t.row_ptr.p, (t.nbins + 1) * sizeof(uint32_t),
                        hipMemcpyDeviceToHost));
    if (col_idx || w64 || w32) require(capacity >= t.nnz, DIO_ERR_INVALID, "capacity too small");
    if (t.nnz) {
      if (col_idx)
        DIO_HIP(hipMemcpy(col_idx, t.col.p, t.nnz * sizeof(uint32_t), hipMemcpyDeviceToHost));
      if (w64) DIO_HIP(hipMemcpy(w64, t.w64.p, t.nnz * sizeof(double), hipMemcpyDeviceToHost));
      if (w32) DIO_HIP(hipMemcpy(w32, t.w32.p, t.nnz * sizeof(float), hipMemcpyDeviceToHost));
    }
  });
}

int dio_plan_denominators(dio_plan* p, double* out) {
  return guarded([&] {
    require(p && out, DIO_ERR_INVALID, "null argument");
    DIO_HIP(hipSetDevice(p->device));
    DIO_HIP(hipMemcpy(out, p->b.den.p, p->b.den.bytes(), hipMemcpyDeviceToHost));
  });
}

int dio_plan_counts(dio_plan* p, double* out) {
  return guarded([&] {
    require(p && out, DIO_ERR_INVALID, "null argument");
    DIO_HIP(hipSetDevice(p->device));
    dio::DevBuf<double> c(std::max<uint32_t>(p->gt.nbins, 1));
    dio::bin_counts(p->b, c.p, p->stream);
    DIO_HIP(hipMemcpyAsync(out, c.p, (size_t)p->gt.nbins * sizeof(double),
                           hipMemcpyDeviceToHost, p->stream));
    DIO_HIP(hipStreamSynchronize(p->stream));
  });
}

int dio_plan_normalization(dio_plan* p, double* out) {
  return guarded([&] {
    require(p && out, DIO_ERR_INVALID, "null argument");
    DIO_HIP(hipSetDevice(p->device));
    DIO_HIP(hipMemcpy(out, p->b.norm.p, p->b.norm.bytes(), hipMemcpyDeviceToHost));
  });
}

int dio_integrate(dio_plan* p, const void* frames, int dtype, uint64_t n_frames, uint32_t flags,
                  double* intensity, double* sigma, double* std_out) {
  return guarded([&] {
    require(p != nullptr, DIO_ERR_INVALID, "null plan");
    if (n_frames == 0) return;
    require(frames && intensity, DIO_ERR_INVALID, "null frames or output");
    require(!(flags & DIO_WITH_ERRORS) || sigma, DIO_ERR_INVALID, "sigma output missing");
    require(!(flags & DIO_WITH_AZIMUTHAL_STD) || std_out, DIO_ERR_INVALID,
            "azimuthal std output missing");
    std::lock_guard<std::mutex> lk(p->mu);
    DIO_HIP(hipSetDevice(p->device));
    size_t esz = dtype_bytes(dtype);
    uint64_t fbytes = p->gt.npix * esz;
    uint32_t nb = p->gt.nbins;
    // Chunks of at most 1 GiB of frames (and at least one frame).
    uint64_t chunk = std::max<uint64_t>(1, (1ull << 30) / fbytes);
    chunk = std::min<uint64_t>(chunk, n_frames);
    if (p->stage.n < chunk * fbytes) p->stage.alloc(chunk * fbytes);
    ensure_out(p->out_i, chunk * nb);
    ensure_out(p->out_s, chunk * nb);
    ensure_out(p->out_d, chunk * nb);
    const char* src = static_cast<const char*>(frames);
    for (uint64_t f0 = 0; f0 < n_frames; f0 += chunk) {
      uint64_t nf = std::min<uint64_t>(chunk, n_frames - f0);
      DIO_HIP(hipMemcpyAsync(p->stage.p, src + f0 * fbytes, nf * fbytes, hipMemcpyHostToDevice,
                             p->stream));
      run_device(p, p->sc, p->stage.p, dtype, nf, flags, p->out_i.p, p->out_s.p, p->out_d.p, p->stream);
      DIO_HIP(hipMemcpyAsync(intensity + f0 * nb, p->out_i.p, nf * nb * 8, hipMemcpyDeviceToHost,
                             p->stream));
      if (flags & DIO_WITH_ERRORS)
        DIO_HIP(hipMemcpyAsync(sigma + f0 * nb, p->out_s.p, nf * nb * 8, hipMemcpyDeviceToHost,
                               p->stream));
      if (flags & DIO_WITH_AZIMUTHAL_STD)
        DIO_HIP(hipMemcpyAsync(std_out + f0 * nb, p->out_d.p, nf * nb * 8,
                               hipMemcpyDeviceToHost, p->stream));
      DIO_HIP(hipStreamSynchronize(p->stream));
    }
  });
}

int dio_integrate_device(dio_plan* p, const void* d_frames, int dtype, uint64_t n_frames,
                         uint32_t flags, double* d_intensity, double* d_sigma, double* d_std,
                         void* stream) {
  return guarded([&] {
    require(p != nullptr, DIO_ERR_INVALID, "null plan");
    if (n_frames == 0) return;
    require(d_frames && d_intensity, DIO_ERR_INVALID, "null frames or output");
    std::lock_guard<std::mutex> lk(p->mu);
    DIO_HIP(hipSetDevice(p->device));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : p->stream;
    run_device(p, p->sc, d_frames, dtype, n_frames, flags, d_intensity, d_sigma, d_std, s);
  });
}

int dio_integrate_device_noalloc(dio_plan* p, const void* d_frames, int dtype,
                                 uint64_t n_frames, uint32_t flags, double* d_intensity,
                                 double* d_sigma, double* d_std, void* stream) {
  return guarded([&] {
    require(p != nullptr, DIO_ERR_INVALID, "null plan");
    if (n_frames == 0) return;
    require(d_frames && d_intensity, DIO_ERR_INVALID, "null frames or output");
    std::lock_guard<std::mutex> lk(p->mu);
    DIO_HIP(hipSetDevice(p->device));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : p->stream;
    if (flags & DIO_WITH_AZIMUTHAL_STD) {
      // The spread's segment partials in the fixed scratch: passes of as
      // many frames as it holds (at most the usual pass).
      size_t cap = 0;
      void* scratch = dio::noalloc_scratch(p->device, &cap);
      if (dio::std_pass_bytes(p->tt, 1, flags & DIO_WITH_ERRORS) > cap)
        dio::fail(DIO_ERR_INVALID,
                  "the azimuthal-spread partials of one frame exceed the fixed scratch of "
                  "dio_integrate_device_noalloc (64 MiB)");
      dio::TileScratch sc;
      sc.ext = static_cast<double*>(scratch);
      sc.ext_n = cap / sizeof(double);
      run_device(p, sc, d_frames, dtype, n_frames, flags, d_intensity, d_sigma, d_std, s);
      return;
    }
    const bool var = flags & DIO_WITH_ERRORS;
    require(!var || d_sigma != nullptr, DIO_ERR_INVALID, "sigma output missing");
    (void)dtype_bytes(dtype);
    size_t cap = 0;
    void* scratch = dio::noalloc_scratch(p->device, &cap);
    // The plan's pass shape if one pass of partials fits the fixed scratch,
    // else the largest smaller one that does (a frame's bits do not depend
    // on the pass shape).
    static const uint32_t kShapes[][2] = {{16, 4}, {8, 4}, {4, 4}, {2, 2}, {1, 1}};
    uint32_t F = p->tile, G = p->lanes;
    const int prec = p->desc.precision;
    if (dio::pass_bytes(p->tt, F, var, prec) > cap) {
      F = 0;
      for (const auto& sh : kShapes) {
        if (sh[0] < p->tile && dio::valid_tile(prec, sh[0], sh[1]) &&
            dio::pass_bytes(p->tt, sh[0], var, prec) <= cap) {
          F = sh[0];
          G = sh[1];
          break;
        }
      }
      if (!F)
        dio::fail(DIO_ERR_INVALID,
                  "the segment partials of one frame exceed the fixed scratch of "
                  "dio_integrate_device_noalloc (64 MiB)");
    }
    const uint32_t passes = (uint32_t)std::min<size_t>(
        dio::passes_per_launch(p->tt, F, dtype, prec), cap / dio::pass_bytes(p->tt, F, var, prec));
    dio::TileScratch sc;
    sc.ext = static_cast<double*>(scratch);
    sc.ext_n = cap / sizeof(double);
    dio::tiled_gather(p->tt, sc, d_frames, dtype, n_frames, var, prec, F, G, d_intensity,
                      d_sigma, s, nullptr, (flags & DIO_RAW_SUMS) != 0, std::max(1u, passes));
  });
}

// Dynamic masks. With a plan mask the execution table is compacted, so the
// CSR kernel runs on the unmasked table; otherwise the tile kernel's raw sums
// of the zeroed signal and of the zeroed normalization give numerator,
// variance and the per-frame denominator, in chunks of 64 frames.
static void run_masked(dio_plan* p, const void* frames, int dtype, uint64_t n_frames,
                       const uint8_t* masks, bool var, double* I, double* S, hipStream_t s) {
  if (p->b.has_exec) {
    masked_frames(p->st, frames, dtype, masks, n_frames, var, I, S, s);
    return;
  }
  const uint64_t npix = p->gt.npix, esz = dtype_bytes(dtype);
  const uint32_t nb = p->st.nbins;
  const bool f64 = p->desc.precision == DIO_F64;
  const int ndt = f64 ? DIO_F64T : DIO_F32T;
  const uint64_t chunk = std::min<uint64_t>(n_frames, 64);
  // Frames per lane of the masked-row kernel: 2 with the variance sums (4
  // needs more registers than the kernel holds), else the plan's.
  const uint32_t lanes = var ? 2 : p->tile == 16 ? p->lanes : 4;
  if (masked_rows_supported(p->tt, frames, dtype, p->desc.precision, lanes)) {
    // The masks folded into the tile kernel's staging: 8 frames and their
    // masked normalization per 16-row pass, no masked copies in HBM.
    const uint32_t FP = dio::masked_rows_frames_per_pass();
    const uint64_t npass = (chunk + FP - 1) / FP;
    if (p->norm32.n < npix) {
      p->norm32.alloc(npix);
      dio::norm_to_f32(p->st.norm, npix, p->norm32.p, s);
    }
    const uint64_t mstride = dio::mask_bits_stride(npix);
    if (p->dyn_m.n < npass * mstride) p->dyn_m.alloc(npass * mstride);
    if (p->dyn_r.n < 2 * npass * 16 * nb) p->dyn_r.alloc(2 * npass * 16 * nb);
    const char* src = static_cast<const char*>(frames);
    for (uint64_t f0 = 0; f0 < n_frames; f0 += chunk) {
      const uint64_t nf = std::min<uint64_t>(chunk, n_frames - f0);
      dio::mask_bits(masks + f0 * npix, npix, (uint32_t)nf, p->dyn_m.p, s);
      double* rs = p->dyn_r.p;
      double* rv = p->dyn_r.p + npass * 16 * nb;
      tiled_gather_masked(p->tt, p->sc, src + f0 * npix * esz, dtype, p->dyn_m.p, mstride,
                          p->norm32.p, nf, var, p->desc.precision, lanes, rs, rv, s);
      dio::masked_rows_finish(rs, rv, nf, nb, I + f0 * nb, var ? S + f0 * nb : nullptr, s);
    }
    return;
  }
  if (p->dyn_x.n < chunk * npix * esz) p->dyn_x.alloc(chunk * npix * esz);
  if (p->dyn_n.n < chunk * npix * (f64 ? 8 : 4)) p->dyn_n.alloc(chunk * npix * (f64 ? 8 : 4));
  if (p->dyn_d.n < chunk * nb) p->dyn_d.alloc(chunk * nb);
  const char* src = static_cast<const char*>(frames);
  for (uint64_t f0 = 0; f0 < n_frames; f0 += chunk) {
    const uint64_t nf = std::min<uint64_t>(chunk, n_frames - f0);
    masked_prepare(p->st, src + f0 * npix * esz, dtype, masks + f0 * npix, nf, f64, p->dyn_x.p,
                   p->dyn_n.p, s);
    tiled_gather(p->tt, p->sc, p->dyn_x.p, dtype, nf, var, p->desc.precision, p->tile, p->lanes,
                 I + f0 * nb, var ? S + f0 * nb : nullptr, s, nullptr, true);
    tiled_gather(p->tt, p->sc, p->dyn_n.p, ndt, nf, false, p->desc.precision, p->tile, p->lanes,
                 p->dyn_d.p, nullptr, s, nullptr, true);
    dio::masked_finish(I + f0 * nb, var ? S + f0 * nb : nullptr, p->dyn_d.p, nf * nb, var, s);
  }
}

static void check_masked(dio_plan* p, uint32_t flags) {
  require((flags & ~(uint32_t)DIO_WITH_ERRORS) == 0, DIO_ERR_UNSUPPORTED,
          "per-frame masks take DIO_WITH_ERRORS only");
  require(p->desc.ndim == 1, DIO_ERR_UNSUPPORTED, "per-frame masks are defined for 1D plans");
}

int dio_integrate_masked(dio_plan* p, const void* frames, int dtype, uint64_t n_frames,
                         const uint8_t* masks, uint32_t flags, double* intensity, double* sigma) {
  return guarded([&] {
    require(p != nullptr, DIO_ERR_INVALID, "null plan");
    check_masked(p, flags);
    if (n_frames == 0) return;
    require(frames && masks && intensity, DIO_ERR_INVALID, "null frames, masks or output");
    const bool var = flags & DIO_WITH_ERRORS;
    require(!var || sigma, DIO_ERR_INVALID, "sigma output missing");
    std::lock_guard<std::mutex> lk(p->mu);
    DIO_HIP(hipSetDevice(p->device));
    const uint64_t npix = p->gt.npix, fbytes = npix * dtype_bytes(dtype);
    const uint32_t nb = p->st.nbins;
    // Chunks of at most 1 GiB of frames and masks (and at least one frame).
    uint64_t chunk = std::max<uint64_t>(1, (1ull << 30) / (fbytes + npix));
    chunk = std::min<uint64_t>(chunk, n_frames);
    if (p->stage.n < chunk * fbytes) p->stage.alloc(chunk * fbytes);
    if (p->mstage.n < chunk * npix) p->mstage.alloc(chunk * npix);
    ensure_out(p->out_i, chunk * nb);
    ensure_out(p->out_s, chunk * nb);
    const char* src = static_cast<const char*>(frames);
    for (uint64_t f0 = 0; f0 < n_frames; f0 += chunk) {
      const uint64_t nf = std::min<uint64_t>(chunk, n_frames - f0);
      DIO_HIP(hipMemcpyAsync(p->stage.p, src + f0 * fbytes, nf * fbytes, hipMemcpyHostToDevice,
                             p->stream));
      DIO_HIP(hipMemcpyAsync(p->mstage.p, masks + f0 * npix, nf * npix, hipMemcpyHostToDevice,
                             p->stream));
      run_masked(p, p->stage.p, dtype, nf, p->mstage.p, var, p->out_i.p, p->out_s.p, p->stream);
      DIO_HIP(hipMemcpyAsync(intensity + f0 * nb, p->out_i.p, nf * nb * 8, hipMemcpyDeviceToHost,
                             p->stream));
      if (var)
        DIO_HIP(hipMemcpyAsync(sigma + f0 * nb, p->out_s.p, nf * nb * 8, hipMemcpyDeviceToHost,
                               p->stream));
      DIO_HIP(hipStreamSynchronize(p->stream));
    }
  });
}

int dio_integrate_masked_device(dio_plan* p, const void* d_frames, int dtype, uint64_t n_frames,
                                const uint8_t* d_masks, uint32_t flags, double* d_intensity,
                                double* d_sigma, void* stream) {
  return guarded([&] {
    require(p != nullptr, DIO_ERR_INVALID, "null plan");
    check_masked(p, flags);
    if (n_frames == 0) return;
    require(d_frames && d_masks && d_intensity, DIO_ERR_INVALID, "null frames, masks or output");
    const bool var = flags & DIO_WITH_ERRORS;
    require(!var || d_sigma, DIO_ERR_INVALID, "sigma output missing");
    dtype_bytes(dtype);
    std::lock_guard<std::mutex> lk(p->mu);
    DIO_HIP(hipSetDevice(p->device));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : p->stream;
    run_masked(p, d_frames, dtype, n_frames, d_masks, var, d_intensity, d_sigma, s);
  });
}

int dio_plan_set_tile(dio_plan* p, uint32_t tile, uint32_t lanes) {
  return guarded([&] {
    require(p != nullptr, DIO_ERR_INVALID, "null plan");
    std::lock_guard<std::mutex> lk(p->mu);
    if (lanes == 0) lanes = std::min<uint32_t>(tile, 4);
    require(dio::valid_tile(p->desc.precision, tile, lanes), DIO_ERR_INVALID,
            "frames per pass must be a power of two in 1..32 (1..16 for f64) and frames per "
            "lane a power of two in 1..min(4, frames per pass)");
    p->tile = tile;
    p->lanes = lanes;
  });
}

int dio_plan_autotune(dio_plan* p, int dtype, uint64_t n_frames_hint, uint32_t flags,
                      uint32_t* frames_per_tile, uint32_t* frames_per_lane) {
  return guarded([&] {
    require(p != nullptr, DIO_ERR_INVALID, "null plan");
    std::lock_guard<std::mutex> lk(p->mu);
    DIO_HIP(hipSetDevice(p->device));
    size_t esz = dtype_bytes(dtype);
    bool var = flags & DIO_WITH_ERRORS;
    // Sweep on 64 zero frames (or fewer when the hint is smaller): timing of
    // a gather does not depend on pixel values.
    uint64_t nf = std::max<uint64_t>(1, std::min<uint64_t>(64, n_frames_hint ? n_frames_hint : 64));
    dio::DevBuf<uint8_t> fr(nf * p->gt.npix * esz);
    DIO_HIP(hipMemsetAsync(fr.p, 0, fr.bytes(), p->stream));
    dio::DevBuf<double> oi(nf * p->gt.nbins), os(var ? nf * p->gt.nbins : 1);
    hipEvent_t e0, e1;
    DIO_HIP(hipEventCreate(&e0));
    DIO_HIP(hipEventCreate(&e1));
    // Two sweeps over the candidates; each keeps its faster sweep, so a
    // candidate timed during a transient slowdown is not ruled out by it.
    std::vector<std::pair<uint32_t, uint32_t>> cand;
    for (uint32_t tile : {1u, 2u, 4u, 8u, 16u, 32u}) {
      if (tile > 1 && tile > nf * 2) break;
      for (uint32_t lanes : {1u, 2u, 4u})
        if (dio::valid_tile(p->desc.precision, tile, lanes)) cand.emplace_back(tile, lanes);
    }
    std::vector<double> t_best(cand.size(), 1e300);
    for (int sweep = 0; sweep < 2; sweep++) {
      for (size_t c = 0; c < cand.size(); c++) {
        const uint32_t tile = cand[c].first, lanes = cand[c].second;
        dio::tiled_gather(p->tt, p->sc, fr.p, dtype, nf, var, p->desc.precision, tile, lanes,
                          oi.p, os.p, p->stream);  // warm-up
        DIO_HIP(hipEventRecord(e0, p->stream));
        for (int rep = 0; rep < 3; rep++)
          dio::tiled_gather(p->tt, p->sc, fr.p, dtype, nf, var, p->desc.precision, tile, lanes,
                            oi.p, os.p, p->stream);
        DIO_HIP(hipEventRecord(e1, p->stream));
        DIO_HIP(hipEventSynchronize(e1));
        float ms = 0;
        DIO_HIP(hipEventElapsedTime(&ms, e0, e1));
        t_best[c] = std::min(t_best[c], (double)ms);
      }
    }
    uint32_t best_tile = p->tile, best_lanes = p->lanes;
    double best = 1e300;
    for (size_t c = 0; c < cand.size(); c++)
      if (t_best[c] < best) {
        best = t_best[c];
        best_tile = cand[c].first;
        best_lanes = cand[c].second;
      }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    p->tile = best_tile;
    p->lanes = best_lanes;
    if (frames_per_tile) *frames_per_tile = best_tile;
    if (frames_per_lane) *frames_per_lane = best_lanes;
  });
}

int dio_graph_create(dio_plan* p, const void* d_frames, int dtype, uint64_t n_frames,
                     uint32_t flags, double* d_intensity, double* d_sigma, double* d_std,
                     dio_graph** out) {
  return guarded([&] {
    require(p && out && d_frames && d_intensity, DIO_ERR_INVALID, "null argument");
    *out = nullptr;
    std::lock_guard<std::mutex> lk(p->mu);
    DIO_HIP(hipSetDevice(p->device));
    // Scratch must exist before capture: allocation is not capturable.
    auto g = std::make_unique<dio_graph>();
    g->plan = p;
    g->sc.ensure(p->tt, p->tile, flags & DIO_WITH_ERRORS);
    if (flags & DIO_WITH_AZIMUTHAL_STD)
      g->sc.ensure_bytes(dio::std_pass_bytes(p->tt, dio::std_frames_per_pass(),
                                             flags & DIO_WITH_ERRORS));
    DIO_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    try {
      run_device(p, g->sc, d_frames, dtype, n_frames, flags, d_intensity, d_sigma, d_std,
                 p->stream);
    } catch (...) {
      hipGraph_t tmp = nullptr;
      (void)hipStreamEndCapture(p->stream, &tmp);
      if (tmp) (void)hipGraphDestroy(tmp);
      throw;
    }
    DIO_HIP(hipStreamEndCapture(p->stream, &g->graph));
    DIO_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    *out = g.release();
  });
}

int dio_graph_launch(dio_graph* g, void* stream) {
  return guarded([&] {
    require(g != nullptr, DIO_ERR_INVALID, "null graph");
    DIO_HIP(hipSetDevice(g->plan->device));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->plan->stream;
    DIO_HIP(hipGraphLaunch(g->exec, s));
  });
}

int dio_graph_destroy(dio_graph* g) {
  return guarded([&] { delete g; });
}

int dio_plan_time_passes(dio_plan* p, const void* d_frames, int dtype, uint64_t n_frames,
                         uint32_t flags, double* d_intensity, double* d_sigma, float* ms,
                         uint32_t* frames, uint32_t cap, uint32_t* n_launches) {
  return guarded([&] {
    require(p && d_frames && d_intensity && n_frames > 0, DIO_ERR_INVALID, "null argument");
    require(!(flags & DIO_WITH_AZIMUTHAL_STD), DIO_ERR_UNSUPPORTED,
            "pass timing covers the tiled gather only");
    std::lock_guard<std::mutex> lk(p->mu);
    DIO_HIP(hipSetDevice(p->device));
    const uint64_t passes = (n_frames + p->tile - 1) / p->tile;  // launches <= passes
    std::vector<hipEvent_t> ev(2 * passes, nullptr);
    struct Events {
      std::vector<hipEvent_t>& v;
      ~Events() {
        for (hipEvent_t e : v)
          if (e) (void)hipEventDestroy(e);
      }
    } guard{ev};
    for (hipEvent_t& e : ev) DIO_HIP(hipEventCreate(&e));
    std::vector<uint32_t> fr(passes, 0);
    dio::LaunchEvents le;
    le.ev = ev.data();
    le.cap = passes;
    le.frames = fr.data();
    run_device(p, p->sc, d_frames, dtype, n_frames, flags, d_intensity, d_sigma, nullptr,
               p->stream, &le);
    require(le.n > 0, DIO_ERR_DEVICE, "no tile kernel launch was timed");
    DIO_HIP(hipEventSynchronize(ev[2 * le.n - 1]));
    if (n_launches) *n_launches = (uint32_t)le.n;
    for (uint64_t k = 0; k < le.n && k < cap; k++) {
      float t = 0.f;
      DIO_HIP(hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]));
      ms[k] = t;
      if (frames) frames[k] = fr[k];
    }
  });
}

int dio_plan_time_gather(dio_plan* p, const void* d_frames, int dtype, uint64_t n_frames,
                         uint32_t flags, double* d_intensity, double* d_sigma, int reps,
                         double* avg_ms) {
  return guarded([&] {
    require(p && d_frames && avg_ms && reps > 0, DIO_ERR_INVALID, "null argument");
    (void)d_intensity;
    (void)d_sigma;
    std::lock_guard<std::mutex> lk(p->mu);
    DIO_HIP(hipSetDevice(p->device));
    bool var = flags & DIO_WITH_ERRORS;
    uint32_t nf = (uint32_t)std::min<uint64_t>(n_frames, p->tile);
    hipEvent_t e0, e1;
    DIO_HIP(hipEventCreate(&e0));
    DIO_HIP(hipEventCreate(&e1));
    dio::tiled_pass_only(p->tt, p->sc, d_frames, dtype, nf, var, p->desc.precision, p->tile,
                         p->lanes, p->stream);
    DIO_HIP(hipEventRecord(e0, p->stream));
    for (int r = 0; r < reps; r++)
      dio::tiled_pass_only(p->tt, p->sc, d_frames, dtype, nf, var, p->desc.precision, p->tile,
                           p->lanes, p->stream);
    DIO_HIP(hipEventRecord(e1, p->stream));
    DIO_HIP(hipEventSynchronize(e1));
    float ms = 0;
    DIO_HIP(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  });
}

}  // extern "C"
