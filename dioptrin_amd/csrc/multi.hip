// Supplied pixel variances and multi-detector merging.
//
// Supplied variances (Prepared1d::integrate_with_variances,
// src/lut/prepared.rs:737-766; Integrator::integrate2d_raw_on_grid_with_
// variances_from_slice, src/integration/grid.rs:347-392): per bin
// sum_signal = sum w x and sum_variance = sum w^2 v, v the caller's variance
// plane. The tile kernel already forms sum w^2 max(x, 0) of whatever it
// stages, and a supplied variance is finite and nonnegative (checked here,
// grid.rs:1102-1107), so the variance plane runs through the same kernel as
// a second frame set whose "sigma" output is sum w^2 v, normalised by
// k_tfinal exactly as normalize_bin does (prepared.rs:985-996).
//
// Multi-detector merging (MultiIntegrator, src/multi.rs:198-316, 346-510):
// every detector's raw sums come from its own plan on one shared bin grid;
// k_merge adds them on the device in detector order and divides once.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "plan.hpp"

namespace {

using dio::DevBuf;
using dio::require;

// One detector's device sums for k_merge.
struct MergeSrc {
  const double* sig;  // n_frames x nbins
  const double* var;  // n_frames x nbins, or null
  const double* den;  // nbins: sum_norm of the detector's table
  const double* cnt;  // nbins: count
  double factor;      // multiplies den (monitor x solid-angle scale)
};

// Per (frame, bin): S = sum_d sig, N = sum_d den * factor, V = sum_d var,
// C = sum_d cnt, each summed from 0.0 in detector order (multi.rs:274-297);
// I = S / N and sigma = sqrt(max(V / N^2, 0)) (1D, multi.rs:396-418) or
// sqrt(max(V, 0)) / N (2D, multi.rs:298-311), 0 where N == 0.
__global__ void k_merge(const MergeSrc* src, uint32_t ndet, uint64_t nbins, uint64_t total,
                        int form2d, double* I, double* S, double* raw_sig, double* raw_var,
                        double* raw_norm, double* raw_cnt) {
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const uint64_t b = i % nbins;
  const bool var = S || raw_var;
  double s = 0.0, n = 0.0, v = 0.0, c = 0.0;
  for (uint32_t d = 0; d < ndet; d++) {
    s += src[d].sig[i];
    n += src[d].den[b] * src[d].factor;
    if (var) v += src[d].var[i];
    c += src[d].cnt[b];
  }
  I[i] = n == 0.0 ? 0.0 : s / n;
  if (S) {
    double e = 0.0;
    if (n != 0.0) e = form2d ? sqrt(fmax(v, 0.0)) / n : sqrt(fmax(v / (n * n), 0.0));
    S[i] = e;
  }
  if (raw_sig) raw_sig[i] = s;
  if (raw_var) raw_var[i] = v;
  if (i < nbins) {
    if (raw_norm) raw_norm[i] = n;
    if (raw_cnt) raw_cnt[i] = c;
  }
}

// A supplied variance plane must be finite and nonnegative (grid.rs:1102-1107).
void check_variances(const void* v, int dtype, uint64_t n) {
  bool ok = true;
  switch (dtype) {
    case DIO_F64T: {
      const double* x = static_cast<const double*>(v);
      for (uint64_t i = 0; i < n && ok; i++) ok = std::isfinite(x[i]) && x[i] >= 0.0;
      break;
    }
    case DIO_F32T: {
      const float* x = static_cast<const float*>(v);
      for (uint64_t i = 0; i < n && ok; i++) ok = std::isfinite(x[i]) && x[i] >= 0.0f;
      break;
    }
    case DIO_I32: {
      const int32_t* x = static_cast<const int32_t*>(v);
      for (uint64_t i = 0; i < n && ok; i++) ok = x[i] >= 0;
      break;
    }
    default:  // unsigned planes
      (void)dio::dtype_bytes(dtype);
  }
  require(ok, DIO_ERR_INVALID,
          "pixel variances must match the image and be finite and nonnegative");
}

// Locks each distinct plan once, in address order.
struct PlanLocks {
  std::vector<std::unique_lock<std::mutex>> held;
  explicit PlanLocks(std::vector<dio_plan*> ps) {
    std::sort(ps.begin(), ps.end());
    ps.erase(std::unique(ps.begin(), ps.end()), ps.end());
    for (dio_plan* p : ps) held.emplace_back(p->mu);
  }
};

}  // namespace

extern "C" {

int dio_integrate_variances(dio_plan* p, const void* frames, int dtype, const void* variances,
                            int vdtype, uint64_t n_frames, uint32_t flags, double* intensity,
                            double* sigma) {
  return dio::guarded([&] {
    require(p != nullptr, DIO_ERR_INVALID, "null plan");
    if (n_frames == 0) return;
    require(frames && variances && intensity && sigma, DIO_ERR_INVALID,
            "null frames, variances or output");
    require(!(flags & ~(uint32_t)(DIO_WITH_ERRORS | DIO_RAW_SUMS)), DIO_ERR_INVALID,
            "supplied variances take DIO_WITH_ERRORS and DIO_RAW_SUMS only");
    const uint64_t npix = p->gt.npix;
    const size_t esz = dio::dtype_bytes(dtype), vsz = dio::dtype_bytes(vdtype);
    check_variances(variances, vdtype, n_frames * npix);
    std::lock_guard<std::mutex> lk(p->mu);
    DIO_HIP(hipSetDevice(p->device));
    const uint32_t nb = p->gt.nbins;
    const uint32_t raw = flags & DIO_RAW_SUMS;
    // Chunks of at most 1 GiB of frames and variances (at least one frame).
    uint64_t chunk = std::max<uint64_t>(1, (1ull << 30) / (npix * (esz + vsz)));
    chunk = std::min<uint64_t>(chunk, n_frames);
    DevBuf<uint8_t> fx(chunk * npix * esz), vx(chunk * npix * vsz);
    DevBuf<double> oi(chunk * nb), os(chunk * nb), junk(chunk * nb);
    const char* fsrc = static_cast<const char*>(frames);
    const char* vsrc = static_cast<const char*>(variances);
    for (uint64_t f0 = 0; f0 < n_frames; f0 += chunk) {
      const uint64_t nf = std::min<uint64_t>(chunk, n_frames - f0);
      DIO_HIP(hipMemcpyAsync(fx.p, fsrc + f0 * npix * esz, nf * npix * esz,
                             hipMemcpyHostToDevice, p->stream));
      DIO_HIP(hipMemcpyAsync(vx.p, vsrc + f0 * npix * vsz, nf * npix * vsz,
                             hipMemcpyHostToDevice, p->stream));
      // Intensity (or sum_signal) from the frames, sigma (or sum_variance)
      // from the variance planes.
      dio::run_device(p, p->sc, fx.p, dtype, nf, raw, oi.p, nullptr, nullptr, p->stream);
      dio::run_device(p, p->sc, vx.p, vdtype, nf, raw | DIO_WITH_ERRORS, junk.p, os.p, nullptr,
                      p->stream);
      DIO_HIP(hipMemcpyAsync(intensity + f0 * nb, oi.p, nf * nb * 8, hipMemcpyDeviceToHost,
                             p->stream));
      DIO_HIP(hipMemcpyAsync(sigma + f0 * nb, os.p, nf * nb * 8, hipMemcpyDeviceToHost,
                             p->stream));
      DIO_HIP(hipStreamSynchronize(p->stream));
    }
  });
}

int dio_multi_integrate(uint32_t n_det, dio_plan* const* plans, const void* const* frames,
                        const int* dtypes, const void* const* variances, const int* vdtypes,
                        const double* factors, uint64_t n_frames, uint32_t flags,
                        double* intensity, double* sigma, dio_raw_sums* raw) {
  return dio::guarded([&] {
    require(n_det > 0 && plans && frames && dtypes, DIO_ERR_INVALID,
            "at least one detector with a plan, frames and dtype is required");
    require(!(flags & ~(uint32_t)DIO_WITH_ERRORS), DIO_ERR_INVALID,
            "merging takes DIO_WITH_ERRORS only");
    const bool supplied = variances != nullptr;
    require(!supplied || vdtypes, DIO_ERR_INVALID, "variance dtypes missing");
    const bool errors = supplied || (flags & DIO_WITH_ERRORS);
    require(!errors || sigma || (raw && raw->sum_variance), DIO_ERR_INVALID,
            "sigma output missing");
    if (n_frames == 0) return;
    require(intensity != nullptr, DIO_ERR_INVALID, "null output");
    std::vector<dio_plan*> ps(plans, plans + n_det);
    for (uint32_t d = 0; d < n_det; d++) {
      require(ps[d] != nullptr && frames[d] != nullptr, DIO_ERR_INVALID,
              "null plan or frames");
      require(!supplied || variances[d] != nullptr, DIO_ERR_INVALID, "null variance plane");
      require(ps[d]->device == ps[0]->device, DIO_ERR_INVALID,
              "all detector plans must be on one device");
      require(ps[d]->gt.nbins == ps[0]->gt.nbins && ps[d]->desc.ndim == ps[0]->desc.ndim,
              DIO_ERR_INVALID, "all detector plans must share one bin grid");
      const double f = factors ? factors[d] : 1.0;
      require(std::isfinite(f) && f > 0.0, DIO_ERR_INVALID,
              "normalization factors must be finite and positive");
      (void)dio::dtype_bytes(dtypes[d]);
      if (supplied) check_variances(variances[d], vdtypes[d], n_frames * ps[d]->gt.npix);
    }
    PlanLocks locks(ps);
    DIO_HIP(hipSetDevice(ps[0]->device));
    hipStream_t s = ps[0]->stream;
    const uint64_t nb = ps[0]->gt.nbins;
    // Frames per chunk: staged frames and variances of every detector within
    // 1 GiB (at least one frame).
    uint64_t per_frame = 0;
    for (uint32_t d = 0; d < n_det; d++)
      per_frame += ps[d]->gt.npix *
                   (dio::dtype_bytes(dtypes[d]) + (supplied ? dio::dtype_bytes(vdtypes[d]) : 0));
    const uint64_t chunk =
        std::min<uint64_t>(n_frames, std::max<uint64_t>(1, (1ull << 30) / per_frame));
    std::vector<DevBuf<uint8_t>> fx(n_det), vx(n_det);
    std::vector<DevBuf<double>> sig(n_det), var(n_det), cnt(n_det);
    std::vector<MergeSrc> host_src(n_det);
    for (uint32_t d = 0; d < n_det; d++) {
      dio_plan* p = ps[d];
      fx[d].alloc(chunk * p->gt.npix * dio::dtype_bytes(dtypes[d]));
      if (supplied) vx[d].alloc(chunk * p->gt.npix * dio::dtype_bytes(vdtypes[d]));
      sig[d].alloc(chunk * nb);
      if (errors) var[d].alloc(chunk * nb);
      cnt[d].alloc(nb);
      dio::bin_counts(p->b, cnt[d].p, s);
      host_src[d] = MergeSrc{sig[d].p, errors ? var[d].p : nullptr, p->b.den.p, cnt[d].p,
                             factors ? factors[d] : 1.0};
    }
    DevBuf<MergeSrc> src(n_det);
    DevBuf<double> oi(chunk * nb), os(errors ? chunk * nb : 0), junk(supplied ? chunk * nb : 0);
    DevBuf<double> rs(raw && raw->sum_signal ? chunk * nb : 0);
    DevBuf<double> rv(raw && raw->sum_variance && errors ? chunk * nb : 0);
    DevBuf<double> rn(raw && raw->sum_norm ? nb : 0), rc(raw && raw->count ? nb : 0);
    DIO_HIP(hipMemcpyAsync(src.p, host_src.data(), n_det * sizeof(MergeSrc),
                           hipMemcpyHostToDevice, s));
    const int form2d = ps[0]->desc.ndim == 2;
    for (uint64_t f0 = 0; f0 < n_frames; f0 += chunk) {
      const uint64_t nf = std::min<uint64_t>(chunk, n_frames - f0);
      for (uint32_t d = 0; d < n_det; d++) {
        dio_plan* p = ps[d];
        const size_t esz = dio::dtype_bytes(dtypes[d]);
        DIO_HIP(hipMemcpyAsync(fx[d].p,
                               static_cast<const char*>(frames[d]) + f0 * p->gt.npix * esz,
                               nf * p->gt.npix * esz, hipMemcpyHostToDevice, s));
        if (supplied) {
          const size_t vsz = dio::dtype_bytes(vdtypes[d]);
          DIO_HIP(hipMemcpyAsync(vx[d].p,
                                 static_cast<const char*>(variances[d]) + f0 * p->gt.npix * vsz,
                                 nf * p->gt.npix * vsz, hipMemcpyHostToDevice, s));
          dio::run_device(p, p->sc, fx[d].p, dtypes[d], nf, DIO_RAW_SUMS, sig[d].p, nullptr,
                          nullptr, s);
          dio::run_device(p, p->sc, vx[d].p, vdtypes[d], nf, DIO_RAW_SUMS | DIO_WITH_ERRORS,
                          junk.p, var[d].p, nullptr, s);
        } else {
          dio::run_device(p, p->sc, fx[d].p, dtypes[d], nf,
                          DIO_RAW_SUMS | (errors ? DIO_WITH_ERRORS : 0), sig[d].p,
                          errors ? var[d].p : nullptr, nullptr, s);
        }
      }
      const uint64_t total = nf * nb;
      k_merge<<<dio::blocks_for(total, 256), 256, 0, s>>>(
          src.p, n_det, nb, total, form2d, oi.p, sigma ? os.p : nullptr, rs.p, rv.p,
          f0 == 0 ? rn.p : nullptr, f0 == 0 ? rc.p : nullptr);
      DIO_HIP(hipGetLastError());
      DIO_HIP(hipMemcpyAsync(intensity + f0 * nb, oi.p, total * 8, hipMemcpyDeviceToHost, s));
      if (sigma && errors)
        DIO_HIP(hipMemcpyAsync(sigma + f0 * nb, os.p, total * 8, hipMemcpyDeviceToHost, s));
      if (rs.p)
        DIO_HIP(hipMemcpyAsync(raw->sum_signal + f0 * nb, rs.p, total * 8,
                               hipMemcpyDeviceToHost, s));
      if (rv.p)
        DIO_HIP(hipMemcpyAsync(raw->sum_variance + f0 * nb, rv.p, total * 8,
                               hipMemcpyDeviceToHost, s));
      if (f0 == 0 && rn.p)
        DIO_HIP(hipMemcpyAsync(raw->sum_norm, rn.p, nb * 8, hipMemcpyDeviceToHost, s));
      if (f0 == 0 && rc.p)
        DIO_HIP(hipMemcpyAsync(raw->count, rc.p, nb * 8, hipMemcpyDeviceToHost, s));
      DIO_HIP(hipStreamSynchronize(s));
    }
  });
}

}  // extern "C"
