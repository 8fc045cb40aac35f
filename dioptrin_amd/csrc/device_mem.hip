// Device memory, synchronisation and seeded synthetic frames.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

struct dio_ctx;
extern "C" int dio_ctx_device(dio_ctx* ctx);

namespace {

thread_local std::string g_err;

template <class Fn>
int guarded_local(Fn&& fn) {
  try {
    fn();
    return DIO_OK;
  } catch (const dio::Error& e) {
    dio_set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    dio_set_last_error(e.what());
    return DIO_ERR_DEVICE;
  }
}

// SplitMix64 finaliser (Steele, Lea & Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014): a counter-based stream per (frame, pixel).
__device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ inline double uniform01(uint64_t key, uint32_t& ctr) {
  uint64_t r = mix64(key + 0x9e3779b97f4a7c15ull * (uint64_t)(++ctr));
  return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Poisson variate: inversion by sequential search for small means, the
// transformed rejection method PTRS of W. Hormann, "The transformed rejection
// method for generating Poisson random variables", Insurance: Mathematics and
// Economics 12 (1993) 39-45, for means >= 10.
__device__ inline uint32_t poisson(double lam, uint64_t key) {
  uint32_t ctr = 0;
  if (!(lam > 0.0)) return 0;
  if (lam < 10.0) {
    double p = exp(-lam), s = p, u = uniform01(key, ctr);
    uint32_t k = 0;
    while (u > s && k < 1000) {
      k++;
      p *= lam / (double)k;
      s += p;
    }
    return k;
  }
  double slam = sqrt(lam), loglam = log(lam);
  double b = 0.931 + 2.53 * slam;
  double a = -0.059 + 0.02483 * b;
  double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
  double vr = 0.9277 - 3.6224 / (b - 2.0);
  for (int it = 0; it < 64; it++) {
    double U = uniform01(key, ctr) - 0.5;
    double V = uniform01(key, ctr);
    double us = 0.5 - fabs(U);
    double k = floor((2.0 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return (uint32_t)k;
    if (k < 0.0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(inv_alpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0))
      return (uint32_t)k;
  }
  return (uint32_t)floor(lam + 0.5);
}

// Plain streaming copy, 16 bytes per lane (also the reference kernel for
// calibrating memory-traffic counters).
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_copy1(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void k_poisson(const float* rate, uint64_t npix, const double* bg, uint64_t n_frames,
                          uint64_t seed, uint32_t* out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix * n_frames) return;
  uint64_t f = i / npix, p = i % npix;
  uint64_t key = mix64(seed * 0x100000001b3ull + f) ^ mix64(p + 0x632be59bd9b4e019ull);
  out[i] = poisson((double)rate[p] + bg[f], key);
}

}  // namespace

extern "C" {

int dio_device_alloc(dio_ctx* ctx, uint64_t bytes, void** d_ptr) {
  return guarded_local([&] {
    if (!ctx || !d_ptr) dio::fail(DIO_ERR_INVALID, "null argument");
    DIO_HIP(hipSetDevice(dio_ctx_device(ctx)));
    *d_ptr = nullptr;
    if (bytes) DIO_HIP(hipMalloc(d_ptr, bytes));
  });
}

int dio_device_free(dio_ctx* ctx, void* d_ptr) {
  return guarded_local([&] {
    if (!ctx) dio::fail(DIO_ERR_INVALID, "null context");
    DIO_HIP(hipSetDevice(dio_ctx_device(ctx)));
    if (d_ptr) DIO_HIP(hipFree(d_ptr));
  });
}

int dio_host_alloc(dio_ctx* ctx, uint64_t bytes, void** h_ptr) {
  return guarded_local([&] {
    if (!ctx || !h_ptr) dio::fail(DIO_ERR_INVALID, "null argument");
    DIO_HIP(hipSetDevice(dio_ctx_device(ctx)));
    *h_ptr = nullptr;
    if (bytes) DIO_HIP(hipHostMalloc(h_ptr, bytes, hipHostMallocDefault));
  });
}

int dio_host_free(dio_ctx* ctx, void* h_ptr) {
  return guarded_local([&] {
    if (!ctx) dio::fail(DIO_ERR_INVALID, "null context");
    DIO_HIP(hipSetDevice(dio_ctx_device(ctx)));
    if (h_ptr) DIO_HIP(hipHostFree(h_ptr));
  });
}

int dio_memcpy(dio_ctx* ctx, void* dst, const void* src, uint64_t bytes, int kind) {
  return guarded_local([&] {
    if (!ctx || (bytes && (!dst || !src))) dio::fail(DIO_ERR_INVALID, "null argument");
    DIO_HIP(hipSetDevice(dio_ctx_device(ctx)));
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                    : kind == 1 ? hipMemcpyDeviceToHost
                                : hipMemcpyDeviceToDevice;
    if (kind < 0 || kind > 2) dio::fail(DIO_ERR_INVALID, "unknown copy kind");
    if (!bytes) return;
    if (kind == 2 && (reinterpret_cast<uintptr_t>(dst) % 16 == 0) &&
        (reinterpret_cast<uintptr_t>(src) % 16 == 0)) {
      uint64_t n16 = bytes / 16, tail = bytes % 16;
      if (n16)
        k_copy16<<<2048, 256>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16);
      if (tail)
        k_copy1<<<1, 64>>>(static_cast<const uint8_t*>(src) + n16 * 16,
                           static_cast<uint8_t*>(dst) + n16 * 16, tail);
      DIO_HIP(hipGetLastError());
      DIO_HIP(hipDeviceSynchronize());
      return;
    }
    DIO_HIP(hipMemcpy(dst, src, bytes, k));
  });
}

int dio_synchronize(dio_ctx* ctx) {
  return guarded_local([&] {
    if (!ctx) dio::fail(DIO_ERR_INVALID, "null context");
    DIO_HIP(hipSetDevice(dio_ctx_device(ctx)));
    DIO_HIP(hipDeviceSynchronize());
  });
}

int dio_synthetic_poisson_u32(dio_ctx* ctx, const float* d_rate, uint64_t npix,
                              const double* background, uint64_t n_frames, uint64_t seed,
                              uint32_t* d_out) {
  return guarded_local([&] {
    if (!ctx || !d_rate || !background || !d_out) dio::fail(DIO_ERR_INVALID, "null argument");
    DIO_HIP(hipSetDevice(dio_ctx_device(ctx)));
    if (!npix || !n_frames) return;
    dio::DevBuf<double> bg(n_frames);
    DIO_HIP(hipMemcpy(bg.p, background, n_frames * 8, hipMemcpyHostToDevice));
    uint64_t n = npix * n_frames;
    k_poisson<<<dio::blocks_for(n, 256), 256>>>(d_rate, npix, bg.p, n_frames, seed, d_out);
    DIO_HIP(hipGetLastError());
    DIO_HIP(hipDeviceSynchronize());
  });
}

}  // extern "C"
