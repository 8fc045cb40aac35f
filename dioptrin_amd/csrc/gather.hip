// Azimuthal-spread kernel (gfx950): one wavefront per bin row of the
// unmasked CSR table, FP64 arithmetic on float32 weights. The path is a
// sparse, bandwidth-bound reduction, so it uses the vector ALUs, not MFMA.
#include <cfloat>

#include "gather.hpp"

namespace dio {

namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;

// --------------------------------------------------------- azimuthal std

__device__ inline double wave_sum(double x) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) x = x + __shfl_down(x, off, kWave);
  return __shfl(x, 0, kWave);
}

// azimuthal.rs:48-118: weighted population spread of corrected pixels on the
// unmasked table with the mask applied per entry. One wavefront per row;
// FP64 arithmetic on float32 weights.
template <class P>
__global__ __launch_bounds__(kBlock) void k_std(const uint32_t* __restrict__ row_ptr,
                                                const uint32_t* __restrict__ col,
                                                const float* __restrict__ w32,
                                                const uint8_t* __restrict__ mask,
                                                const double* __restrict__ norm,
                                                const P* __restrict__ frames, uint64_t npix,
                                                uint32_t nbins, uint32_t dim1, uint32_t dim2,
                                                int flip, int errors, double* I, double* S,
                                                double* SD) {
  uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  uint32_t r = blockIdx.x * kWavesPerBlock + wave;
  if (r >= nbins) return;
  const P* fr = frames + (uint64_t)blockIdx.y * npix;
  uint64_t ob = (uint64_t)blockIdx.y * nbins + r;
  uint32_t k0 = row_ptr[r], k1 = row_ptr[r + 1];
  auto contrib = [&](uint32_t k, double& c, double& s, double& x, double& w) -> bool {
    uint32_t j = col[k];
    if (mask && mask[j]) return false;
    double n = norm[j];
    c = (double)w32[k];
    s = (double)fr[flip ? (dim1 - 1 - j / dim2) * dim2 + j % dim2 : j];
    x = s / n;
    w = c * n;
    return isfinite(s) && isfinite(n) && n > 0.0 && isfinite(c) && c > 0.0 && isfinite(x) &&
           isfinite(w) && w > 0.0;
  };
  // Origin: the first valid contribution in row order.
  bool found = false;
  double origin = 0.0;
  for (uint32_t base = k0; base < k1 && !found; base += kWave) {
    uint32_t k = base + lane;
    double c, s, x, w;
    bool ok = k < k1 && contrib(k, c, s, x, w);
    unsigned long long m = __ballot(ok);
    if (m) {
      int first = __ffsll((long long)m) - 1;
      origin = __shfl(ok ? x : 0.0, first, kWave);
      found = true;
    }
  }
  if (!found) {
    if (lane == 0) {
      I[ob] = 0.0;
      if (errors) S[ob] = 0.0;
      SD[ob] = NAN;
    }
    return;
  }
  double ss = 0.0, sw = 0.0, sh = 0.0, sv = 0.0;
  for (uint32_t k = k0 + lane; k < k1; k += kWave) {
    double c, s, x, w;
    if (!contrib(k, c, s, x, w)) continue;
    ss = ss + c * s;
    sw = sw + w;
    sh = sh + w * (x - origin);
    sv = sv + c * c * fmax(s, 0.0);
  }
  ss = wave_sum(ss);
  sw = wave_sum(sw);
  sh = wave_sum(sh);
  sv = wave_sum(sv);
  double mean_shifted = sh / sw;
  double m2 = 0.0;
  for (uint32_t k = k0 + lane; k < k1; k += kWave) {
    double c, s, x, w;
    if (!contrib(k, c, s, x, w)) continue;
    double delta = (x - origin) - mean_shifted;
    m2 = m2 + w * delta * delta;
  }
  m2 = wave_sum(m2);
  if (lane == 0) {
    I[ob] = ss / sw;
    if (errors) S[ob] = sqrt(fmax(sv, 0.0)) / sw;
    SD[ob] = sqrt(fmax(m2 / sw, 0.0));
  }
}

}  // namespace

void azimuthal_std_frames(const StdTable& t, const void* frames, int dtype, uint64_t n_frames,
                          bool errors, double* intensity, double* sigma, double* std_out,
                          hipStream_t s) {
  if (n_frames == 0) return;
  dim3 grid((t.nbins + kWavesPerBlock - 1) / kWavesPerBlock, (uint32_t)n_frames);
  switch (dtype) {
#define DIO_STD(T)                                                                            \
  k_std<T><<<grid, kBlock, 0, s>>>(t.row_ptr, t.col, t.w32, t.mask, t.norm,                  \
                                   static_cast<const T*>(frames), t.npix, t.nbins, t.dim1,     \
                                   t.dim2, t.flip ? 1 : 0, errors,                             \
                                   intensity, sigma, std_out);                                 \
  break;
    case DIO_U16: DIO_STD(uint16_t)
    case DIO_U32: DIO_STD(uint32_t)
    case DIO_I32: DIO_STD(int32_t)
    case DIO_F32T: DIO_STD(float)
    case DIO_F64T: DIO_STD(double)
#undef DIO_STD
    default: fail(DIO_ERR_INVALID, "unknown frame dtype");
  }
  DIO_HIP(hipGetLastError());
}

}  // namespace dio
