// Per-frame-mask kernels (gfx950): the CSR route (one wavefront per bin row
// of the unmasked table) and the staging and finishing passes of the tile
// route. The path is a sparse, bandwidth-bound reduction, so it uses the
// vector ALUs, not MFMA.
#include <algorithm>
#include <cfloat>

#include "gather.hpp"

namespace dio {

namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;

// --------------------------------------------------------- per-frame masks

__device__ inline double wave_sum(double x) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) x = x + __shfl_down(x, off, kWave);
  return __shfl(x, 0, kWave);
}

// prepared.rs:1011-1121 (row, strict order per lane) with a per-frame mask:
// one wavefront per row and NF frames; entry k's index, weight and
// normalization are read once for the NF frames, and its NF mask bytes and
// signals are loaded together before the sums. FP64 sums of FP64 weights;
// lane partial sums combined by the fixed tree of wave_sum.
template <class P, int NF>
__global__ __launch_bounds__(kBlock) void k_masked(const uint32_t* __restrict__ row_ptr,
                                                   const uint32_t* __restrict__ col,
                                                   const double* __restrict__ w64, int weighted,
                                                   const double* __restrict__ norm,
                                                   const P* __restrict__ frames,
                                                   const uint8_t* __restrict__ masks,
                                                   uint64_t npix, uint64_t n_frames,
                                                   uint32_t nbins, uint32_t dim1, uint32_t dim2,
                                                   int flip, int errors, double* I, double* S) {
  uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  uint32_t r = blockIdx.x * kWavesPerBlock + wave;
  if (r >= nbins) return;
  const uint64_t f0 = (uint64_t)blockIdx.y * NF;
  const int nf = (int)min<uint64_t>(NF, n_frames - f0);
  const P* fr = frames + f0 * npix;
  const uint8_t* mk = masks + f0 * npix;
  const uint32_t k0 = row_ptr[r], k1 = row_ptr[r + 1];
  double sx[NF], sn[NF], sv[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) sx[f] = sn[f] = sv[f] = 0.0;
  for (uint32_t k = k0 + lane; k < k1; k += kWave) {
    const uint32_t j = col[k];
    const double w = weighted ? w64[k] : 1.0;
    const double n = norm[j];
    // Frames and masks hold the rows in input order (detector order flipped
    // for bottom-up input).
    const uint64_t jf = flip ? (uint64_t)(dim1 - 1 - j / dim2) * dim2 + j % dim2 : j;
    // All loads of the entry first, then the sums: a masked entry is
    // skipped (its signal is never read into a sum, NaN or not).
    bool on[NF];
    double xs[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) {
      on[f] = f < nf && mk[(uint64_t)f * npix + jf] == 0;
      xs[f] = f < nf ? (double)fr[(uint64_t)f * npix + jf] : 0.0;
    }
#pragma unroll
    for (int f = 0; f < NF; f++) {
      if (!on[f]) continue;
      sx[f] = sx[f] + w * xs[f];
      sn[f] = sn[f] + w * n;
      if (errors) sv[f] = sv[f] + w * w * fmax(xs[f], 0.0);
    }
  }
#pragma unroll
  for (int f = 0; f < NF; f++) {
    sx[f] = wave_sum(sx[f]);
    sn[f] = wave_sum(sn[f]);
    if (errors) sv[f] = wave_sum(sv[f]);
    if (f >= nf || lane != 0) continue;
    const uint64_t ob = (f0 + f) * nbins + r;
    const double d = sn[f];
    I[ob] = d == 0.0 ? 0.0 : sx[f] / d;
    if (errors) S[ob] = d == 0.0 ? 0.0 : sqrt(fmax(sv[f] / (d * d), 0.0));
  }
}

// One frame per blockIdx.y: masked pixels' signal and normalization set to 0
// (a masked pixel then adds exactly 0 to every sum of the tile kernel).
template <class T, class N>
__global__ __launch_bounds__(256) void k_masked_prepare(const T* __restrict__ frames,
                                                        const uint8_t* __restrict__ masks,
                                                        const double* __restrict__ norm,
                                                        uint64_t npix, uint32_t dim1,
                                                        uint32_t dim2, int flip, T* xo, N* no) {
  const uint64_t base = (uint64_t)blockIdx.y * npix;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const bool m = masks[base + p] != 0;
    const uint64_t pd = flip ? (uint64_t)(dim1 - 1 - p / dim2) * dim2 + p % dim2 : p;
    xo[base + p] = m ? T(0) : frames[base + p];
    no[base + p] = m ? N(0) : (N)norm[pd];
  }
}

__global__ __launch_bounds__(256) void k_masked_finish(double* I, double* S, const double* D,
                                                       uint64_t n, int errors) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const double d = D[i];
    I[i] = d == 0.0 ? 0.0 : I[i] / d;
    if (errors) S[i] = d == 0.0 ? 0.0 : sqrt(fmax(S[i] / (d * d), 0.0));
  }
}

// One byte per pixel and 8-frame pass of the masks of n frames: bit i of
// byte [h][p] (plane h at h * stride) is frame 8 h + i's mask at pixel p
// (input order). A thread packs 16 consecutive pixels: per frame it reads
// the 5 aligned words that cover them and takes the 16 bytes out with byte
// alignment; a group whose words would reach outside the masks, and a
// frame's last partial group, are read byte by byte.
__global__ __launch_bounds__(256) void k_mask_bits(const uint8_t* __restrict__ masks,
                                                   uint64_t npix, uint32_t n, uint64_t stride,
                                                   uint8_t* out) {
  const uint32_t h = blockIdx.y;
  const uint32_t f0 = 8 * h, nf = min(8u, n - f0);
  const uint64_t total = (uint64_t)n * npix, groups = (npix + 15) / 16;
  const uintptr_t base = reinterpret_cast<uintptr_t>(masks);
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < groups;
       j += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t p0 = 16 * j;
    uint32_t bits[16];
#pragma unroll
    for (int k = 0; k < 16; k++) bits[k] = 0;
    for (uint32_t i = 0; i < nf; i++) {
      const uint64_t b = (uint64_t)(f0 + i) * npix + p0;
      const uintptr_t aw = (base + b) & ~(uintptr_t)3;  // the aligned word holding byte b
      if (p0 + 16 <= npix && aw >= base && aw + 20 <= base + total) {
        const uint32_t* w = reinterpret_cast<const uint32_t*>(aw);
        const uint32_t sh = (uint32_t)((base + b) & 3) * 8;
        uint32_t v[5];
#pragma unroll
        for (int k = 0; k < 5; k++) v[k] = w[k];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const uint32_t x = sh ? (v[k] >> sh) | (v[k + 1] << (32 - sh)) : v[k];
#pragma unroll
          for (int c = 0; c < 4; c++) bits[4 * k + c] |= ((x >> (8 * c)) & 0xffu ? 1u : 0u) << i;
        }
      } else {
        for (uint32_t k = 0; k < 16 && p0 + k < npix; k++)
          bits[k] |= (masks[b + k] != 0 ? 1u : 0u) << i;
      }
    }
    uint8_t* o = out + (uint64_t)h * stride + p0;
    if (p0 + 16 <= stride) {  // planes are padded to 16 bytes: whole 16-byte stores
      uint4 q;
      q.x = bits[0] | bits[1] << 8 | bits[2] << 16 | bits[3] << 24;
      q.y = bits[4] | bits[5] << 8 | bits[6] << 16 | bits[7] << 24;
      q.z = bits[8] | bits[9] << 8 | bits[10] << 16 | bits[11] << 24;
      q.w = bits[12] | bits[13] << 8 | bits[14] << 16 | bits[15] << 24;
      *reinterpret_cast<uint4*>(o) = q;
    }
  }
}

__global__ __launch_bounds__(256) void k_norm32(const double* __restrict__ norm, uint64_t n,
                                                float* out) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (uint64_t)gridDim.x * blockDim.x)
    out[p] = (float)norm[p];
}

// Frame k of masked-row passes (pass k / 8, row k % 8 holds its sums, row
// 8 + k % 8 its denominator): the arithmetic of k_masked_finish.
__global__ __launch_bounds__(256) void k_masked_rows_finish(const double* R, const double* V,
                                                            uint64_t n_frames, uint32_t nbins,
                                                            double* I, double* S) {
  const uint64_t n = n_frames * nbins;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = j / nbins, b = j % nbins, h = k / 8, r = k % 8;
    const double x = R[(16 * h + r) * nbins + b], d = R[(16 * h + 8 + r) * nbins + b];
    I[j] = d == 0.0 ? 0.0 : x / d;
    if (S) S[j] = d == 0.0 ? 0.0 : sqrt(fmax(V[(16 * h + r) * nbins + b] / (d * d), 0.0));
  }
}

}  // namespace

uint64_t mask_bits_stride(uint64_t npix) { return (npix + 15) / 16 * 16; }

void mask_bits(const uint8_t* masks, uint64_t npix, uint32_t n_frames, uint8_t* out,
               hipStream_t s) {
  if (n_frames == 0) return;
  const uint64_t groups = (npix + 15) / 16;
  const dim3 grid((uint32_t)std::min<uint64_t>((groups + 255) / 256, 2048), (n_frames + 7) / 8);
  k_mask_bits<<<grid, 256, 0, s>>>(masks, npix, n_frames, mask_bits_stride(npix), out);
  DIO_HIP(hipGetLastError());
}

void norm_to_f32(const double* norm, uint64_t n, float* out, hipStream_t s) {
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
  k_norm32<<<blocks, 256, 0, s>>>(norm, n, out);
  DIO_HIP(hipGetLastError());
}

void masked_rows_finish(const double* rows_sum, const double* rows_var, uint64_t n_frames,
                        uint32_t nbins, double* intensity, double* sigma, hipStream_t s) {
  const uint64_t n = n_frames * nbins;
  if (n == 0) return;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
  k_masked_rows_finish<<<blocks, 256, 0, s>>>(rows_sum, rows_var, n_frames, nbins, intensity,
                                              sigma);
  DIO_HIP(hipGetLastError());
}

void masked_frames(const StdTable& t, const void* frames, int dtype, const uint8_t* masks,
                   uint64_t n_frames, bool errors, double* intensity, double* sigma,
                   hipStream_t s) {
  if (n_frames == 0) return;
  constexpr int NF = 8;  // frames per wavefront
  dim3 grid((t.nbins + kWavesPerBlock - 1) / kWavesPerBlock, (uint32_t)((n_frames + NF - 1) / NF));
  switch (dtype) {
#define DIO_MASKED(T)                                                                         \
  k_masked<T, NF><<<grid, kBlock, 0, s>>>(t.row_ptr, t.col, t.w64, t.weighted, t.norm,       \
                                          static_cast<const T*>(frames), masks, t.npix,        \
                                          n_frames, t.nbins, t.dim1, t.dim2, t.flip ? 1 : 0,   \
                                          errors, intensity, sigma);                           \
  break;
    case DIO_U16: DIO_MASKED(uint16_t)
    case DIO_U32: DIO_MASKED(uint32_t)
    case DIO_I32: DIO_MASKED(int32_t)
    case DIO_F32T: DIO_MASKED(float)
    case DIO_F64T: DIO_MASKED(double)
#undef DIO_MASKED
    default: fail(DIO_ERR_INVALID, "unknown frame dtype");
  }
  DIO_HIP(hipGetLastError());
}

void masked_prepare(const StdTable& t, const void* frames, int dtype, const uint8_t* masks,
                    uint64_t n_frames, bool norm_f64, void* xo, void* no, hipStream_t s) {
  if (n_frames == 0) return;
  if (n_frames > 65535) fail(DIO_ERR_INVALID, "at most 65535 frames per prepare");
  dim3 grid((uint32_t)std::min<uint64_t>((t.npix + 255) / 256, 1024), (uint32_t)n_frames);
  const int flip = t.flip ? 1 : 0;
  switch (dtype) {
#define DIO_PREP(T)                                                                          \
  if (norm_f64)                                                                              \
    k_masked_prepare<T, double><<<grid, 256, 0, s>>>(static_cast<const T*>(frames), masks,   \
                                                     t.norm, t.npix, t.dim1, t.dim2, flip,    \
                                                     static_cast<T*>(xo),                     \
                                                     static_cast<double*>(no));               \
  else                                                                                       \
    k_masked_prepare<T, float><<<grid, 256, 0, s>>>(static_cast<const T*>(frames), masks,    \
                                                    t.norm, t.npix, t.dim1, t.dim2, flip,     \
                                                    static_cast<T*>(xo),                      \
                                                    static_cast<float*>(no));                 \
  break;
    case DIO_U16: DIO_PREP(uint16_t)
    case DIO_U32: DIO_PREP(uint32_t)
    case DIO_I32: DIO_PREP(int32_t)
    case DIO_F32T: DIO_PREP(float)
    case DIO_F64T: DIO_PREP(double)
#undef DIO_PREP
    default: fail(DIO_ERR_INVALID, "unknown frame dtype");
  }
  DIO_HIP(hipGetLastError());
}

void masked_finish(double* intensity, double* sigma, const double* den, uint64_t n, bool errors,
                   hipStream_t s) {
  if (n == 0) return;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
  k_masked_finish<<<blocks, 256, 0, s>>>(intensity, sigma, den, n, errors ? 1 : 0);
  DIO_HIP(hipGetLastError());
}

}  // namespace dio
