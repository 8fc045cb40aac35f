// Bitshuffle + LZ4 chunks (HDF5 filter 32008, bitshuffle 0.5.2 with its
// bundled LZ4 block format) decoded on the device and fed to the gather.
//
// Chunk layout (the reference's pre-scan, src/integration/cuda/device.rs:
// 2029-2180): 8-byte big-endian uncompressed size, 4-byte big-endian block
// size in bytes (0 = 8192), then per full block of block_bytes a 4-byte
// big-endian compressed length and an LZ4 block; the elements left over
// (fewer than a block) form one more compressed block of the largest
// multiple of 8 elements, and the last (< 8) elements follow verbatim.
// A decompressed block of n elements (n a multiple of 8) is bit-transposed:
// bit k of element i sits in byte k * n/8 + i/8, bit i % 8.
//
// Device design (wave64): one wavefront (and workgroup) per block, decoding
// in place: the compressed block is staged with coalesced dword loads at the
// end of an LDS region one block plus LZ4's in-place margin long, and
// decoded forward from the region's start, so a wave holds ~8.3 KB of LDS
// and about 19 blocks are in flight per CU to hide the serial LZ4 latency
// (twice as many as with separate input and output areas). Every lane
// parses the sequence stream in lockstep, the parse state kept scalar
// (readfirstlane), and the 64 lanes copy literals and matches. A match byte
// j equals out[op - offset + (j mod offset)], all of it written before the
// match starts, so a whole match is copied in one parallel pass whatever its
// offset. The bit transposition is done per lane, eight elements at a time,
// as 8x8 bit-matrix transposes of 64-bit words, stored straight to the
// frames. Every index is checked against the descriptor and buffer sizes,
// and the write position against the read position: a corrupt stream sets
// the error flag, never reads or writes outside its region and block.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "plan.hpp"

namespace {

using dio::DevBuf;
using dio::require;

constexpr uint32_t kRawFlag = 0x80000000u;
constexpr uint32_t kMaxBlockBytes = 49152;  // largest block accepted (device.rs:2076-2080)

uint32_t elem_bytes(int dtype) {
  switch (dtype) {
    case DIO_U16: return 2;
    case DIO_U32: case DIO_I32: return 4;
    default:
      dio::fail(DIO_ERR_INVALID, "Bitshuffle/LZ4 frames must be uint16, uint32 or int32");
  }
}

// LZ4's worst-case size of n incompressible bytes (what the reference
// accepts per block, device.rs:2084, 2121).
uint64_t lz4_bound(uint64_t n) { return n + n / 255 + 16; }

uint64_t be64(const uint8_t* p) {
  uint64_t v = 0;
  for (int i = 0; i < 8; i++) v = v << 8 | p[i];
  return v;
}
uint32_t be32(const uint8_t* p) {
  return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}
void put_be32(std::vector<uint8_t>& v, uint32_t x) {
  for (int s = 24; s >= 0; s -= 8) v.push_back((uint8_t)(x >> s));
}

// Transposes the 8x8 bit matrix held in x (row r = byte r, column c = bit c):
// bit 8r + c moves to bit 8c + r (three delta swaps).
__device__ inline uint64_t transpose8(uint64_t x) {
  uint64_t t;
  t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
  x ^= t ^ (t << 7);
  t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
  x ^= t ^ (t << 14);
  t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
  x ^= t ^ (t << 28);
  return x;
}

__device__ inline uint32_t sgpr(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// One block per wavefront; cap_o bytes of LDS per wave.
template <int EB>
__global__ __launch_bounds__(256) void k_bslz4(const uint8_t* __restrict__ comp, uint64_t comp_bytes,
                                                const dio_bslz4_block* __restrict__ desc,
                                                uint64_t nblocks, uint32_t cap_o,
                                                uint8_t* __restrict__ out, uint64_t n_elems,
                                                uint32_t* __restrict__ err) {
  extern __shared__ __align__(16) uint8_t lds[];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint64_t b = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wave;
  if (b >= nblocks) return;
  // One region of cap_o bytes per wave: the decoded block from its start,
  // the compressed block staged at its end (in-place decoding: the margin
  // cap_o - ub >= len / 256 + 32 keeps every write behind the read position
  // of a well-formed stream; a malformed one decodes wrongly but in bounds).
  uint8_t* o = lds + (size_t)wave * cap_o;
  const uint32_t src = sgpr(desc[b].src_off), len_raw = sgpr(desc[b].comp_len);
  const uint32_t dst = sgpr(desc[b].dst_off), ub = sgpr(desc[b].uncomp_bytes);
  auto fail = [&] {
    if (lane == 0) atomicOr(err, 1u);
  };
  if (len_raw & kRawFlag) {  // verbatim tail of fewer than 8 elements
    const uint32_t raw = len_raw & ~kRawFlag;
    if ((uint64_t)src + raw > comp_bytes || raw != ub || ub % EB ||
        (uint64_t)dst + ub / EB > n_elems) {
      fail();
      return;
    }
    for (uint32_t i = lane; i < raw; i += 64) out[(uint64_t)dst * EB + i] = comp[src + i];
    return;
  }
  const uint32_t len = len_raw;
  const uint32_t n = ub / EB;  // elements of the block, a multiple of 8
  const uint32_t L4 = (len + 7) & ~3u;
  if ((uint64_t)src + len > comp_bytes || ub % (8 * EB) || (uint64_t)dst + n > n_elems ||
      L4 + 4 > cap_o) {
    fail();
    return;
  }
  // Staged input: dword aligned, ending at the end of the region; byte k of
  // the block sits at region offset in_base + k.
  const uint32_t s0 = (cap_o - L4 - 4) & ~3u;
  uint8_t* c = o + s0;
  const uint32_t in_base = s0 + (src & 3u);
  if ((uint64_t)ub + len / 256 + 32 > (uint64_t)in_base + len) {  // in-place margin
    fail();
    return;
  }
  // Stage [src, src + len) in LDS from aligned dwords; byte k of the block
  // is c[head + k]. The last dwords are read byte by byte when they would
  // pass the end of the compressed buffer.
  const uint64_t a0 = src & ~3ull;
  const uint32_t head = src & 3u, nwords = (head + len + 3) / 4;
  const uint32_t* comp32 = reinterpret_cast<const uint32_t*>(comp + a0);
  uint32_t* c32 = reinterpret_cast<uint32_t*>(c);
  const uint32_t full = a0 + 4ull * nwords <= comp_bytes ? nwords : nwords - 1;
  for (uint32_t i0 = 0; i0 < full; i0 += 64 * 8) {
    uint32_t w[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const uint32_t i = i0 + lane + 64 * u;
      w[u] = comp32[i < full ? i : 0u];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const uint32_t i = i0 + lane + 64 * u;
      if (i < full) c32[i] = w[u];
    }
  }
  if (full < nwords && lane == 0) {
    const uint64_t a = a0 + 4ull * full;
    uint32_t w = 0;
    for (uint32_t k = 0; k < 4 && a + k < comp_bytes; k++) w |= (uint32_t)comp[a + k] << (8 * k);
    c32[full] = w;
  }
  __builtin_amdgcn_wave_barrier();
  const uint8_t* in = c + head;
  // LZ4 sequences (token, literal length, literals, 2-byte offset, match
  // length); the block ends after the literals of its last sequence.
  uint32_t ip = 0, op = 0;
  bool bad = false;
  while (ip < len) {
    const uint32_t token = sgpr(in[ip++]);
    uint32_t lit = token >> 4;
    if (lit == 15) {
      uint32_t s;
      do {
        if (ip >= len) { bad = true; break; }
        s = sgpr(in[ip++]);
        if (s > ub - op - lit) { bad = true; break; }
        lit += s;
      } while (s == 255);
    }
    if (bad || lit > len - ip || lit > ub - op || op > in_base + ip) { bad = true; break; }
    for (uint32_t k = lane; k < lit; k += 64) o[op + k] = in[ip + k];
    ip += lit;
    op += lit;
    if (ip >= len) break;
    if (ip + 2 > len) { bad = true; break; }
    const uint32_t off = sgpr((uint32_t)in[ip] | (uint32_t)in[ip + 1] << 8);
    ip += 2;
    if (off == 0 || off > op) { bad = true; break; }
    uint32_t ml = (token & 15u) + 4u;
    if ((token & 15u) == 15u) {
      uint32_t s;
      do {
        if (ip >= len) { bad = true; break; }
        s = sgpr(in[ip++]);
        if (ml > ub - op || s > ub - op - ml) { bad = true; break; }
        ml += s;
      } while (s == 255);
    }
    if (bad || ml > ub - op || op + ml > in_base + ip) { bad = true; break; }
    // Byte j of the match is o[op - off + j mod off]: all sources precede op.
    uint32_t r = off >= 64 ? lane : lane % off;
    const uint32_t step = off >= 64 ? (off == 64 ? 0u : 64u) : 64u % off;
    for (uint32_t j = lane; j < ml; j += 64) {
      o[op + j] = o[op - off + r];
      r += step;
      if (r >= off) r -= off;
    }
    op += ml;
  }
  if (bad || op != ub) {
    fail();
    return;
  }
  __builtin_amdgcn_wave_barrier();
  // Bit transposition: lane g handles elements 8g .. 8g + 7. For each group
  // q of 8 bit planes, byte p of x is byte g of plane 8q + p (bit e of it is
  // bit 8q + p of element 8g + e); after the transpose byte e of y holds
  // bits 8q .. 8q + 7 of element 8g + e.
  const uint32_t plane = n / 8;
  for (uint32_t g = lane; g < plane; g += 64) {
    uint64_t y[EB];
#pragma unroll
    for (int q = 0; q < EB; q++) {
      uint64_t x = 0;
#pragma unroll
      for (int p = 0; p < 8; p++) x |= (uint64_t)o[(8 * q + p) * plane + g] << (8 * p);
      y[q] = transpose8(x);
    }
#pragma unroll
    for (int e = 0; e < 8; e++) {
      uint32_t v = 0;
#pragma unroll
      for (int q = 0; q < EB; q++) v |= (uint32_t)(uint8_t)(y[q] >> (8 * e)) << (8 * q);
      if (EB == 4)
        reinterpret_cast<uint32_t*>(out)[(uint64_t)dst + 8 * g + e] = v;
      else
        reinterpret_cast<uint16_t*>(out)[(uint64_t)dst + 8 * g + e] = (uint16_t)v;
    }
  }
}

// Block descriptors of n chunks (validation as device.rs:2046-2180).
struct Scan {
  std::vector<dio_bslz4_block> blocks;
  uint32_t max_comp = 0, max_block = 0;
  uint64_t comp_bytes = 0;
};

void scan_chunks(const uint8_t* const* chunks, const uint64_t* sizes, uint64_t n_frames,
                 uint64_t pixels, int dtype, Scan& out) {
  const uint32_t eb = elem_bytes(dtype);
  const uint64_t expect = pixels * eb;
  require(n_frames * pixels < (1ull << 32), DIO_ERR_INVALID,
          "Bitshuffle/LZ4 frames of one call exceed 2^32 elements");
  uint64_t base = 0;
  auto bad = [](const std::string& m) { dio::fail(DIO_ERR_INVALID, m); };
  for (uint64_t f = 0; f < n_frames; f++) {
    const uint8_t* ch = chunks[f];
    const uint64_t sz = sizes[f];
    if (!ch) bad("null Bitshuffle/LZ4 chunk");
    if (sz < 12) bad("Bitshuffle/LZ4 chunk shorter than 12-byte header");
    const uint64_t uncomp = be64(ch);
    if (uncomp != expect)
      bad("Bitshuffle/LZ4 chunk uncompressed size " + std::to_string(uncomp) +
          " does not match plan frame size " + std::to_string(expect) + " bytes");
    uint32_t block = be32(ch + 8);
    if (block == 0) block = 8192;
    if (block > kMaxBlockBytes || block % (8 * eb))
      bad("unsupported Bitshuffle/LZ4 block size " + std::to_string(block) + " bytes");
    out.max_block = std::max(out.max_block, block);
    const uint64_t block_elems = block / eb, n_full = pixels / block_elems;
    const uint64_t rem = pixels % block_elems, rem_comp = rem / 8 * 8, tail = rem % 8;
    const uint64_t frame_dst = f * pixels;
    uint64_t off = 12;
    auto add_block = [&](uint64_t elems_at, uint64_t ub, const char* what) {
      if (off + 4 > sz) bad(std::string("truncated Bitshuffle/LZ4 ") + what + " header");
      const uint64_t cl = be32(ch + off);
      off += 4;
      if (cl > lz4_bound(ub))
        bad(std::string("Bitshuffle/LZ4 ") + what + " length " + std::to_string(cl) +
            " exceeds LZ4 bound " + std::to_string(lz4_bound(ub)));
      if (off + cl > sz) bad(std::string("truncated Bitshuffle/LZ4 ") + what + " payload");
      out.blocks.push_back({(uint32_t)(base + off), (uint32_t)cl, (uint32_t)(frame_dst + elems_at),
                            (uint32_t)ub});
      out.max_comp = std::max<uint32_t>(out.max_comp, (uint32_t)cl);
      off += cl;
    };
    for (uint64_t k = 0; k < n_full; k++) add_block(k * block_elems, block, "compressed block");
    if (rem_comp) add_block(n_full * block_elems, rem_comp * eb, "partial compressed block");
    if (tail) {
      const uint64_t raw = tail * eb;
      if (off + raw > sz) bad("truncated Bitshuffle/LZ4 tail bytes");
      out.blocks.push_back({(uint32_t)(base + off), kRawFlag | (uint32_t)raw,
                            (uint32_t)(frame_dst + n_full * block_elems + rem_comp),
                            (uint32_t)raw});
    }
    base += sz;
    if (base >= (1ull << 32) - 16) bad("compressed chunks of one call exceed 4 GiB");
  }
  out.comp_bytes = base;
}

void launch_decode(const uint8_t* d_comp, uint64_t comp_bytes, const dio_bslz4_block* d_desc,
                   uint64_t nblocks, uint32_t max_comp, uint32_t max_block, int dtype,
                   void* d_frames, uint64_t n_elems, uint32_t* d_err, hipStream_t s) {
  if (nblocks == 0) return;
  const uint32_t eb = elem_bytes(dtype);
  require(max_block <= kMaxBlockBytes && max_comp <= lz4_bound(kMaxBlockBytes), DIO_ERR_INVALID,
          "Bitshuffle/LZ4 block exceeds the decoder's limits");
  // Compressed staging holds the block and its 3-byte alignment head, and
  // later the decoded elements (max_block bytes).
  // In-place region: the largest block, the in-place margin of the largest
  // compressed block and its alignment.
  const uint32_t cap_o = (std::max(max_block + max_comp / 256 + 32, max_comp) + 16 + 15) / 16 * 16;
  const uint32_t per_wave = cap_o;
  const uint32_t waves = 1;  // one wave per workgroup measured fastest (LDS-bound occupancy)
  const uint64_t grid = (nblocks + waves - 1) / waves;
  require(grid < (1ull << 31), DIO_ERR_INVALID, "too many Bitshuffle/LZ4 blocks in one call");
  const size_t shmem = (size_t)waves * per_wave;
  auto* out = static_cast<uint8_t*>(d_frames);
  if (eb == 4)
    k_bslz4<4><<<(uint32_t)grid, 64 * waves, shmem, s>>>(d_comp, comp_bytes, d_desc, nblocks, cap_o,
                                                        out, n_elems, d_err);
  else
    k_bslz4<2><<<(uint32_t)grid, 64 * waves, shmem, s>>>(d_comp, comp_bytes, d_desc, nblocks, cap_o,
                                                        out, n_elems, d_err);
  DIO_HIP(hipGetLastError());
}

__global__ void k_zero_u32(uint32_t* p) { *p = 0; }

// ---------------------------------------------------------------- encoder
// Greedy LZ4 block compressor (hash of 4-byte sequences, the block format's
// end rules: the last 5 bytes are literals and no match starts in the last
// 12 bytes).
void lz4_block(const uint8_t* src, uint32_t n, std::vector<uint8_t>& dst) {
  constexpr uint32_t kMinMatch = 4, kLastLiterals = 5, kMfLimit = 12, kHashBits = 12;
  std::vector<int32_t> table(1u << kHashBits, -1);
  auto read32 = [&](uint32_t i) {
    uint32_t v;
    std::memcpy(&v, src + i, 4);
    return v;
  };
  auto hash = [&](uint32_t i) { return (read32(i) * 2654435761u) >> (32 - kHashBits); };
  auto put_len = [&](uint32_t v) {
    while (v >= 255) {
      dst.push_back(255);
      v -= 255;
    }
    dst.push_back((uint8_t)v);
  };
  uint32_t anchor = 0, i = 0;
  if (n >= kMfLimit + 1) {
    const uint32_t match_limit = n - kLastLiterals;  // matches end before here
    while (i + kMfLimit <= n) {
      const uint32_t h = hash(i);
      const int32_t cand = table[h];
      table[h] = (int32_t)i;
      if (cand < 0 || i - (uint32_t)cand > 65535 || read32((uint32_t)cand) != read32(i)) {
        i++;
        continue;
      }
      uint32_t m = (uint32_t)cand, len = kMinMatch;
      while (i + len < match_limit && src[m + len] == src[i + len]) len++;
      const uint32_t lit = i - anchor, ml = len - kMinMatch;
      dst.push_back((uint8_t)(std::min(lit, 15u) << 4 | std::min(ml, 15u)));
      if (lit >= 15) put_len(lit - 15);
      dst.insert(dst.end(), src + anchor, src + i);
      const uint32_t off = i - m;
      dst.push_back((uint8_t)off);
      dst.push_back((uint8_t)(off >> 8));
      if (ml >= 15) put_len(ml - 15);
      i += len;
      anchor = i;
    }
  }
  const uint32_t lit = n - anchor;
  dst.push_back((uint8_t)(std::min(lit, 15u) << 4));
  if (lit >= 15) put_len(lit - 15);
  dst.insert(dst.end(), src + anchor, src + n);
}

// Bit-transposes n elements (n a multiple of 8) of eb bytes.
void bitshuffle(const uint8_t* src, uint32_t n, uint32_t eb, uint8_t* dst) {
  const uint32_t plane = n / 8;
  std::memset(dst, 0, (size_t)n * eb);
  for (uint32_t i = 0; i < n; i++)
    for (uint32_t byte = 0; byte < eb; byte++) {
      const uint8_t v = src[(size_t)i * eb + byte];
      if (!v) continue;
      for (uint32_t bit = 0; bit < 8; bit++)
        if (v >> bit & 1) dst[(size_t)(8 * byte + bit) * plane + i / 8] |= (uint8_t)(1u << (i % 8));
    }
}

}  // namespace

extern "C" {

uint64_t dio_bslz4_bound(uint64_t n_elems, int dtype, uint32_t block_bytes) {
  uint32_t eb;
  switch (dtype) {
    case DIO_U16: eb = 2; break;
    case DIO_U32: case DIO_I32: eb = 4; break;
    default: return 0;
  }
  if (block_bytes == 0) block_bytes = 8192;
  const uint64_t bytes = n_elems * eb, blocks = bytes / block_bytes + 2;
  return 12 + blocks * 4 + lz4_bound(bytes) + blocks * 32;
}

int dio_bslz4_compress(const void* frame, uint64_t n_elems, int dtype, uint32_t block_bytes,
                       uint8_t* out, uint64_t cap, uint64_t* out_len) {
  return dio::guarded([&] {
    require(frame && out && out_len, DIO_ERR_INVALID, "null argument");
    const uint32_t eb = elem_bytes(dtype);
    const uint32_t block = block_bytes ? block_bytes : 8192;
    require(block <= kMaxBlockBytes && block % (8 * eb) == 0, DIO_ERR_INVALID,
            "block size must be a multiple of 8 elements and at most 49152 bytes");
    const uint8_t* src = static_cast<const uint8_t*>(frame);
    std::vector<uint8_t> v;
    v.reserve(n_elems * eb / 2 + 64);
    for (int s = 56; s >= 0; s -= 8) v.push_back((uint8_t)((n_elems * eb) >> s));
    put_be32(v, block);
    const uint64_t be = block / eb, n_full = n_elems / be, rem = n_elems % be;
    const uint64_t rem_comp = rem / 8 * 8, tail = rem % 8;
    std::vector<uint8_t> shuf(block), lz;
    auto emit = [&](const uint8_t* p, uint32_t n) {
      bitshuffle(p, n, eb, shuf.data());
      lz.clear();
      lz4_block(shuf.data(), n * eb, lz);
      put_be32(v, (uint32_t)lz.size());
      v.insert(v.end(), lz.begin(), lz.end());
    };
    for (uint64_t k = 0; k < n_full; k++) emit(src + k * block, (uint32_t)be);
    if (rem_comp) emit(src + n_full * block, (uint32_t)rem_comp);
    if (tail)
      v.insert(v.end(), src + (n_full * be + rem_comp) * eb, src + n_elems * eb);
    require(v.size() <= cap, DIO_ERR_INVALID, "compressed output exceeds the buffer");
    std::memcpy(out, v.data(), v.size());
    *out_len = v.size();
  });
}

int dio_bslz4_scan(const uint8_t* const* chunks, const uint64_t* sizes, uint64_t n_frames,
                   uint64_t pixels, int dtype, dio_bslz4_block* blocks, uint64_t cap,
                   uint64_t* n_blocks, uint32_t* max_comp, uint32_t* max_block) {
  return dio::guarded([&] {
    require(chunks && sizes && n_blocks, DIO_ERR_INVALID, "null argument");
    Scan sc;
    scan_chunks(chunks, sizes, n_frames, pixels, dtype, sc);
    *n_blocks = sc.blocks.size();
    if (max_comp) *max_comp = sc.max_comp;
    if (max_block) *max_block = sc.max_block;
    if (!blocks) return;  // count only
    require(cap >= sc.blocks.size(), DIO_ERR_INVALID, "block descriptor buffer too small");
    std::memcpy(blocks, sc.blocks.data(), sc.blocks.size() * sizeof(dio_bslz4_block));
  });
}

int dio_bslz4_decode_device(dio_ctx* ctx, const uint8_t* d_comp, uint64_t comp_bytes,
                            const dio_bslz4_block* d_blocks, uint64_t n_blocks, uint32_t max_comp,
                            uint32_t max_block, int dtype, void* d_frames, uint64_t n_elems,
                            uint32_t* d_err, void* stream) {
  return dio::guarded([&] {
    require(ctx && d_comp && d_blocks && d_frames && d_err, DIO_ERR_INVALID, "null argument");
    DIO_HIP(hipSetDevice(ctx->device));
    launch_decode(d_comp, comp_bytes, d_blocks, n_blocks, max_comp, max_block, dtype, d_frames,
                  n_elems, d_err, static_cast<hipStream_t>(stream));
  });
}

int dio_integrate_bslz4(dio_plan* p, const uint8_t* const* chunks, const uint64_t* sizes,
                        uint64_t n_frames, int dtype, uint32_t flags, double* intensity,
                        double* sigma) {
  return dio::guarded([&] {
    require(p != nullptr, DIO_ERR_INVALID, "null plan");
    if (n_frames == 0) return;
    require(chunks && sizes && intensity, DIO_ERR_INVALID, "null chunks or output");
    require(!(flags & DIO_WITH_ERRORS) || sigma, DIO_ERR_INVALID, "sigma output missing");
    require(!(flags & ~(uint32_t)(DIO_WITH_ERRORS | DIO_RAW_SUMS)), DIO_ERR_INVALID,
            "Bitshuffle/LZ4 integration takes DIO_WITH_ERRORS and DIO_RAW_SUMS only");
    const uint32_t eb = elem_bytes(dtype);
    const uint64_t npix = p->gt.npix;
    const uint32_t nb = p->gt.nbins;
    std::lock_guard<std::mutex> lk(p->mu);
    DIO_HIP(hipSetDevice(p->device));
    hipStream_t s = p->stream;
    // Groups of frames whose decoded frames stay within 1 GiB.
    const uint64_t group = std::min<uint64_t>(n_frames, std::max<uint64_t>(1, (1ull << 30) / (npix * eb)));
    DevBuf<uint8_t> frames(group * npix * eb);
    DevBuf<uint32_t> err(1);
    if (p->out_i.n < group * nb) p->out_i.alloc(group * nb);
    if ((flags & DIO_WITH_ERRORS) && p->out_s.n < group * nb) p->out_s.alloc(group * nb);
    for (uint64_t f0 = 0; f0 < n_frames; f0 += group) {
      const uint64_t nf = std::min(group, n_frames - f0);
      Scan sc;
      scan_chunks(chunks + f0, sizes + f0, nf, npix, dtype, sc);
      // Compressed bytes of the group, concatenated: one upload.
      std::vector<uint8_t> host(sc.comp_bytes);
      uint64_t at = 0;
      for (uint64_t k = 0; k < nf; k++) {
        std::memcpy(host.data() + at, chunks[f0 + k], sizes[f0 + k]);
        at += sizes[f0 + k];
      }
      DevBuf<uint8_t> dcomp(std::max<uint64_t>(sc.comp_bytes, 1));
      DevBuf<dio_bslz4_block> ddesc(std::max<size_t>(sc.blocks.size(), 1));
      DIO_HIP(hipMemcpyAsync(dcomp.p, host.data(), sc.comp_bytes, hipMemcpyHostToDevice, s));
      DIO_HIP(hipMemcpyAsync(ddesc.p, sc.blocks.data(), sc.blocks.size() * sizeof(dio_bslz4_block),
                             hipMemcpyHostToDevice, s));
      k_zero_u32<<<1, 1, 0, s>>>(err.p);
      launch_decode(dcomp.p, sc.comp_bytes, ddesc.p, sc.blocks.size(), sc.max_comp, sc.max_block,
                    dtype, frames.p, nf * npix, err.p, s);
      uint32_t herr = 0;
      DIO_HIP(hipMemcpyAsync(&herr, err.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      DIO_HIP(hipStreamSynchronize(s));
      require(herr == 0, DIO_ERR_INVALID, "corrupt or truncated Bitshuffle/LZ4 compressed block");
      dio::run_device(p, p->sc, frames.p, dtype, nf, flags, p->out_i.p, p->out_s.p, nullptr, s);
      DIO_HIP(hipMemcpyAsync(intensity + f0 * nb, p->out_i.p, nf * nb * 8, hipMemcpyDeviceToHost, s));
      if (flags & DIO_WITH_ERRORS)
        DIO_HIP(hipMemcpyAsync(sigma + f0 * nb, p->out_s.p, nf * nb * 8, hipMemcpyDeviceToHost, s));
      DIO_HIP(hipStreamSynchronize(s));
    }
  });
}

}  // extern "C"
