/* CPU oracle of Bitshuffle + LZ4 chunks (TEST INFRASTRUCTURE ONLY: tests
 * link it to check the device decoder and the product's encoder; nothing in
 * the product calls it).
 *
 * The codec is a third-party dependency of the reference's data path, not
 * part of /root/reference: the HDF5 Bitshuffle filter (id 32008) of
 * bitshuffle 0.5.2 (pinned in the reference's uv.lock), which bundles the
 * LZ4 block format. Restated here from their published definitions:
 *   chunk   8-byte big-endian decoded size, 4-byte big-endian block size in
 *           bytes (0 = 8192); per full block a 4-byte big-endian compressed
 *           length and an LZ4 block; the leftover elements' largest multiple
 *           of 8 as one more block; the last (< 8) elements verbatim (the
 *           reference's reading of it: src/integration/cuda/device.rs:
 *           2046-2180);
 *   LZ4     sequences of token (high nibble literal count, low nibble match
 *           length - 4, 15 meaning continuation bytes of 255 until one below
 *           255), literals, 2-byte little-endian offset, continuation of the
 *           match length; the last sequence stops after its literals; a match
 *           may overlap its own output (offset < length);
 *   shuffle a block of n elements (n a multiple of 8) stores bit k of element
 *           i in byte k * n/8 + i/8, bit i % 8 (bit k counted from the least
 *           significant bit of the little-endian element).
 * The decoder is written for clarity (byte at a time, every bound checked).
 * Two encoders restate the reference's own test encoders: literal-only LZ4
 * (src/integration/metal/tests.rs:297-357) and byte-run LZ4 with offset-1
 * matches (src/integration/metal/bench.rs:150-230).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static uint64_t rd_be64(const uint8_t *p) {
  uint64_t v = 0;
  for (int i = 0; i < 8; i++) v = (v << 8) | p[i];
  return v;
}
static uint32_t rd_be32(const uint8_t *p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

/* One LZ4 block: returns the decoded byte count, or -1 when the stream is
 * malformed or does not fit out_cap. */
long or_lz4_decode_block(const uint8_t *in, size_t in_len, uint8_t *out, size_t out_cap) {
  size_t ip = 0, op = 0;
  while (ip < in_len) {
    unsigned token = in[ip++];
    size_t lit = token >> 4;
    if (lit == 15) {
      unsigned s;
      do {
        if (ip >= in_len) return -1;
        s = in[ip++];
        lit += s;
      } while (s == 255);
    }
    if (lit > in_len - ip || lit > out_cap - op) return -1;
    for (size_t k = 0; k < lit; k++) out[op + k] = in[ip + k];
    ip += lit;
    op += lit;
    if (ip == in_len) break; /* last sequence: literals only */
    if (in_len - ip < 2) return -1;
    size_t off = (size_t)in[ip] | ((size_t)in[ip + 1] << 8);
    ip += 2;
    if (off == 0 || off > op) return -1;
    size_t ml = (token & 15u) + 4;
    if ((token & 15u) == 15u) {
      unsigned s;
      do {
        if (ip >= in_len) return -1;
        s = in[ip++];
        ml += s;
      } while (s == 255);
    }
    if (ml > out_cap - op) return -1;
    for (size_t k = 0; k < ml; k++) out[op + k] = out[op + k - off]; /* in order: overlap */
    op += ml;
  }
  return (long)op;
}

/* Inverse bit transposition of one block of n elements of eb bytes. */
void or_bitunshuffle(const uint8_t *src, size_t n, size_t eb, uint8_t *dst) {
  size_t plane = n / 8;
  memset(dst, 0, n * eb);
  for (size_t k = 0; k < 8 * eb; k++)
    for (size_t i = 0; i < n; i++)
      if ((src[k * plane + i / 8] >> (i % 8)) & 1u) dst[i * eb + k / 8] |= (uint8_t)(1u << (k % 8));
}

/* Forward transposition (the definition read the other way). */
void or_bitshuffle(const uint8_t *src, size_t n, size_t eb, uint8_t *dst) {
  size_t plane = n / 8;
  memset(dst, 0, n * eb);
  for (size_t k = 0; k < 8 * eb; k++)
    for (size_t i = 0; i < n; i++)
      if ((src[i * eb + k / 8] >> (k % 8)) & 1u) dst[k * plane + i / 8] |= (uint8_t)(1u << (i % 8));
}

/* Whole chunk into out (pixels elements of eb bytes). 0 on success; 1 bad
 * header or sizes; 2 truncated; 3 malformed LZ4 block. */
int or_bslz4_decode(const uint8_t *ch, size_t len, size_t pixels, size_t eb, uint8_t *out) {
  if (len < 12) return 1;
  if (rd_be64(ch) != (uint64_t)pixels * eb) return 1;
  size_t block = rd_be32(ch + 8);
  if (block == 0) block = 8192;
  if (block % (8 * eb)) return 1;
  size_t be = block / eb, n_full = pixels / be, rem = pixels % be;
  size_t rem_comp = rem / 8 * 8, tail = rem % 8, off = 12;
  uint8_t *tmp = malloc(block);
  int rc = 0;
  for (size_t b = 0; b <= n_full && rc == 0; b++) {
    size_t n = b < n_full ? be : rem_comp;
    if (n == 0) break;
    if (len - off < 4) { rc = 2; break; }
    size_t cl = rd_be32(ch + off);
    off += 4;
    if (cl > len - off) { rc = 2; break; }
    long got = or_lz4_decode_block(ch + off, cl, tmp, n * eb);
    if (got != (long)(n * eb)) { rc = 3; break; }
    or_bitunshuffle(tmp, n, eb, out + b * be * eb);
    off += cl;
  }
  if (rc == 0 && tail) {
    if (len - off < tail * eb) rc = 2;
    else memcpy(out + (n_full * be + rem_comp) * eb, ch + off, tail * eb);
  }
  free(tmp);
  return rc;
}

/* ------------------------------------------------------------ encoders */

static size_t put_len(uint8_t *dst, size_t v) {
  size_t k = 0;
  while (v >= 255) { dst[k++] = 255; v -= 255; }
  dst[k++] = (uint8_t)v;
  return k;
}

/* Literal-only LZ4 block of n bytes; returns its length (dst holds at
 * least n + n / 255 + 16 bytes). */
static size_t lz4_literals(const uint8_t *src, size_t n, uint8_t *dst) {
  size_t k = 0;
  dst[k++] = (uint8_t)((n < 15 ? n : 15) << 4);
  if (n >= 15) k += put_len(dst + k, n - 15);
  memcpy(dst + k, src, n);
  return k + n;
}

/* Byte runs of 6 or more (not reaching the end) become an offset-1 match
 * after their first byte; everything else is literal. */
static size_t lz4_runs(const uint8_t *src, size_t n, uint8_t *dst) {
  size_t k = 0, i = 0, anchor = 0;
  while (i < n) {
    size_t run = 1;
    while (i + run < n && src[i + run] == src[i]) run++;
    if (run >= 6 && i + run < n) {
      size_t lit = i + 1 - anchor, ml = run - 1;
      dst[k++] = (uint8_t)(((lit < 15 ? lit : 15) << 4) | (ml - 4 < 15 ? ml - 4 : 15));
      if (lit >= 15) k += put_len(dst + k, lit - 15);
      memcpy(dst + k, src + anchor, lit);
      k += lit;
      dst[k++] = 1;
      dst[k++] = 0;
      if (ml - 4 >= 15) k += put_len(dst + k, ml - 4 - 15);
      i += run;
      anchor = i;
    } else {
      i += run;
    }
  }
  size_t lit = n - anchor;
  dst[k++] = (uint8_t)((lit < 15 ? lit : 15) << 4);
  if (lit >= 15) k += put_len(dst + k, lit - 15);
  memcpy(dst + k, src + anchor, lit);
  return k + lit;
}

static void put_be32(uint8_t *p, uint32_t x) {
  p[0] = (uint8_t)(x >> 24); p[1] = (uint8_t)(x >> 16); p[2] = (uint8_t)(x >> 8); p[3] = (uint8_t)x;
}

/* A chunk of pixels elements with blocks of block_bytes; mode 0 literal
 * LZ4, mode 1 byte runs. Returns its length, 0 if out_cap is too small. */
size_t or_bslz4_encode(const uint8_t *raw, size_t pixels, size_t eb, size_t block_bytes, int mode,
                       uint8_t *out, size_t out_cap) {
  size_t be = block_bytes / eb, n_full = pixels / be, rem = pixels % be;
  size_t rem_comp = rem / 8 * 8, tail = rem % 8;
  size_t need = 12 + (n_full + 2) * (4 + block_bytes + block_bytes / 255 + 16) + tail * eb;
  if (out_cap < need) return 0;
  uint64_t total = (uint64_t)pixels * eb;
  for (int i = 0; i < 8; i++) out[i] = (uint8_t)(total >> (56 - 8 * i));
  put_be32(out + 8, (uint32_t)block_bytes);
  size_t k = 12;
  uint8_t *shuf = malloc(block_bytes);
  for (size_t b = 0; b <= n_full; b++) {
    size_t n = b < n_full ? be : rem_comp;
    if (n == 0) break;
    or_bitshuffle(raw + b * be * eb, n, eb, shuf);
    size_t cl = mode ? lz4_runs(shuf, n * eb, out + k + 4) : lz4_literals(shuf, n * eb, out + k + 4);
    put_be32(out + k, (uint32_t)cl);
    k += 4 + cl;
  }
  free(shuf);
  memcpy(out + k, raw + (n_full * be + rem_comp) * eb, tail * eb);
  return k + tail * eb;
}
