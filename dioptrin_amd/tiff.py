"""Single-channel TIFF frames for the integration path (Frame::from_tiff,
src/frame.rs:32-190 of the reference).

Reads baseline and BigTIFF files whose image is stored in strips: unsigned,
signed and floating-point samples of 8 to 64 bits, uncompressed, PackBits,
LZW or Deflate, with or without the horizontal predictor. Like the
reference, only single-channel (grayscale) images are accepted, and the
rows are flipped up-down while loading (convert_and_flip_into,
frame.rs:186-214), so row 0 of the returned array is the file's last row.

The array keeps the file's sample type; the integration converts it (the
reference converts every sample to f64 first, which is exact for all types
up to 32 bits).
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

SINGLE_CHANNEL_ONLY = "only single-channel (grayscale) TIFFs are supported"

_TYPES = {1: ("B", 1), 2: ("c", 1), 3: ("H", 2), 4: ("I", 4), 5: ("II", 8), 6: ("b", 1),
          7: ("c", 1), 8: ("h", 2), 9: ("i", 4), 10: ("ii", 8), 11: ("f", 4), 12: ("d", 8),
          16: ("Q", 8), 17: ("q", 8), 18: ("Q", 8)}

WIDTH, HEIGHT, BITS, COMPRESSION, PHOTOMETRIC = 256, 257, 258, 259, 262
STRIP_OFFSETS, SAMPLES, ROWS_PER_STRIP, STRIP_COUNTS = 273, 277, 278, 279
PREDICTOR, SAMPLE_FORMAT, TILE_WIDTH = 317, 339, 322


class TiffError(ValueError):
    """A file that is not a TIFF this reader supports."""


def _ifd(buf: bytes, bo: str, big: bool, off: int) -> dict:
    n_fmt, n_sz = ("Q", 8) if big else ("H", 2)
    ent_sz, val_sz = (20, 8) if big else (12, 4)
    if off + n_sz > len(buf):
        raise TiffError("IFD offset beyond the end of the file")
    (n,) = struct.unpack_from(bo + n_fmt, buf, off)
    tags = {}
    for i in range(n):
        e = off + n_sz + i * ent_sz
        if big:
            tag, typ, cnt = struct.unpack_from(bo + "HHQ", buf, e)
        else:
            tag, typ, cnt = struct.unpack_from(bo + "HHI", buf, e)
        if typ not in _TYPES:
            continue
        fmt, size = _TYPES[typ]
        total = size * cnt
        if total <= val_sz:
            data_off = e + 4 + (8 if big else 4)
        else:
            (data_off,) = struct.unpack_from(bo + ("Q" if big else "I"), buf, e + 4 +
                                             (8 if big else 4))
        if data_off + total > len(buf):
            raise TiffError(f"tag {tag} data beyond the end of the file")
        if typ in (2, 7):
            continue
        vals = struct.unpack_from(bo + fmt * cnt, buf, data_off)
        if typ in (5, 10):
            vals = tuple(vals[2 * k] / vals[2 * k + 1] if vals[2 * k + 1] else 0.0
                         for k in range(cnt))
        tags[tag] = vals
    return tags


def _packbits(data: bytes) -> bytes:
    out = bytearray()
    i, n = 0, len(data)
    while i < n:
        h = data[i]
        i += 1
        if h < 128:
            out += data[i:i + h + 1]
            i += h + 1
        elif h > 128:
            if i < n:
                out += bytes([data[i]]) * (257 - h)
            i += 1
    return bytes(out)


def _lzw(data: bytes) -> bytes:
    """TIFF LZW: MSB-first codes of 9 to 12 bits, ClearCode 256, EOI 257,
    code width growing one code early."""
    out = bytearray()
    table = [bytes([k]) for k in range(256)] + [b"", b""]
    width, bitpos, nbits = 9, 0, len(data) * 8
    prev = None
    while bitpos + width <= nbits:
        byte = bitpos >> 3
        chunk = int.from_bytes(data[byte:byte + 3].ljust(3, b"\0"), "big")
        code = (chunk >> (24 - (bitpos & 7) - width)) & ((1 << width) - 1)
        bitpos += width
        if code == 257:
            break
        if code == 256:
            table = table[:258]
            width, prev = 9, None
            continue
        if prev is None:
            entry = table[code]
            out += entry
            prev = entry
            continue
        if code < len(table):
            entry = table[code]
            table.append(prev + entry[:1])
        elif code == len(table):
            entry = prev + prev[:1]
            table.append(entry)
        else:
            raise TiffError("corrupt LZW stream")
        out += entry
        prev = entry
        if len(table) + 1 >= (1 << width) and width < 12:
            width += 1
    return bytes(out)


def _dtype(bits: int, fmt: int, bo: str) -> np.dtype:
    kind = {1: "u", 2: "i", 3: "f"}.get(fmt)
    if kind is None or bits not in (8, 16, 32, 64) or (kind == "f" and bits < 32):
        raise TiffError(f"unsupported TIFF sample format {fmt} with {bits} bits per sample")
    return np.dtype(f"{bo}{kind}{bits // 8}")


def read_tiff(path: str) -> np.ndarray:
    """The first image of a TIFF file as a (dim1, dim2) array in the
    reference's frame order (rows flipped up-down)."""
    with open(path, "rb") as f:
        buf = f.read()
    if len(buf) < 8 or buf[:2] not in (b"II", b"MM"):
        raise TiffError(f"{path}: not a TIFF file")
    bo = "<" if buf[:2] == b"II" else ">"
    (magic,) = struct.unpack_from(bo + "H", buf, 2)
    if magic == 42:
        big = False
        (off,) = struct.unpack_from(bo + "I", buf, 4)
    elif magic == 43:
        big = True
        (off,) = struct.unpack_from(bo + "Q", buf, 8)
    else:
        raise TiffError(f"{path}: not a TIFF file")
    t = _ifd(buf, bo, big, off)
    try:
        width, height = int(t[WIDTH][0]), int(t[HEIGHT][0])
    except KeyError:
        raise TiffError(f"{path}: image width or height missing") from None
    samples = int(t.get(SAMPLES, (1,))[0])
    bits = [int(b) for b in t.get(BITS, (1,))]
    photometric = int(t.get(PHOTOMETRIC, (1,))[0])
    if samples != 1 or len(bits) != 1 or photometric not in (0, 1):
        raise TiffError(f"unsupported TIFF: photometric interpretation {photometric}, "
                        f"{samples} sample{'' if samples == 1 else 's'} per pixel (bits per "
                        f"sample {bits}), image shape ({height}, {width}); "
                        f"{SINGLE_CHANNEL_ONLY}")
    if TILE_WIDTH in t:
        raise TiffError(f"{path}: tiled TIFF images are not supported (strips only)")
    dt = _dtype(bits[0], int(t.get(SAMPLE_FORMAT, (1,))[0]), bo)
    comp = int(t.get(COMPRESSION, (1,))[0])
    offsets = [int(x) for x in t[STRIP_OFFSETS]]
    counts = [int(x) for x in t.get(STRIP_COUNTS, ())]
    if len(counts) != len(offsets):
        raise TiffError(f"{path}: strip offsets and byte counts differ in number")
    rows_per_strip = int(t.get(ROWS_PER_STRIP, (height,))[0])
    row_bytes = width * dt.itemsize
    raw = bytearray()
    for k, (o, c) in enumerate(zip(offsets, counts)):
        if o + c > len(buf):
            raise TiffError(f"{path}: strip {k} beyond the end of the file")
        chunk = buf[o:o + c]
        if comp == 1:
            pass
        elif comp == 32773:
            chunk = _packbits(chunk)
        elif comp == 5:
            chunk = _lzw(chunk)
        elif comp in (8, 32946):
            chunk = zlib.decompress(chunk)
        else:
            raise TiffError(f"{path}: unsupported TIFF compression {comp}")
        rows = min(rows_per_strip, height - k * rows_per_strip)
        raw += chunk[:rows * row_bytes]
    if len(raw) != height * row_bytes:
        raise TiffError(f"unsupported TIFF: decoded {len(raw) // dt.itemsize} samples for an "
                        f"image of shape ({height}, {width}); {SINGLE_CHANNEL_ONLY}")
    img = np.frombuffer(bytes(raw), dtype=dt).reshape(height, width)
    predictor = int(t.get(PREDICTOR, (1,))[0])
    if predictor == 2:
        if dt.kind == "f":
            raise TiffError(f"{path}: horizontal differencing of floating-point samples is not "
                            f"supported")
        img = np.cumsum(img, axis=1, dtype=dt)
    elif predictor != 1:
        raise TiffError(f"{path}: unsupported TIFF predictor {predictor}")
    return np.ascontiguousarray(img[::-1].astype(dt.newbyteorder("=")))
