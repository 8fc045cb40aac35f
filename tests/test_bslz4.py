"""Bitshuffle + LZ4 on the CPU: the oracle decoder pinned, the product's
encoder and block pre-scan checked against it (no GPU).

The codec (HDF5 filter 32008: bitshuffle 0.5.2 with the LZ4 block format)
is a third-party dependency absent from the reference; its definitions are
pinned here by:
  - LZ4 blocks assembled by hand from the block format (literal runs,
    offset-1 and overlapping matches, 255-continuation lengths, a match
    reaching back across earlier matches) with the bytes they must give,
    and malformed blocks that must be rejected;
  - bit-transposition closed forms (elements 0..15 give planes of 0xAA,
    0xCC, 0xF0, 0x00FF patterns; a single set bit lands in one plane byte);
  - the reference's own test encoders (src/integration/metal/tests.rs:
    297-357 literal-only, src/integration/metal/bench.rs:150-230 byte
    runs) on the reference's own test frames (metal/tests.rs:362-401);
  - round trips of the product encoder's chunks through the oracle.
"""
import numpy as np
import pytest

import oracle as O
from dioptrin_amd import bslz4


# ------------------------------------------------------------ LZ4 blocks

def test_lz4_hand_assembled_blocks():
    cases = [
        (b"\x40abcd", b"abcd"),                                   # literals only
        (b"\x16a\x01\x00", b"a" * 11),                             # offset 1, length 10
        (b"\x35abc\x03\x00", b"abc" * 4),                          # overlap: offset 3, 9
        (b"\x32xyz\x03\x00" + b"\x20pq", b"xyz" * 3 + b"pq"),       # match then literals
        (b"\xf0\xff\x1e" + bytes(range(256)) + bytes(44), bytes(range(256)) + bytes(44)),
        (b"\x1f" + b"z" + b"\x01\x00" + b"\xff\x01", b"z" * (1 + 4 + 15 + 255 + 1)),
        # "ab", match offset 2 length 6 -> "abababab"; then "c" and a match
        # of offset 5 length 5 reaching back over the first match.
        (b"\x22ab\x02\x00\x11c\x05\x00", b"abababab" + b"c" + b"ababc"),
    ]
    for block, want in cases:
        assert O.lz4_decode_block(block, len(want)) == want, block[:8]
        # Exact capacity is enough; one byte less is rejected.
        assert O.lz4_decode_block(block, len(want) - 1) is None


def test_lz4_malformed_blocks_are_rejected():
    for block in [b"\x16a\x00\x00",       # offset 0
                  b"\x16a\x02\x00",       # offset beyond the output
                  b"\x40abc",             # literals past the end
                  b"\xf0",                # missing length continuation
                  b"\x16a\x01",           # half an offset
                  b"\x1fa\x01\x00\xff"]:  # match continuation cut off
        assert O.lz4_decode_block(block, 4096) is None, block


# ------------------------------------------------------------ bit transposition

def test_bitshuffle_closed_forms():
    x = np.arange(16, dtype=np.uint16)
    planes = np.frombuffer(O.bitshuffle(x), np.uint8).reshape(16, 2)
    assert planes[0].tolist() == [0xAA, 0xAA]
    assert planes[1].tolist() == [0xCC, 0xCC]
    assert planes[2].tolist() == [0xF0, 0xF0]
    assert planes[3].tolist() == [0x00, 0xFF]
    assert not planes[4:].any()
    # One set bit k of element i lands in plane k, byte i / 8, bit i % 8.
    for i, k in [(0, 0), (13, 5), (63, 31), (40, 17)]:
        y = np.zeros(64, np.uint32)
        y[i] = np.uint32(1) << k
        p = np.frombuffer(O.bitshuffle(y), np.uint8).reshape(32, 8)
        assert p.sum() == 1 << (i % 8) and p[k, i // 8] == 1 << (i % 8)
    rng = np.random.default_rng(1)
    for dt, n in [(np.uint16, 8), (np.uint32, 2048), (np.int32, 24)]:
        z = rng.integers(np.iinfo(dt).min, np.iinfo(dt).max, n, endpoint=True).astype(dt)
        assert np.array_equal(O.bitunshuffle(O.bitshuffle(z), dt, n), z)


# ------------------------------------------------------------ chunks

def reference_frames():
    """The reference's own bslz4 test frames (metal/tests.rs:362-401)."""
    pixels = 65 * 67
    i = np.arange(pixels)
    u16 = ((i * 73 + 11) % 65000).astype(np.uint16).reshape(65, 67)
    u32 = ((i * 197 + 5) % 500_000).astype(np.uint32).reshape(65, 67)
    return u16, u32


def test_reference_encoders_round_trip():
    u16, u32 = reference_frames()
    for frame, block in [(u16, 1024), (u32, 2048)]:
        for mode in (0, 1):
            ch = O.bslz4_encode(frame, block, mode)
            rc, out = O.bslz4_decode(ch, frame.size, frame.dtype)
            assert rc == 0 and np.array_equal(out, frame.ravel())
    # Byte runs compress the zero middle planes (a run reaching the block's
    # end stays literal in that encoder, so the top plane is set).
    small = ((np.arange(4096) % 7) | 0x80000000).astype(np.uint32)
    assert len(O.bslz4_encode(small, 8192, 1)) < len(O.bslz4_encode(small, 8192, 0)) / 4


@pytest.mark.parametrize("dtype", [np.uint16, np.uint32, np.int32])
def test_product_encoder_round_trips(dtype):
    rng = np.random.default_rng(2)
    info = np.iinfo(dtype)
    frames = [
        np.zeros((40, 33), dtype),
        rng.poisson(40.0, (65, 67)).astype(dtype),
        rng.integers(info.min, info.max, (64, 128), endpoint=True).astype(dtype),
        np.full((1, 5), info.max, dtype),            # only a verbatim tail
        np.arange(3 * 1031, dtype=np.int64).astype(dtype).reshape(3, 1031),
    ]
    for frame in frames:
        for block in (0, 16 * frame.itemsize, 1024, 8192, 49152):
            ch = bslz4.compress(frame, block)
            rc, out = O.bslz4_decode(ch, frame.size, dtype)
            assert rc == 0 and np.array_equal(out, frame.ravel()), (frame.shape, block)
    z = bslz4.compress(np.zeros((256, 256), dtype))
    assert len(z) < np.zeros((256, 256), dtype).nbytes / 50


def test_scan_descriptors_and_validation():
    frame = np.arange(65 * 67, dtype=np.uint32).reshape(65, 67)
    ch = bslz4.compress(frame, 2048)
    desc, max_comp, max_block = bslz4.scan([ch, ch], frame.size, "uint32")
    per = 65 * 67 // 512 + 1 + 1   # 8 full blocks, one partial block, a tail of 3
    assert desc.shape == (2 * per, 4) and max_block == 2048
    assert desc[0, 0] == 16 and desc[per, 0] == len(ch) + 16   # after the headers
    assert desc[per - 1, 1] == 0x80000000 | 12 and desc[per - 1, 3] == 12
    assert desc[per - 2, 3] == (65 * 67 % 512) // 8 * 8 * 4
    assert desc[per, 2] == frame.size
    assert max_comp == max(int(d) for d in desc[:, 1] if not d & 0x80000000)
    bad = bytearray(ch)
    cases = {
        "shorter than 12": bytes(ch[:8]),
        "does not match plan frame size": bytes(bytearray(b"\0" * 8) + ch[8:]),
        "unsupported Bitshuffle/LZ4 block size": bytes(ch[:8] + b"\0\0\0\x0c" + ch[12:]),
        "truncated Bitshuffle/LZ4 compressed block payload": bytes(ch[:40]),
        "truncated Bitshuffle/LZ4 tail": bytes(ch[:-1]),
    }
    bad[12:16] = (100000).to_bytes(4, "big")
    cases["exceeds LZ4 bound"] = bytes(bad)
    for msg, chunk in cases.items():
        with pytest.raises(ValueError, match=msg):
            bslz4.scan([chunk], frame.size, "uint32")
    with pytest.raises(ValueError, match="uint16, uint32 or int32"):
        bslz4.scan([ch], frame.size, "float32")
