"""Bitshuffle + LZ4 chunks decoded on the GPU.

  - Device decode equals the raw frames bit for bit: uint16, uint32 and
    int32; block sizes from 16 elements to 49152 bytes; shapes with full
    blocks, a partial block and a verbatim tail; chunks from the product's
    LZ4 encoder and from the reference's own test encoders (literal-only,
    byte runs: src/integration/metal/tests.rs:297-357,
    src/integration/metal/bench.rs:150-230).
  - Decode-and-integrate equals integrating the raw frames bit for bit, 1D
    and 2D, with errors and raw sums (the reference's test,
    metal/tests.rs:362-401, on its own frames), and at full size on 3269 x
    3110 Eiger 9M frames.
  - A corrupt LZ4 stream fails with ValueError and leaves the device usable.
"""
import numpy as np
import pytest

import oracle as O
from dioptrin_amd import GpuPlan, bslz4, synth
from dioptrin_amd.device import DeviceArray, synchronize

pytestmark = pytest.mark.gpu

PONI = dict(pixel1=172e-6, pixel2=172e-6, distance=0.1, poni1=30 * 172e-6, poni2=35 * 172e-6,
            rot1=0.02, rot2=-0.01, wavelength=1e-10)


def frames_for(dtype, shape, n, seed):
    rng = np.random.default_rng(seed)
    f = rng.poisson(120.0, size=(n, *shape)).astype(np.int64)
    f[:, ::7, ::5] += rng.integers(0, 60000, size=f[:, ::7, ::5].shape)  # spots
    if dtype == np.int32:
        f[:, 0, :3] = -17  # negative pixels (gaps marked below zero)
    return f.astype(dtype)


def decode(chunks, shape, dtype):
    dc = bslz4.DeviceChunks(chunks, shape, np.dtype(dtype).name)
    out = DeviceArray((len(chunks), *shape), dtype)
    dc.decode(out)
    synchronize()
    dc.check()
    return out.numpy()


@pytest.mark.parametrize("dtype", [np.uint16, np.uint32, np.int32])
def test_device_decode_equals_raw_frames(dtype):
    eb = np.dtype(dtype).itemsize
    for shape in [(65, 67), (1, 5), (128, 96), (301, 257)]:
        fr = frames_for(dtype, shape, 3, 4)
        for block in (16 * eb, 1024, 2048, 8192, 49152):
            chunks = [bslz4.compress(f, block) for f in fr]
            assert np.array_equal(decode(chunks, shape, dtype), fr), (shape, block)
        for mode in (0, 1):
            chunks = [O.bslz4_encode(f, 2048, mode) for f in fr]
            assert np.array_equal(decode(chunks, shape, dtype), fr), (shape, mode)
    z = np.zeros((3, 200, 100), dtype)
    assert np.array_equal(decode([bslz4.compress(f) for f in z], z.shape[1:], dtype), z)


def test_reference_frames_integrate_bit_for_bit():
    shape = (65, 67)
    i = np.arange(65 * 67)
    u16 = ((i * 73 + 11) % 65000).astype(np.uint16).reshape(shape)
    u32 = ((i * 197 + 5) % 500_000).astype(np.uint32).reshape(shape)
    for frame, block, na in [(u16, 1024, None), (u32, 2048, 18)]:
        plan = GpuPlan(PONI, shape, n_radial=120 if na is None else 32, n_azimuthal=na)
        I, S, _ = plan.integrate_stack(frame[None], errors=True)
        for chunk in (O.bslz4_encode(frame, block, 0), bslz4.compress(frame, block)):
            bI, bS = plan.integrate_bslz4([chunk], frame.dtype.name, errors=True)
            assert np.array_equal(bI, I) and np.array_equal(bS, S)
        if na is not None:
            stk, err = plan.batch2d_bslz4_stack([chunk], "uint32", errors=True)
            assert stk.shape == (1, na, 32) and np.array_equal(stk.reshape(1, -1), I)


@pytest.mark.parametrize("precision", ["f32_compensated", "f32_wide"])
@pytest.mark.parametrize("dtype", [np.uint16, np.uint32, np.int32])
def test_decode_and_integrate_equals_direct(dtype, precision):
    shape = (301, 257)
    fr = frames_for(dtype, shape, 20, 6)
    chunks = [bslz4.compress(f) for f in fr]
    for na in (None, 24):
        plan = GpuPlan(PONI, shape, n_radial=200 if na is None else 60, n_azimuthal=na,
                       precision=precision)
        I, S, _ = plan.integrate_stack(fr, errors=True)
        bI, bS = plan.integrate_bslz4(chunks, np.dtype(dtype).name, errors=True)
        assert np.array_equal(bI, I) and np.array_equal(bS, S)
        raw = plan.integrate_raw(fr, errors=True)
        rI, rS = plan.integrate_bslz4(chunks, np.dtype(dtype).name, errors=True, raw=True)
        assert np.array_equal(rI, raw["sum_signal"]) and np.array_equal(rS, raw["sum_variance"])
        res = plan.batch1d_bslz4(chunks[:2], np.dtype(dtype).name)
        assert np.array_equal(res[1].intensity.ravel(), I[1])


def test_corrupt_stream_fails_and_device_stays_usable():
    shape = (128, 96)
    fr = frames_for(np.uint32, shape, 2, 8)
    chunks = [bslz4.compress(f, 2048) for f in fr]
    plan = GpuPlan(PONI, shape, n_radial=80)
    I, _, _ = plan.integrate_stack(fr)
    desc, _, _ = bslz4.scan(chunks[1:], fr[0].size, "uint32")
    bad = bytearray(chunks[1])
    # Zero the first match offset of block 3 of the second frame: the host
    # scan sees consistent sizes, the device decoder an invalid offset.
    src, cl = int(desc[3, 0]), int(desc[3, 1])
    blk = bytes(bad[src:src + cl])
    ip, found = 0, None
    while ip < cl:
        tok = blk[ip]
        ip += 1
        lit = tok >> 4
        if lit == 15:
            while True:
                s = blk[ip]
                ip += 1
                lit += s
                if s != 255:
                    break
        ip += lit
        if ip >= cl:
            break
        found = src + ip
        break
    assert found is not None, "block 3 has no match"
    bad[found:found + 2] = b"\x00\x00"
    with pytest.raises(ValueError, match="corrupt or truncated"):
        plan.integrate_bslz4([chunks[0], bytes(bad)])
    with pytest.raises(ValueError, match="truncated"):
        plan.integrate_bslz4([chunks[0], chunks[1][:-200]])
    bI, _ = plan.integrate_bslz4(chunks)
    assert np.array_equal(bI, I)


def test_eiger9m_full_size_bit_for_bit():
    name = "eiger9m_batch1d"
    w = synth.WORKLOADS[name]
    print("case bslz4 eiger9m: 8 frames 3269x3110", flush=True)
    dev = synth.frames_device(name, 8, seed=21)
    synchronize()
    fr = dev.numpy()
    chunks = [bslz4.compress(f) for f in fr]
    ratio = fr.nbytes / sum(len(c) for c in chunks)
    print(f"bslz4 eiger9m compression ratio {ratio:.2f}", flush=True)
    assert np.array_equal(decode(chunks, w["shape"], np.uint32), fr)
    plan = GpuPlan(synth.workload_poni(name), w["shape"], n_radial=w["n_radial"],
                   precision="f32_wide")
    I, S, _ = plan.integrate_stack(fr, errors=True)
    bI, bS = plan.integrate_bslz4(chunks, errors=True)
    assert np.array_equal(bI, I) and np.array_equal(bS, S)
