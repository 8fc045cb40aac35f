"""16-bit frames in the double-buffered tile kernel: one 4-byte load per
pixel pair when rows have an even width and the frames start 4-byte aligned,
two 2-byte loads otherwise. Either way a uint16 frame gives the bits of the
same values as uint32 (whose path is checked against the oracle elsewhere),
for intensity and sigma, every float32-weight precision, and batches of more
than one launch."""
import functools

import numpy as np
import pytest

from dioptrin_amd import GpuPlan, synth
from dioptrin_amd.device import DeviceArray, synchronize

pytestmark = pytest.mark.gpu


def misaligned_u16(frames: np.ndarray) -> DeviceArray:
    """A device copy of uint16 frames starting 2 bytes past a 4-byte
    boundary (a view into a buffer one element longer)."""
    n = frames.size
    buf = DeviceArray((n + 2,), np.uint16)
    host = np.zeros(n + 2, np.uint16)
    host[1:n + 1] = frames.ravel()
    buf.copy_from(host)
    view = DeviceArray.__new__(DeviceArray)
    view.shape, view.dtype, view.device, view._ctx = frames.shape, np.dtype(np.uint16), 0, buf._ctx
    view.nbytes = frames.nbytes
    view.ptr = buf.ptr + 2
    view._owner = buf
    assert view.ptr % 4 == 2
    return view


@functools.lru_cache(maxsize=1)
def frames_of(name, n=70):
    """n uint16 frames (a 64-frame launch and a 6-frame one) with the top
    value at one pixel, and their uint32 copy, on the device."""
    h = np.random.default_rng(5).integers(0, 2000, size=(n,) + synth.WORKLOADS[name]["shape"],
                                          dtype=np.uint16)
    h[:, 7, 11] = 65535
    return h, DeviceArray.from_numpy(h), DeviceArray.from_numpy(h.astype(np.uint32))


def run(plan, frames, n, errors):
    oi = DeviceArray((n, plan.nbins), np.float64)
    os_ = DeviceArray((n, plan.nbins), np.float64) if errors else None
    plan.integrate_device(frames, oi, os_, errors=errors)
    synchronize()
    return oi.numpy(), (os_.numpy() if errors else None)


@pytest.mark.parametrize("precision", ["f32_wide", "f32_compensated", "f32"])
@pytest.mark.parametrize("name", ["eiger4m_ss2d", "pilatus1m_ps1d"])
def test_u16_frames_equal_u32(name, precision):
    w = synth.WORKLOADS[name]
    print(f"case u16 {name} {precision}", flush=True)
    plan = GpuPlan(synth.workload_poni(name), w["shape"], n_radial=1000, precision=precision)
    h, d16, d32 = frames_of(name)
    n = h.shape[0]
    for lanes in (4, 2):
        plan.set_tile(16, lanes)
        for errors in (False, True):
            ri, rs = run(plan, d32, n, errors)
            gi, gs = run(plan, d16, n, errors)
            assert np.array_equal(gi, ri)
            if errors:
                assert np.array_equal(gs, rs)
    if w["shape"][1] % 2 == 0:
        # Even rows but a 2-byte aligned start: the two-load path.
        plan.set_tile(16, 4)
        ri, rs = run(plan, d32, n, True)
        gi, gs = run(plan, misaligned_u16(h), n, True)
        assert np.array_equal(gi, ri) and np.array_equal(gs, rs)


@pytest.mark.parametrize("precision", ["f32_wide", "f32_compensated"])
def test_u16_masked_rows_equal_u32(precision):
    """Per-frame masks on 16-bit frames with even rows take the fused
    masked-row kernel (4-byte pair loads): the bits of the same values as
    uint32 frames, whose route is checked against the oracle elsewhere; a
    2-byte aligned start takes the other route (masked copies, then two
    gathers), equal within the gather tolerance."""
    name = "eiger4m_ss2d"
    w = synth.WORKLOADS[name]
    print(f"case u16 masked {precision}", flush=True)
    plan = GpuPlan(synth.workload_poni(name), w["shape"], n_radial=1000, precision=precision)
    h, d16, d32 = frames_of(name)
    n = 37
    rng = np.random.default_rng(8)
    masks = DeviceArray.from_numpy((rng.random((n,) + w["shape"]) < 0.05).astype(np.uint8))
    for errors in (False, True):
        out = []
        for fr in (d32[:n], d16[:n]):
            oi = DeviceArray((n, plan.nbins), np.float64)
            os_ = DeviceArray((n, plan.nbins), np.float64) if errors else None
            plan.integrate_masked_device(fr, masks, oi, os_, errors=errors)
            synchronize()
            out.append((oi.numpy(), os_.numpy() if errors else None))
        assert np.array_equal(out[0][0], out[1][0])
        if errors:
            assert np.array_equal(out[0][1], out[1][1])
    mis = misaligned_u16(h[:n])
    oi = DeviceArray((n, plan.nbins), np.float64)
    plan.integrate_masked_device(mis, masks, oi)
    synchronize()
    ref = DeviceArray((n, plan.nbins), np.float64)
    plan.integrate_masked_device(d32[:n], masks, ref)
    synchronize()
    assert np.allclose(oi.numpy(), ref.numpy(), rtol=2e-6, atol=0)
