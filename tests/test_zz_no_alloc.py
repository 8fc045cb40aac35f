"""dio_integrate_device_noalloc (the entry point that allocates nothing)
against dio_integrate_device: the same bits, on two calls in a row (the
second with fewer frames), for a plan whose partials fit the fixed scratch
as they are and for one whose passes must be shortened to fit; plans whose
partials of one frame exceed the usable scratch are refused."""
import numpy as np
import pytest

from dioptrin_amd import GpuPlan, synth
from dioptrin_amd.device import DeviceArray, synchronize

pytestmark = pytest.mark.gpu


def outputs(plan, n, errors):
    oi = DeviceArray((n, plan.nbins), np.float64)
    os_ = DeviceArray((n, plan.nbins), np.float64) if errors else None
    return oi, os_


def run_both(plan, frames, errors, raw=False):
    n = frames.shape[0]
    ri, rs = outputs(plan, n, errors)
    flags = dict(errors=errors)
    if raw:
        from dioptrin_amd import _native as N
        dt = plan._check_device(frames, ri, rs, None, errors)
        N.check(N.lib().dio_integrate_device(
            plan._h, frames.ptr, N.DTYPES[dt], n, N.RAW_SUMS | (N.WITH_ERRORS if errors else 0),
            ri.ptr, None if rs is None else rs.ptr, None, None))
    else:
        plan.integrate_device(frames, ri, rs, **flags)
    ni, ns = outputs(plan, n, errors)
    plan.integrate_device_noalloc(frames, ni, ns, errors=errors, raw=raw)
    synchronize()
    assert np.array_equal(ni.numpy(), ri.numpy())
    if errors:
        assert np.array_equal(ns.numpy(), rs.numpy())


@pytest.mark.parametrize("errors", [False, True])
def test_noalloc_equals_device_entry_pilatus(errors):
    w = synth.WORKLOADS["pilatus1m_ps1d"]
    print(f"case noalloc pilatus errors={errors}", flush=True)
    plan = GpuPlan(synth.workload_poni("pilatus1m_ps1d"), w["shape"], n_radial=w["n_radial"],
                   polarization_factor=0.99, mask=synth.mask_pilatus(w["shape"], 0),
                   precision="f32_wide")
    plan.set_tile(16, 4)
    frames = DeviceArray.from_numpy(
        np.random.default_rng(41).poisson(150.0, size=(37,) + w["shape"]).astype(np.uint32))
    run_both(plan, frames, errors)
    run_both(plan, frames[:5], errors)          # a second, smaller call
    run_both(plan, frames[:19], errors, raw=True)


def test_noalloc_shortens_passes_at_batch1d_shape():
    """3269 x 3110, 2000 bins: a 16-frame pass of partials with errors is
    85 MB, beyond the 64 MiB scratch, so the call runs 8-frame passes; the
    bits equal the device entry point's 16-frame passes."""
    name = "eiger9m_batch1d"
    w = synth.WORKLOADS[name]
    print("case noalloc batch1d", flush=True)
    plan = GpuPlan(synth.workload_poni(name), w["shape"], n_radial=w["n_radial"],
                   precision="f32_wide")
    plan.set_tile(16, 4)
    frames = synth.frames_device(name, 21, seed=9)
    run_both(plan, frames, True)
    run_both(plan, frames[:3], True)
    run_both(plan, frames, False)


def test_noalloc_azimuthal_std_and_refusal(monkeypatch):
    from dioptrin_amd import _native as N
    w = synth.WORKLOADS["pilatus1m_ps1d"]
    print("case noalloc std and refusal", flush=True)
    plan = GpuPlan(synth.workload_poni("pilatus1m_ps1d"), w["shape"], n_radial=w["n_radial"],
                   precision="f32_wide")
    frames = DeviceArray.from_numpy(
        np.random.default_rng(42).poisson(80.0, size=(6,) + w["shape"]).astype(np.float32))
    outs = [tuple(DeviceArray((6, plan.nbins), np.float64) for _ in range(3)) for _ in range(2)]
    plan.integrate_device(frames, *outs[0], errors=True, azimuthal_std=True)
    plan.integrate_device_noalloc(frames, *outs[1], errors=True, azimuthal_std=True)
    synchronize()
    for a, b in zip(*outs):
        assert np.array_equal(a.numpy(), b.numpy(), equal_nan=True)
    # Usable scratch below one frame's partials: refused, nothing launched.
    monkeypatch.setenv("DIOPTRIN_NOALLOC_LIMIT", "4096")
    oi, _ = outputs(plan, 6, False)
    with pytest.raises(ValueError, match="fixed scratch"):
        plan.integrate_device_noalloc(frames, oi)
    monkeypatch.delenv("DIOPTRIN_NOALLOC_LIMIT")
    ri, _ = outputs(plan, 6, False)
    plan.integrate_device(frames, ri)
    plan.integrate_device_noalloc(frames, oi)
    synchronize()
    assert np.array_equal(oi.numpy(), ri.numpy())
    assert N.DIO_ERR_INVALID == 3


@pytest.mark.parametrize("errors", [False, True])
def test_noalloc_three_and_two_passes_per_launch(monkeypatch, errors):
    """A usable scratch of 3 (then 2) pass blocks: launches of 48 (32)
    frames whose partial rows hold 3 (2) passes side by side, reduced in
    groups of 32 frames; the bits equal the device entry point's 64-frame
    launches."""
    w = synth.WORKLOADS["pilatus1m_ps1d"]
    plan = GpuPlan(synth.workload_poni("pilatus1m_ps1d"), w["shape"], n_radial=w["n_radial"],
                   precision="f32_wide")
    plan.set_tile(16, 4)
    frames = DeviceArray.from_numpy(
        np.random.default_rng(43).poisson(90.0, size=(100,) + w["shape"]).astype(np.uint32))
    pass_bytes = plan.info["n_segments"] * 16 * 4 * (2 if errors else 1)
    for passes in (3, 2):
        print(f"case noalloc {passes} passes errors={errors}", flush=True)
        monkeypatch.setenv("DIOPTRIN_NOALLOC_LIMIT", str(passes * pass_bytes + 256))
        run_both(plan, frames, errors)
        run_both(plan, frames[:37], errors)
    monkeypatch.delenv("DIOPTRIN_NOALLOC_LIMIT")
