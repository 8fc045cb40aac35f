"""GPU parity at the full sizes of the benchmarked workloads.

Each workload's hot path runs on the device at its own shape and is compared
with the FP64 oracle:
  eiger9m_batch1d   256 device-resident uint32 frames of 3269 x 3110, 2000
                    bins pixel splitting, autotuned tile, hipGraph replay;
                    frames 0, 63, 64 and 255 against the oracle and bit for
                    bit against their single-frame runs; per-frame masks on
                    64 frames of that shape.
  eiger4m_ss2d      2167 x 2070, supersampling n_ss = 3: integrate2d with
                    errors at 1000 x 360 bins, and the azimuthal std at 1000
                    bins.
  pilatus1m_tilted  pyFAI-compatible splitting on the tilted geometry, 1D
                    3000 bins and 2D 500 x 360: every precision on intensity
                    and sigma (and the azimuthal std of the 1D table).

Tolerances (reference src/integration/cuda/tests.rs:206-250): per bin
|gpu - oracle| <= tol * magnitude, magnitude = the gather of |x|, tol 2e-6
for float32 storage and 1e-12 for float64; sigma is its own magnitude. The
frames are non-negative, so magnitude = |oracle| and the contract is a
relative error bound; f32_compensated, f32_wide and the azimuthal std are
held to relative error 1e-6 (f64: 1e-12). Each test prints its measured
maximum relative errors on lines starting with "maxrel".
"""
import numpy as np
import pytest

import oracle as O
from dioptrin_amd import GpuPlan, synth

pytestmark = pytest.mark.gpu

TOL = {"f32": 2e-6, "f32_compensated": 2e-6, "f32_wide": 2e-6, "f64": 1e-12}
REL = {"f32": 2e-6, "f32_compensated": 1e-6, "f32_wide": 1e-6, "f64": 1e-12}


def max_rel(a, b):
    """Largest |a - b| / |b| over bins where b != 0; a must be 0 where b is."""
    a, b = np.asarray(a), np.asarray(b)
    z = b == 0
    assert np.all(a[z] == 0), "nonzero output where the oracle gives 0"
    if not np.any(~z):
        return 0.0
    return float(np.max(np.abs(a[~z] - b[~z]) / np.abs(b[~z])))


def report(config, precision, output, rel):
    print(f"maxrel {config} {precision} {output} {rel:.3e}", flush=True)


def check_i_sigma(config, precision, I, S, i_ref, s_ref, mag):
    tol = TOL[precision]
    assert np.all(np.abs(I - i_ref) <= tol * mag), (config, precision, "intensity contract")
    rel_i = max_rel(I, i_ref)
    report(config, precision, "intensity", rel_i)
    assert rel_i <= REL[precision], (config, precision, rel_i)
    if S is not None:
        assert np.all(np.abs(S - s_ref) <= tol * s_ref), (config, precision, "sigma contract")
        rel_s = max_rel(S, s_ref)
        report(config, precision, "sigma", rel_s)
        assert rel_s <= REL[precision], (config, precision, rel_s)


def check_std(config, precision, D, d_ref):
    assert np.array_equal(np.isnan(D), np.isnan(d_ref)), (config, "empty bins differ")
    ok = ~np.isnan(d_ref)
    rel = max_rel(D[ok], d_ref[ok])
    report(config, precision, "std", rel)
    assert rel <= 1e-6, (config, precision, rel)


# ------------------------------------------------------------ eiger9m_batch1d

@pytest.fixture(scope="module")
def batch1d():
    name = "eiger9m_batch1d"
    w = synth.WORKLOADS[name]
    poni = synth.workload_poni(name)
    print(f"case {name}: oracle table", flush=True)
    P = O.Poni(**poni)
    lut = O.build_lut1d(P, *w["shape"], w["n_radial"])
    norm = O.normalization(P, *w["shape"], w["polarization"])
    return dict(w=w, poni=poni, lut=lut, norm=norm, den=O.denominators(lut, norm))


def test_batch1d_full_size_graph_matches_oracle(batch1d):
    """The benchmarked step: 256 device frames, autotuned tile, hipGraph."""
    from dioptrin_amd.device import DeviceArray, synchronize
    d = batch1d
    w = d["w"]
    plan = GpuPlan(d["poni"], w["shape"], n_radial=w["n_radial"], precision="f32_wide")
    assert plan.info["nnz"] == d["lut"].nnz
    tile = plan.autotune("uint32", 64)
    print(f"case eiger9m_batch1d: 256 frames, tile {tile}", flush=True)
    frames = synth.frames_device("eiger9m_batch1d", 256, seed=2026)
    out = DeviceArray((256, plan.nbins), np.float64)
    graph = plan.capture_graph(frames, out)
    graph.launch()
    synchronize()
    I = out.numpy()
    graph.launch()  # a replay gives the same bits
    synchronize()
    assert np.array_equal(out.numpy(), I)
    plan.set_tile(1, 1)
    one = DeviceArray((1, plan.nbins), np.float64)
    for k in (0, 63, 64, 255):
        frame = frames[k:k + 1].numpy()[0].astype(np.float64)
        i_ref, _ = O.gather(d["lut"], d["den"], frame, False)
        check_i_sigma("eiger9m_batch1d", "f32_wide", I[k], None, i_ref, None, np.abs(i_ref))
        plan.integrate_device(frames[k:k + 1], one)
        synchronize()
        assert np.array_equal(one.numpy()[0], I[k]), f"frame {k}: batch differs from alone"


@pytest.mark.parametrize("precision", ["f32_compensated", "f64"])
def test_batch1d_full_size_errors_each_precision(batch1d, precision):
    """Intensity and sigma at the batch1d shape for the default API precision
    and float64, 20 frames (a 16-frame pass and a partial one)."""
    from dioptrin_amd.device import DeviceArray, synchronize
    d = batch1d
    w = d["w"]
    print(f"case eiger9m_batch1d errors {precision}", flush=True)
    plan = GpuPlan(d["poni"], w["shape"], n_radial=w["n_radial"], precision=precision)
    frames = synth.frames_device("eiger9m_batch1d", 20, seed=77)
    oi = DeviceArray((20, plan.nbins), np.float64)
    os_ = DeviceArray((20, plan.nbins), np.float64)
    plan.integrate_device(frames, oi, os_, errors=True)
    synchronize()
    I, S = oi.numpy(), os_.numpy()
    for k in (0, 19):
        frame = frames[k:k + 1].numpy()[0].astype(np.float64)
        i_ref, s_ref = O.gather(d["lut"], d["den"], frame, True)
        check_i_sigma("eiger9m_batch1d", precision, I[k], S[k], i_ref, s_ref, np.abs(i_ref))


def test_batch1d_full_size_dynamic_masks(batch1d):
    """Per-frame masks on 64 device frames of the batch1d shape (the tile
    route of an unmasked plan): frames 0, 31 and 63 against the oracle's
    gather on the table compacted by each frame's mask."""
    from dioptrin_amd.device import DeviceArray, synchronize
    d = batch1d
    w = d["w"]
    shape = w["shape"]
    print("case eiger9m_batch1d dynamic masks", flush=True)
    plan = GpuPlan(d["poni"], shape, n_radial=w["n_radial"], precision="f32_wide")
    n = 64
    frames = synth.frames_device("eiger9m_batch1d", n, seed=5)
    base = (np.random.default_rng(3).random(shape[0] * shape[1]) < 0.05).astype(np.uint8)
    masks_h = np.stack([np.roll(base, 977 * k).reshape(shape) for k in range(n)])
    masks = DeviceArray.from_numpy(masks_h)
    oi = DeviceArray((n, plan.nbins), np.float64)
    os_ = DeviceArray((n, plan.nbins), np.float64)
    plan.integrate_masked_device(frames, masks, oi, os_, errors=True)
    synchronize()
    I, S = oi.numpy(), os_.numpy()
    for k in (0, 31, 63):
        fixed = O.fixed_table(d["lut"], masks_h[k])
        den = O.denominators(fixed, d["norm"])
        frame = frames[k:k + 1].numpy()[0].astype(np.float64)
        i_ref, s_ref = O.gather(fixed, den, frame, True)
        check_i_sigma("eiger9m_batch1d_dynamic_mask", "f32_wide", I[k], S[k], i_ref, s_ref,
                      np.abs(i_ref))


# ------------------------------------------------------------- eiger4m_ss2d

@pytest.fixture(scope="module")
def ss2d():
    name = "eiger4m_ss2d"
    w = synth.WORKLOADS[name]
    poni = synth.workload_poni(name)
    P = O.Poni(**poni)
    norm = O.normalization(P, *w["shape"], w["polarization"])
    frames = np.stack([synth.frame_eiger4m(s) for s in (1, 2)])
    return dict(w=w, poni=poni, P=P, norm=norm, frames=frames)


@pytest.mark.parametrize("precision", ["f32_compensated", "f32_wide", "f64"])
def test_ss2d_integrate2d_errors_full_size(ss2d, precision):
    d = ss2d
    w = d["w"]
    print(f"case eiger4m_ss2d integrate2d {precision}", flush=True)
    if "lut2d" not in d:
        d["lut2d"], _ = O.build_lut2d_ss(d["P"], *w["shape"], w["n_radial"], w["n_azimuthal"],
                                         w["n_ss"])
        d["den2d"] = O.denominators(d["lut2d"], d["norm"])
    plan = GpuPlan(d["poni"], w["shape"], method="supersampled", n_ss=w["n_ss"],
                   n_radial=w["n_radial"], n_azimuthal=w["n_azimuthal"], precision=precision)
    I, S, _ = plan.integrate_stack(d["frames"], errors=True)
    for k in range(d["frames"].shape[0]):
        i_ref, s_ref = O.gather(d["lut2d"], d["den2d"], d["frames"][k].astype(np.float64), True)
        check_i_sigma("eiger4m_ss2d_2d", precision, I[k], S[k], i_ref, s_ref, np.abs(i_ref))
    r = plan.integrate2d(d["frames"][0], errors=True)
    assert r.intensity.shape == (w["n_azimuthal"], w["n_radial"])
    assert np.array_equal(r.intensity.ravel(), I[0]) and np.array_equal(r.errors.ravel(), S[0])


def test_ss1d_azimuthal_std_full_size(ss2d):
    d = ss2d
    w = d["w"]
    print("case eiger4m_ss2d azimuthal std 1000 bins", flush=True)
    lut = O.build_lut1d_ss(d["P"], *w["shape"], w["n_radial"], w["n_ss"])
    plan = GpuPlan(d["poni"], w["shape"], method="supersampled", n_ss=w["n_ss"],
                   n_radial=w["n_radial"], precision="f32_compensated")
    I, S, D = plan.integrate_stack(d["frames"], errors=True, azimuthal_std=True)
    for k in range(d["frames"].shape[0]):
        i_ref, s_ref, d_ref = O.azimuthal_std(lut, None, d["norm"],
                                              d["frames"][k].astype(np.float64), errors=True)
        for name, a, b in (("intensity", I[k], i_ref), ("sigma", S[k], s_ref)):
            ok = ~np.isnan(d_ref)
            rel = max_rel(a[ok], b[ok])
            report("eiger4m_ss1d_std", "std_kernel", name, rel)
            assert rel <= 1e-6, (name, rel)
        check_std("eiger4m_ss1d_std", "std_kernel", D[k], d_ref)


# --------------------------------------------------------- pilatus1m_tilted

@pytest.fixture(scope="module")
def tilted():
    name = "pilatus1m_tilted"
    w = synth.WORKLOADS[name]
    poni = synth.workload_poni(name)
    P = O.Poni(**poni)
    norm = O.normalization(P, *w["shape"], w["polarization"])
    lut1 = O.build_lut1d_pyfai(P, *w["shape"], w["n_radial"])
    lut2, _ = O.build_lut2d_pyfai(P, *w["shape"], w["n_radial_2d"], w["n_azimuthal"])
    frames = np.stack([synth.frame_tilted(s) for s in (3, 4)])
    return dict(w=w, poni=poni, norm=norm, lut1=lut1, den1=O.denominators(lut1, norm),
                lut2=lut2, den2=O.denominators(lut2, norm), frames=frames)


@pytest.mark.parametrize("ndim", [1, 2])
@pytest.mark.parametrize("precision", ["f32", "f32_compensated", "f32_wide", "f64"])
def test_pyfai_tilted_precisions(tilted, precision, ndim):
    """The FP32 / compensated FP32 / wide / FP64 ablation on the tilted
    pyFAI-compatible tables: intensity and sigma against the FP64 oracle,
    and compensated float32 never worse than plain float32."""
    d = tilted
    w = d["w"]
    print(f"case pilatus1m_tilted {ndim}D {precision}", flush=True)
    if ndim == 1:
        plan = GpuPlan(d["poni"], w["shape"], method="pyfai_full", n_radial=w["n_radial"],
                       precision=precision)
        lut, den = d["lut1"], d["den1"]
    else:
        plan = GpuPlan(d["poni"], w["shape"], method="pyfai_full", n_radial=w["n_radial_2d"],
                       n_azimuthal=w["n_azimuthal"], precision=precision)
        lut, den = d["lut2"], d["den2"]
    I, S, _ = plan.integrate_stack(d["frames"], errors=True)
    for k in range(d["frames"].shape[0]):
        i_ref, s_ref = O.gather(lut, den, d["frames"][k].astype(np.float64), True)
        check_i_sigma(f"pilatus1m_tilted_{ndim}d", precision, I[k], S[k], i_ref, s_ref,
                      np.abs(i_ref))
    errs = d.setdefault(f"err{ndim}", {})
    i_ref, _ = O.gather(lut, den, d["frames"][0].astype(np.float64), False)
    errs[precision] = max_rel(I[0], i_ref)
    if "f32" in errs and "f32_compensated" in errs:
        assert errs["f32_compensated"] <= errs["f32"], errs


def test_pyfai_tilted_azimuthal_std(tilted):
    d = tilted
    w = d["w"]
    print("case pilatus1m_tilted azimuthal std", flush=True)
    plan = GpuPlan(d["poni"], w["shape"], method="pyfai_full", n_radial=w["n_radial"])
    I, S, D = plan.integrate_stack(d["frames"], errors=True, azimuthal_std=True)
    for k in range(d["frames"].shape[0]):
        i_ref, s_ref, d_ref = O.azimuthal_std(d["lut1"], None, d["norm"],
                                              d["frames"][k].astype(np.float64), errors=True)
        ok = ~np.isnan(d_ref)
        for name, a, b in (("intensity", I[k], i_ref), ("sigma", S[k], s_ref)):
            rel = max_rel(a[ok], b[ok])
            report("pilatus1m_tilted_std", "std_kernel", name, rel)
            assert rel <= 1e-6, (name, rel)
        check_std("pilatus1m_tilted_std", "std_kernel", D[k], d_ref)
