"""GPU parity: the device-built tables and the HIP gathers against the FP64
oracle (bit-exact CSR; stated tolerances for floating-point outputs).

Tolerance contract (reference src/integration/cuda/tests.rs:206-250):
|gpu - oracle| <= tol * magnitude per bin, magnitude = gather of |x|;
tol = 2e-6 for float32-storage precisions, 1e-12 for f64. In addition the
relative error <= 1e-6 against the FP64 oracle wherever
the magnitude is not dominated by cancellation.
"""
import math

import numpy as np
import pytest

import oracle as O
from dioptrin_amd import GpuPlan, synth

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps

# name, poni kwargs, shape, bins, method, unit, radial_range
GEOMETRIES = [
    ("tilted_7x9", dict(pixel1=1e-4, pixel2=1.3e-4, distance=0.02, poni1=3e-4, poni2=4e-4,
                        rot1=0.137, rot2=-0.073, rot3=0.419, wavelength=1e-10), (7, 9), 20,
     "pixel_split", "2th_rad", None),
    ("tilted_3x4", dict(pixel1=1e-4, pixel2=1.3e-4, distance=0.02, poni1=-3e-4, poni2=2e-4,
                        rot1=0.137, rot2=-0.073, rot3=0.419, wavelength=1e-10), (3, 4), 7,
     "pixel_split", "2th_rad", None),
    ("beam_64x80", dict(pixel1=172e-6, pixel2=172e-6, distance=0.05, poni1=24 * 172e-6,
                        poni2=36 * 172e-6, wavelength=1e-10), (64, 80), 32, "pixel_split",
     "2th_rad", None),
    ("beam_9x11_deg", dict(pixel1=1e-4, pixel2=1e-4, distance=0.05, poni1=4.3e-4, poni2=5.6e-4,
                           wavelength=1e-10), (9, 11), 200, "pixel_split", "2th_deg", None),
    ("beam_9x11_range", dict(pixel1=1e-4, pixel2=1e-4, distance=0.05, poni1=4.5e-4,
                             poni2=5.5e-4, wavelength=1e-10), (9, 11), 200, "pixel_split",
     "2th_deg", (math.degrees(math.atan(0.5e-4 / 0.05)), 1.2)),
    ("counting_tilted", dict(pixel1=1.7e-4, pixel2=2.1e-4, distance=0.14, poni1=-0.003,
                             poni2=-0.002, rot1=0.007, rot2=-0.009, wavelength=1e-10), (23, 29),
     40, "counting", "2th_rad", None),
    ("counting_range", dict(pixel1=1e-4, pixel2=1e-4, distance=0.05, poni1=4.5e-4, poni2=5.5e-4,
                            wavelength=1e-10), (9, 11), 10, "counting", "2th_deg", (0.1, 0.8)),
]


def oracle_lut(poni, shape, bins, method, unit, rr):
    return O.build_lut1d(O.Poni(**poni), shape[0], shape[1], bins, method, unit, rr)


def assert_same_table(plan, lut, masked=False):
    rp, col, w64, w32 = plan.export_csr(masked)
    assert np.array_equal(rp.astype(np.uint64), lut.row_ptr), "row pointers differ"
    assert np.array_equal(col.astype(np.uint64), lut.col), "column indices differ"
    w = lut.w if lut.w is not None else np.ones(lut.nnz)
    assert np.array_equal(w64.view(np.uint64), w.view(np.uint64)), "f64 weight bits differ"
    assert np.array_equal(w32, w.astype(np.float32)), "f32 weights differ"


@pytest.mark.parametrize("g", GEOMETRIES, ids=[g[0] for g in GEOMETRIES])
def test_table_bit_exact(g):
    name, poni, shape, bins, method, unit, rr = g
    print(f"case {name}", flush=True)
    plan = GpuPlan(poni, shape, method=method, unit=unit, n_radial=bins, radial_range=rr)
    lut = oracle_lut(poni, shape, bins, method, unit, rr)
    assert_same_table(plan, lut)
    assert plan.info["hist_min"] == lut.hist_min and plan.info["hist_step"] == lut.hist_step
    assert np.array_equal(plan.radial, lut.bins())


@pytest.mark.parametrize("workload", ["pilatus1m_ps1d", "eiger9m_batch1d"])
def test_full_size_table_bit_exact(workload):
    print(f"case {workload}", flush=True)
    w = synth.WORKLOADS[workload]
    poni = synth.workload_poni(workload)
    mask = synth.mask_pilatus(w["shape"], 0) if workload == "pilatus1m_ps1d" else None
    plan = GpuPlan(poni, w["shape"], n_radial=w["n_radial"],
                   polarization_factor=w["polarization"], mask=mask)
    lut = oracle_lut(poni, w["shape"], w["n_radial"], "pixel_split", "2th_rad", None)
    assert_same_table(plan, lut)
    if mask is not None:
        assert_same_table(plan, O.fixed_table(lut, mask), masked=True)
    print(f"  nnz {lut.nnz} uncertain angles {plan.info['uncertain_angles']} "
          f"singular {plan.info['singular_pixels']} build {plan.info['build_ms']:.1f} ms",
          flush=True)
    norm = O.normalization(O.Poni(**poni), *w["shape"], w["polarization"])
    assert np.allclose(plan.normalization(), norm, rtol=1e-14, atol=0)
    fixed = O.fixed_table(lut, mask) if mask is not None else lut
    assert np.allclose(plan.denominators(), O.denominators(fixed, norm), rtol=1e-12, atol=0)


# ----------------------------------------------------------------- gathers

@pytest.fixture(scope="module")
def pilatus():
    w = synth.WORKLOADS["pilatus1m_ps1d"]
    poni = synth.workload_poni("pilatus1m_ps1d")
    mask = synth.mask_pilatus(w["shape"], 0)
    lut = oracle_lut(poni, w["shape"], w["n_radial"], "pixel_split", "2th_rad", None)
    norm = O.normalization(O.Poni(**poni), *w["shape"], w["polarization"])
    fixed = O.fixed_table(lut, mask)
    den = O.denominators(fixed, norm)
    frames = np.stack([synth.frame_pilatus(s) for s in range(3)])
    return dict(w=w, poni=poni, mask=mask, lut=lut, fixed=fixed, norm=norm, den=den,
                frames=frames)


def oracle_frame(d, frame):
    i, s = O.gather(d["fixed"], d["den"], frame.astype(np.float64), True)
    mag, _ = O.gather(d["fixed"], d["den"], np.abs(frame.astype(np.float64)), False)
    return i, s, np.abs(mag)


@pytest.mark.parametrize("precision,tol", [("f32", 2e-6), ("f32_compensated", 2e-6),
                                           ("f32_wide", 2e-6), ("f64", 1e-12)])
def test_gather_matches_oracle(pilatus, precision, tol):
    d = pilatus
    print(f"case gather {precision}", flush=True)
    plan = GpuPlan(d["poni"], d["w"]["shape"], n_radial=d["w"]["n_radial"],
                   polarization_factor=0.99, mask=d["mask"], precision=precision)
    I, S, _ = plan.integrate_stack(d["frames"], errors=True)
    for k, frame in enumerate(d["frames"]):
        i, s, mag = oracle_frame(d, frame)
        assert np.all(np.abs(I[k] - i) <= tol * mag), precision
        assert np.all(np.abs(S[k] - s) <= tol * s + 1e-300), precision
        big = mag > 0
        rel = np.abs(I[k][big] - i[big]) / np.abs(i[big])
        if precision != "f32":  # plain float32 is held to the contract above only
            assert rel.max() <= (1e-6 if precision != "f64" else 1e-12), (precision, rel.max())


def test_compensated_beats_plain_f32(pilatus):
    d = pilatus
    errs = {}
    for precision in ("f32", "f32_compensated", "f32_wide"):
        plan = GpuPlan(d["poni"], d["w"]["shape"], n_radial=d["w"]["n_radial"],
                       polarization_factor=0.99, mask=d["mask"], precision=precision)
        I, _, _ = plan.integrate_stack(d["frames"][:1])
        i, _, _ = oracle_frame(d, d["frames"][0])
        ok = np.abs(i) > 0
        errs[precision] = np.max(np.abs(I[0][ok] - i[ok]) / np.abs(i[ok]))
    print("max relative error", errs, flush=True)
    assert errs["f32_compensated"] <= errs["f32"]
    assert errs["f32_wide"] <= 1e-6 and errs["f32_compensated"] <= 1e-6


TILE_SHAPES = [(1, 1), (2, 2), (4, 4), (8, 2), (8, 4), (16, 1), (16, 2), (16, 4)]


@pytest.mark.parametrize("precision", ["f32", "f32_compensated", "f32_wide", "f64"])
@pytest.mark.parametrize("tile", TILE_SHAPES, ids=[f"F{f}G{g}" for f, g in TILE_SHAPES])
def test_batch_equals_single_frame_bitwise(pilatus, tile, precision):
    if precision == "f64" and tile[0] > 8:
        pytest.skip("float64 passes hold at most 8 frames")
    d = pilatus
    plan = GpuPlan(d["poni"], d["w"]["shape"], n_radial=d["w"]["n_radial"],
                   polarization_factor=0.99, mask=d["mask"], precision=precision)
    rng = np.random.default_rng(5)
    # 53 frames: the last pass is partial (5 frames); 16-frame passes of
    # 32-bit frames run as one 4-pass launch with reductions of 32 and 21
    # frames, smaller passes reduce two passes (or one float64 pass) at a time.
    frames = rng.poisson(200.0, size=(53,) + d["w"]["shape"]).astype(np.float32)
    plan.set_tile(1, 1)
    single = [plan.integrate_stack(frames[k:k + 1], errors=True) for k in range(frames.shape[0])]
    plan.set_tile(*tile)
    I, S, _ = plan.integrate_stack(frames, errors=True)
    for k in range(frames.shape[0]):
        assert np.array_equal(I[k], single[k][0][0])
        assert np.array_equal(S[k], single[k][1][0])


@pytest.mark.parametrize("dtype", [np.uint32, np.float64], ids=["u32", "f64"])
def test_capped_grid_many_tiles_per_workgroup(pilatus, dtype, monkeypatch):
    """Eight workgroups walk the Pilatus tiles (about 130 each, so the
    per-workgroup tile metadata is reloaded twice); results equal the
    full-grid run bit for bit, for register-prefetched (u32) and directly
    staged (f64) frames."""
    d = pilatus
    frames = np.random.default_rng(11).poisson(150.0, size=(19,) + d["w"]["shape"]).astype(dtype)

    def run():
        plan = GpuPlan(d["poni"], d["w"]["shape"], n_radial=d["w"]["n_radial"],
                       polarization_factor=0.99, mask=d["mask"], precision="f32_wide")
        out = []
        for tile in ((16, 2), (8, 4), (1, 1)):
            plan.set_tile(*tile)
            out.append(plan.integrate_stack(frames, errors=True)[:2])
        return out

    monkeypatch.delenv("DIOPTRIN_TILE_WORKGROUPS", raising=False)
    full = run()
    monkeypatch.setenv("DIOPTRIN_TILE_WORKGROUPS", "8")
    capped = run()
    for (i0, s0), (i1, s1) in zip(full, capped):
        assert np.array_equal(i0, i1)
        assert np.array_equal(s0, s1)


@pytest.mark.parametrize("tile", [(16, 1), (16, 2), (16, 4)], ids=["G1", "G2", "G4"])
def test_double_buffered_equals_single_buffered(pilatus, tile, monkeypatch):
    """16-frame passes of 32-bit frames run the double-buffered kernel (one
    1024-thread workgroup per CU, two LDS tiles); DIOPTRIN_TILE_DB=0 selects
    the single-buffered one. Both give the same bits, with and without
    errors, for u32 and f32 frames."""
    d = pilatus
    rng = np.random.default_rng(12)
    frames = rng.poisson(120.0, size=(21,) + d["w"]["shape"])

    def run():
        plan = GpuPlan(d["poni"], d["w"]["shape"], n_radial=d["w"]["n_radial"],
                       polarization_factor=0.99, mask=d["mask"], precision="f32_wide")
        plan.set_tile(*tile)
        return [plan.integrate_stack(frames.astype(dt), errors=True)[:2]
                for dt in (np.uint32, np.float32)]

    monkeypatch.delenv("DIOPTRIN_TILE_DB", raising=False)
    db = run()
    monkeypatch.setenv("DIOPTRIN_TILE_DB", "0")
    single = run()
    for (i0, s0), (i1, s1) in zip(db, single):
        assert np.array_equal(i0, i1)
        assert np.array_equal(s0, s1)


FLIP_CASES = [(1, 1, np.uint32), (16, 2, np.uint32), (16, 4, np.float32), (8, 2, np.uint32),
              (16, 2, np.uint16), (4, 4, np.float64)]


@pytest.mark.parametrize("tile,lanes,dtype", FLIP_CASES,
                         ids=[f"F{f}G{g}-{np.dtype(t).name}" for f, g, t in FLIP_CASES])
def test_flip_vertical_input_is_bitwise_detector_order(pilatus, tile, lanes, dtype):
    """InputOrder::FlipVertical (fixed.rs:20-24): frames whose rows arrive
    bottom-up, integrated by a flip-order plan, give the bits of the upright
    frames with a detector-order plan (cuda/tests.rs:528-562), for intensity,
    sigma and azimuthal std, through the per-frame, single- and
    double-buffered tile kernels and every staging path."""
    d = pilatus
    rng = np.random.default_rng(21)
    frames = rng.poisson(150.0, size=(21,) + d["w"]["shape"]).astype(dtype)
    flipped = np.ascontiguousarray(frames[:, ::-1, :])
    out = {}
    for order, fr in (("detector", frames), ("flip_vertical", flipped)):
        plan = GpuPlan(d["poni"], d["w"]["shape"], n_radial=d["w"]["n_radial"],
                       polarization_factor=0.99, mask=d["mask"], precision="f32_wide",
                       input_order=order)
        plan.set_tile(tile, lanes)
        out[order] = (plan.integrate_stack(fr, errors=True)[:2],
                      plan.integrate_stack(fr[:3], errors=True, azimuthal_std=True))
    (i0, s0), std0 = out["detector"]
    (i1, s1), std1 = out["flip_vertical"]
    assert np.array_equal(i0, i1) and np.array_equal(s0, s1)
    for a, b in zip(std0, std1):
        assert np.array_equal(a, b, equal_nan=True)


def test_flip_vertical_2d():
    """The flipped row order for a 2D table (500 x 360 bins, tiles of several
    LDS chunks): the same bits as the upright frames."""
    g = GEOMETRIES_2D[-1]
    name, poni, shape, nr, na = g[:5]
    frames = np.random.default_rng(22).poisson(40.0, size=(17,) + shape).astype(np.uint32)
    res = []
    for order, fr in (("detector", frames),
                      ("flip_vertical", np.ascontiguousarray(frames[:, ::-1, :]))):
        plan = GpuPlan(poni, shape, n_radial=nr, n_azimuthal=na, precision="f32_wide",
                       input_order=order)
        res.append(plan.integrate_stack(fr, errors=True)[:2])
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_raw_sums_match_oracle(pilatus):
    """Sums before division (integrate1d_raw, lut/integrate.rs:372-500):
    sum_signal and the Poisson sum_variance within the gather tolerance of the
    FP64 oracle's sums, sum_norm and count bit-exact (sequential row sums over
    the masked table), and the batched kernel's raw sums equal the per-frame
    path's bit for bit."""
    d = pilatus
    plan = GpuPlan(d["poni"], d["w"]["shape"], n_radial=d["w"]["n_radial"],
                   polarization_factor=0.99, mask=d["mask"], precision="f32_wide")
    npix = d["w"]["shape"][0] * d["w"]["shape"][1]
    ones = np.ones(plan.nbins)
    frames = np.random.default_rng(23).poisson(120.0, size=(19,) + d["w"]["shape"]).astype(
        np.uint32)
    plan.set_tile(16, 4)
    raw = plan.integrate_raw(frames, errors=True)
    # The normalization (cos, powers) is not bit-critical: the device's
    # denominators agree with the oracle's to 1e-12.
    assert np.allclose(raw["sum_norm"], d["den"], rtol=1e-12, atol=0)
    # With the device's own normalization the sequential row sums are exact.
    assert np.array_equal(raw["sum_norm"], O.denominators(d["fixed"], plan.normalization()))
    assert np.array_equal(raw["count"], O.denominators(d["fixed"], np.ones(npix)))
    for k in (0, 7, 18):
        s_ref, sig_ref = O.gather(d["fixed"], ones, frames[k].astype(np.float64), True)
        assert np.all(np.abs(raw["sum_signal"][k] - s_ref) <= 2e-6 * np.abs(s_ref))
        v_ref = sig_ref ** 2
        assert np.all(np.abs(raw["sum_variance"][k] - v_ref) <= 2e-6 * v_ref + 1e-300)
    plan.set_tile(1, 1)
    one = plan.integrate_raw(frames, errors=True)
    assert np.array_equal(one["sum_signal"], raw["sum_signal"])
    assert np.array_equal(one["sum_variance"], raw["sum_variance"])


def test_pinned_host_frames_equal_pageable(pilatus):
    """Frames in page-locked host memory (dio_host_alloc) integrate to the
    same bits as the same frames in pageable memory."""
    from dioptrin_amd.device import PinnedArray
    d = pilatus
    plan = GpuPlan(d["poni"], d["w"]["shape"], n_radial=d["w"]["n_radial"],
                   polarization_factor=0.99, mask=d["mask"], precision="f32_wide")
    frames = np.random.default_rng(24).poisson(80.0, size=(20,) + d["w"]["shape"]).astype(
        np.uint32)
    pin = PinnedArray(frames.shape, frames.dtype)
    pin.array[...] = frames
    a = plan.integrate_stack(frames, errors=True)
    b = plan.integrate_stack(pin.array, errors=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    del pin


def test_pixel_dtypes_equal_float64(pilatus):
    d = pilatus
    plan = GpuPlan(d["poni"], d["w"]["shape"], n_radial=d["w"]["n_radial"],
                   polarization_factor=0.99, mask=d["mask"], precision="f32_compensated")
    base = np.minimum(d["frames"][0], 60000).astype(np.uint16)
    ref, _, _ = plan.integrate_stack(base.astype(np.float64)[None])
    for dt in (np.uint16, np.uint32, np.int32, np.float32):
        got, _, _ = plan.integrate_stack(base.astype(dt)[None])
        assert np.array_equal(got, ref), dt


def test_azimuthal_std_matches_oracle(pilatus):
    d = pilatus
    plan = GpuPlan(d["poni"], d["w"]["shape"], n_radial=d["w"]["n_radial"],
                   polarization_factor=0.99, mask=d["mask"])
    I, S, D = plan.integrate_stack(d["frames"][:2], errors=True, azimuthal_std=True)
    for k in range(2):
        i, s, std = O.azimuthal_std(d["lut"], d["mask"], d["norm"],
                                    d["frames"][k].astype(np.float64), errors=True)
        both_nan = np.isnan(std) & np.isnan(D[k])
        assert np.array_equal(np.isnan(std), np.isnan(D[k]))
        ok = ~both_nan
        for name, a, b in (("I", I[k], i), ("sigma", S[k], s), ("std", D[k], std)):
            rel = np.abs(a[ok] - b[ok]) / np.maximum(np.abs(b[ok]), 1e-300)
            assert rel.max() <= 1e-6, (name, rel.max())


def test_azimuthal_std_special_values_match_oracle(pilatus):
    """The reference's contribution filter (azimuthal.rs:60-79) on the
    device: NaN rows, infinite pixels and negative values are skipped or
    kept exactly as the oracle does, a frame far from zero (1e6 plus Poisson
    counts) keeps the spread, and rows left without a valid contribution
    give (0, 0, NaN)."""
    d = pilatus
    shape = d["w"]["shape"]
    plan = GpuPlan(d["poni"], shape, n_radial=d["w"]["n_radial"],
                   polarization_factor=0.99, mask=d["mask"])
    rng = np.random.default_rng(31)
    frames = rng.poisson(60.0, size=(3,) + shape).astype(np.float64)
    frames[0, 421:622, 390:591] = np.nan        # the inner rings around the beam: no valid pixel
    frames[0, 500:520, 100:700] = np.inf
    frames[1] -= 75.0                           # negative values: kept, sigma uses max(s, 0)
    frames[1, 700:710, :] = -np.inf
    frames[2] += 1.0e6
    I, S, D = plan.integrate_stack(frames, errors=True, azimuthal_std=True)
    for k in range(3):
        i, s, std = O.azimuthal_std(d["lut"], d["mask"], d["norm"], frames[k], errors=True)
        assert np.array_equal(np.isnan(std), np.isnan(D[k])), k
        empty = np.isnan(std)
        assert np.all(I[k][empty] == 0) and np.all(S[k][empty] == 0), k
        ok = ~empty
        for name, a, b in (("I", I[k], i), ("sigma", S[k], s), ("std", D[k], std)):
            rel = np.abs(a[ok] - b[ok]) / np.maximum(np.abs(b[ok]), 1e-300)
            assert rel.max() <= 1e-6, (k, name, rel.max())
    assert np.isnan(D[0]).sum() > 10  # the NaN rows empty whole bins


def test_azimuthal_std_batch_equals_single_frame_bitwise(pilatus):
    """The spread kernel stages 8 frames per pass; each frame's results are
    the bits it gets alone, for batches that fill a pass and ones that do
    not, including frames with NaN and infinite pixels."""
    d = pilatus
    plan = GpuPlan(d["poni"], d["w"]["shape"], n_radial=d["w"]["n_radial"],
                   polarization_factor=0.99, mask=d["mask"], precision="f32_wide")
    frames = np.random.default_rng(25).poisson(90.0, size=(11,) + d["w"]["shape"]).astype(
        np.float32)
    frames[2, :40, :] = np.nan
    frames[5, 100:110, 200:300] = np.inf
    frames[9, 300:340, :] = np.nan
    I, S, D = plan.integrate_stack(frames, errors=True, azimuthal_std=True)
    for k in range(11):
        i, s, dd = plan.integrate_stack(frames[k:k + 1], errors=True, azimuthal_std=True)
        assert np.array_equal(I[k], i[0]) and np.array_equal(S[k], s[0])
        assert np.array_equal(D[k], dd[0], equal_nan=True)


def test_edge_cases_masked_inf_empty():
    poni = dict(pixel1=1e-4, pixel2=1e-4, distance=0.05, poni1=4.5e-4, poni2=5.5e-4,
                wavelength=1e-10)
    full = GpuPlan(poni, (9, 11), n_radial=20, mask=np.ones((9, 11), np.uint8))
    I, S, _ = full.integrate_stack(np.ones((2, 9, 11), np.float32), errors=True)
    assert np.all(I == 0) and np.all(S == 0)
    plan = GpuPlan(poni, (9, 11), n_radial=20)
    I, _, _ = plan.integrate_stack(np.zeros((0, 9, 11), np.float32))
    assert I.shape == (0, 20)
    frame = np.ones((9, 11), np.float32)
    frame[4, 5] = np.inf
    I, _, _ = plan.integrate_stack(frame[None])
    lut = O.build_lut1d(O.Poni(**poni), 9, 11, 20)
    rows = np.repeat(np.arange(20), np.diff(lut.row_ptr).astype(int))
    hit = np.unique(rows[lut.col == 4 * 11 + 5])
    assert np.all(np.isinf(I[0][hit]))
    with pytest.raises(ValueError):
        plan.integrate_stack(np.ones((1, 9, 10), np.float32))


def test_device_resident_path_and_graph_replay(pilatus):
    from dioptrin_amd.device import DeviceArray, synchronize
    d = pilatus
    plan = GpuPlan(d["poni"], d["w"]["shape"], n_radial=d["w"]["n_radial"],
                   polarization_factor=0.99, mask=d["mask"])
    frames = DeviceArray.from_numpy(d["frames"])
    n = frames.shape[0]
    out = DeviceArray((n, plan.nbins), np.float64)
    plan.integrate_device(frames, out)
    synchronize()
    host, _, _ = plan.integrate_stack(d["frames"])
    eager = out.numpy()
    assert np.array_equal(eager, host)
    g_out = DeviceArray((n, plan.nbins), np.float64)
    graph = plan.capture_graph(frames, g_out)
    for rep in range(3):
        g_out.zero_()
        synchronize()
        graph.launch()
        synchronize()
        plan.integrate_stack(d["frames"][:1])  # unrelated eager work between replays
        assert np.array_equal(g_out.numpy(), eager), rep


def test_pass_timing_matches_plain_run(pilatus):
    """The timed integration (HIP events around each launch of the gather
    kernel) writes the same outputs as the plain device run and reports one
    positive kernel duration per launch: one pass per launch for 2-frame
    passes, up to four passes per launch for the double-buffered 16-frame
    kernel."""
    from dioptrin_amd.device import DeviceArray, synchronize
    d = pilatus
    plan = GpuPlan(d["poni"], d["w"]["shape"], n_radial=d["w"]["n_radial"],
                   polarization_factor=0.99, mask=d["mask"])
    plan.set_tile(2, 2)
    frames = DeviceArray.from_numpy(np.concatenate([d["frames"]] * 2)[:5])
    n = frames.shape[0]
    ref = DeviceArray((n, plan.nbins), np.float64)
    out = DeviceArray((n, plan.nbins), np.float64)
    plan.integrate_device(frames, ref)
    synchronize()
    for rep in range(2):
        t = plan.time_passes(frames, out)
        assert [f for _, f in t] == [2, 2, 1] and all(ms > 0 for ms, _ in t), t
        assert np.array_equal(out.numpy(), ref.numpy()), rep
    frames = DeviceArray.from_numpy(
        np.random.default_rng(4).poisson(90.0, size=(37,) + d["w"]["shape"]).astype(np.uint32))
    ref = DeviceArray((37, plan.nbins), np.float64)
    out = DeviceArray((37, plan.nbins), np.float64)
    plan.set_tile(16, 2)
    plan.integrate_device(frames, ref)
    synchronize()
    t = plan.time_passes(frames, out)
    assert [f for _, f in t] == [37] and all(ms > 0 for ms, _ in t), t
    assert np.array_equal(out.numpy(), ref.numpy())


def test_paired_passes_with_multi_chunk_tiles(monkeypatch):
    """The double-buffered kernel runs up to four 16-frame passes per launch
    over one LDS copy of each tile's table. A 500 x 360 2D table has tiles
    whose table takes several LDS chunks (up to 4208 groups); every pass after
    the first of such a tile reloads the first chunk. Results equal the
    single-buffered kernel's (one pass per launch, every chunk loaded per
    pass) bit for bit, for 4 + 2 passes (16, 16, 16, 16 | 16, 3 frames)."""
    g = GEOMETRIES_2D[-1]
    name, poni, shape, nr, na = g[:5]
    frames = np.random.default_rng(13).poisson(60.0, size=(83,) + shape).astype(np.uint32)

    def run():
        plan = GpuPlan(poni, shape, n_radial=nr, n_azimuthal=na, precision="f32_wide")
        assert plan.info["max_tile_groups"] > 3 * 672  # several chunks of 672 groups
        plan.set_tile(16, 2)
        return plan.integrate_stack(frames, errors=True)[:2]

    monkeypatch.delenv("DIOPTRIN_TILE_DB", raising=False)
    i0, s0 = run()
    monkeypatch.setenv("DIOPTRIN_TILE_DB", "0")
    i1, s1 = run()
    assert np.array_equal(i0, i1)
    assert np.array_equal(s0, s1)


def test_device_synthetic_frames_are_poisson():
    from dioptrin_amd.device import synthetic_poisson_u32
    rate = np.full((256, 256), 40.0, np.float32)
    rate[:128] = 3.0
    fr = synthetic_poisson_u32(rate, [0.0, 100.0], seed=7).numpy().astype(np.float64)
    for f, bg in ((0, 0.0), (1, 100.0)):
        for sl, lam in ((slice(0, 128), 3.0 + bg), (slice(128, 256), 40.0 + bg)):
            x = fr[f, sl]
            assert abs(x.mean() - lam) < 5 * math.sqrt(lam / x.size)
            assert abs(x.var() / lam - 1.0) < 0.05
    again = synthetic_poisson_u32(rate, [0.0, 100.0], seed=7).numpy()
    assert np.array_equal(again.astype(np.float64), fr)


# --------------------------------------------------------------- 2D tables

GEOMETRIES_2D = [
    # name, poni, shape, n_radial, n_azimuthal, method, unit, radial_range, azimuthal_range
    ("beam_64x80_ps", dict(pixel1=172e-6, pixel2=172e-6, distance=0.05, poni1=24 * 172e-6,
                           poni2=36 * 172e-6, wavelength=1e-10), (64, 80), 32, 36,
     "pixel_split", "2th_rad", None, None),
    ("beam_64x80_full_circle", dict(pixel1=172e-6, pixel2=172e-6, distance=0.05,
                                    poni1=24 * 172e-6, poni2=36 * 172e-6, wavelength=1e-10),
     (64, 80), 24, 36, "pixel_split", "2th_deg", (0.0, 12.0), (-math.pi, math.pi)),
    ("tilted_ps", dict(pixel1=1e-4, pixel2=1.2e-4, distance=0.05, poni1=-0.002, poni2=-0.001,
                       rot1=0.01, rot2=-0.02, wavelength=1e-10), (15, 17), 20, 72,
     "pixel_split", "2th_rad", None, (-math.pi, math.pi)),
    ("counting_2d", dict(pixel1=1e-4, pixel2=1e-4, distance=0.05, poni1=4.5e-4, poni2=5.5e-4,
                         wavelength=1e-10), (9, 11), 10, 8, "counting", "2th_rad", None, None),
    ("pilatus_tilted_500x360", dict(pixel1=172e-6, pixel2=172e-6, distance=0.18,
                                    poni1=1043 * 172e-6 / 2, poni2=981 * 172e-6 / 2, rot1=0.05,
                                    rot2=-0.03, rot3=0.10, wavelength=0.3344e-10), (1043, 981),
     500, 360, "pixel_split", "2th_rad", None, None),
]


@pytest.mark.parametrize("g", GEOMETRIES_2D, ids=[g[0] for g in GEOMETRIES_2D])
def test_table_2d_bit_exact(g):
    name, poni, shape, nr, na, method, unit, rr, ar = g
    print(f"case 2d {name}", flush=True)
    plan = GpuPlan(poni, shape, method=method, unit=unit, n_radial=nr, n_azimuthal=na,
                   radial_range=rr, azimuthal_range=ar)
    lut, (r0, r1, a0, a1) = O.build_lut2d(O.Poni(**poni), shape[0], shape[1], nr, na, method,
                                          unit, rr, ar)
    assert_same_table(plan, lut)
    assert plan.info["hist_min"] == r0 and plan.info["hist_step"] == (r1 - r0) / nr
    assert np.array_equal(plan.azimuthal, a0 + (np.arange(na) + 0.5) * ((a1 - a0) / na))


def test_integrate2d_matches_oracle():
    poni = dict(pixel1=172e-6, pixel2=172e-6, distance=0.05, poni1=24 * 172e-6,
                poni2=36 * 172e-6, wavelength=1e-10)
    shape = (64, 80)
    mask = np.zeros(shape, np.uint8)
    mask[10:14, :] = 1
    plan = GpuPlan(poni, shape, n_radial=32, n_azimuthal=36, polarization_factor=0.95,
                   mask=mask, precision="f32_compensated")
    lut, _ = O.build_lut2d(O.Poni(**poni), *shape, 32, 36)
    fixed = O.fixed_table(lut, mask)
    norm = O.normalization(O.Poni(**poni), *shape, 0.95)
    den = O.denominators(fixed, norm)
    frame = np.random.default_rng(3).poisson(500.0, size=shape).astype(np.uint32)
    r = plan.integrate2d(frame, errors=True)
    i, s = O.gather(fixed, den, frame.astype(np.float64))
    mag, _ = O.gather(fixed, den, frame.astype(np.float64), False)
    assert r.intensity.shape == (36, 32)
    assert np.all(np.abs(r.intensity.ravel() - i) <= 2e-6 * np.abs(mag))
    assert np.all(np.abs(r.errors.ravel() - s) <= 2e-6 * s)


# ------------------------------------------------------------- supersampled

SS_CASES = [
    # name, poni, shape, nr, na (None = 1D), n_ss, unit, radial_range
    ("ss1d_10x10", dict(pixel1=1e-4, pixel2=1e-4, distance=0.1, poni1=0.0, poni2=0.0,
                        wavelength=1e-10), (10, 10), 7, None, 4, "2th_rad", None),
    ("ss1d_range", dict(pixel1=1e-4, pixel2=1e-4, distance=0.1, poni1=0.0, poni2=0.0,
                        wavelength=1e-10), (10, 10), 6, None, 4, "2th_rad", (0.004, 0.009)),
    ("ss1d_tilted_deg", dict(pixel1=1e-4, pixel2=1.2e-4, distance=0.05, poni1=5e-4, poni2=7e-4,
                             rot1=0.02, rot2=-0.01, wavelength=1e-10), (11, 13), 12, None, 3,
     "2th_deg", None),
    ("ss2d_tilted", dict(pixel1=1e-4, pixel2=1.2e-4, distance=0.05, poni1=5e-4, poni2=7e-4,
                         rot1=0.02, rot2=-0.01, wavelength=1e-10), (11, 13), 12, 16, 3,
     "2th_rad", None),
    ("ss2d_range", dict(pixel1=1e-4, pixel2=1.2e-4, distance=0.05, poni1=5e-4, poni2=7e-4,
                        rot1=0.02, rot2=-0.01, wavelength=1e-10), (11, 13), 12, 16, 3,
     "2th_rad", (0.001, 0.03)),
]


@pytest.mark.parametrize("g", SS_CASES, ids=[g[0] for g in SS_CASES])
def test_supersampled_table_bit_exact(g):
    name, poni, shape, nr, na, nss, unit, rr = g
    print(f"case {name}", flush=True)
    plan = GpuPlan(poni, shape, method="supersampled", n_ss=nss, unit=unit, n_radial=nr,
                   n_azimuthal=na, radial_range=rr)
    P = O.Poni(**poni)
    if na is None:
        lut = O.build_lut1d_ss(P, *shape, nr, nss, unit, rr)
    else:
        lut, _ = O.build_lut2d_ss(P, *shape, nr, na, nss, unit, rr)
    assert_same_table(plan, lut)


@pytest.mark.parametrize("ndim", [1, 2])
def test_supersampled_full_size_bit_exact(ndim):
    w = synth.WORKLOADS["eiger4m_ss2d"]
    poni = synth.workload_poni("eiger4m_ss2d")
    print(f"case eiger4m_ss2d ndim {ndim}", flush=True)
    na = w["n_azimuthal"] if ndim == 2 else None
    plan = GpuPlan(poni, w["shape"], method="supersampled", n_ss=w["n_ss"],
                   n_radial=w["n_radial"], n_azimuthal=na)
    P = O.Poni(**poni)
    if ndim == 1:
        lut = O.build_lut1d_ss(P, *w["shape"], w["n_radial"], w["n_ss"])
    else:
        lut, _ = O.build_lut2d_ss(P, *w["shape"], w["n_radial"], na, w["n_ss"])
    assert_same_table(plan, lut)
    print(f"  nnz {lut.nnz} build {plan.info['build_ms']:.1f} ms", flush=True)


# ------------------------------------------------------- pyFAI-compatible

PF_CASES = [
    # name, poni, shape, nr, na (None = 1D), unit, radial_range
    ("pf1d_beam", dict(pixel1=172e-6, pixel2=172e-6, distance=0.05, poni1=24 * 172e-6,
                       poni2=36 * 172e-6, rot1=0.02, rot2=-0.01, rot3=0.05, wavelength=1e-10),
     (64, 80), 40, None, "2th_rad", None),
    ("pf1d_range_deg", dict(pixel1=172e-6, pixel2=172e-6, distance=0.05, poni1=24 * 172e-6,
                            poni2=36 * 172e-6, wavelength=1e-10), (64, 80), 40, None, "2th_deg",
     (0.5, 10.0)),
    ("pf2d_beam", dict(pixel1=172e-6, pixel2=172e-6, distance=0.05, poni1=24 * 172e-6,
                       poni2=36 * 172e-6, rot1=0.02, rot2=-0.01, rot3=0.05, wavelength=1e-10),
     (64, 80), 32, 36, "2th_rad", None),
    ("pf2d_tilted_full", dict(pixel1=172e-6, pixel2=172e-6, distance=0.18,
                              poni1=1043 * 172e-6 / 2, poni2=981 * 172e-6 / 2, rot1=0.05,
                              rot2=-0.03, rot3=0.10, wavelength=0.3344e-10), (1043, 981), 500,
     360, "2th_rad", None),
    ("pf1d_tilted_full", dict(pixel1=172e-6, pixel2=172e-6, distance=0.18,
                              poni1=1043 * 172e-6 / 2, poni2=981 * 172e-6 / 2, rot1=0.05,
                              rot2=-0.03, rot3=0.10, wavelength=0.3344e-10), (1043, 981), 3000,
     None, "2th_rad", None),
]


@pytest.mark.parametrize("g", PF_CASES, ids=[g[0] for g in PF_CASES])
def test_pyfai_table_bit_exact(g):
    name, poni, shape, nr, na, unit, rr = g
    print(f"case {name}", flush=True)
    plan = GpuPlan(poni, shape, method="pyfai_full", unit=unit, n_radial=nr, n_azimuthal=na,
                   radial_range=rr)
    P = O.Poni(**poni)
    if na is None:
        lut = O.build_lut1d_pyfai(P, *shape, nr, unit, rr)
    else:
        lut, _ = O.build_lut2d_pyfai(P, *shape, nr, na, unit, rr)
    assert_same_table(plan, lut)


# ----------------------------------------------------------- dynamic masks

def dynamic_masks(d, n, seed):
    """Per-frame masks: the plan's mask, random 3% pixels, a masked radial
    band (bins whose denominator is 0), nothing masked, and random mixes."""
    shape = d["w"]["shape"]
    rng = np.random.default_rng(seed)
    masks = np.zeros((n,) + shape, np.uint8)
    for k in range(n):
        kind = k % 4
        if kind == 0:
            masks[k] = d["mask"]
        elif kind == 1:
            masks[k] = rng.random(shape) < 0.03
        elif kind == 2:
            masks[k, :, 300:420] = 1   # whole columns: empties bins of the inner rings
            masks[k, 400:700, :] = 1
        else:
            masks[k] = (rng.random(shape) < 0.5) if k >= 7 else 0
    return masks


def oracle_dynamic(d, frame, mask, errors=True):
    fixed = O.fixed_table(d["lut"], mask)
    den = O.denominators(fixed, d["norm"])
    return O.gather(fixed, den, frame.astype(np.float64), errors)


DYN_CASES = [(np.float32, "plan_mask", "f32_wide", 1e-10), (np.uint32, "plan_mask", "f64", 1e-10),
             (np.uint16, None, "f32_wide", 2e-6), (np.uint32, None, "f32", 2e-6),
             (np.float32, None, "f64", 1e-12), (np.uint32, None, "f32_compensated", 2e-6),
             (np.float32, None, "f32_wide", 2e-6), (np.int32, None, "f32_compensated", 2e-6)]


@pytest.mark.parametrize("dtype,plan_mask,precision,tol", DYN_CASES,
                         ids=[f"{np.dtype(t).name}-{m or 'unmasked'}-{p}"
                              for t, m, p, _ in DYN_CASES])
def test_dynamic_masks_match_oracle(pilatus, dtype, plan_mask, precision, tol):
    """Dynamic1dPlan::integrate with a mask per frame (integration/dynamic.rs:
    143-186): each frame's result is the oracle's strict gather on the table
    compacted by that frame's mask with that mask's denominators; the plan's
    own mask plays no part. A plan with a mask runs the FP64 CSR kernel on
    the unmasked table (rel 1e-10: sum order differs from the oracle's); an
    unmasked plan runs the tile kernel's raw sums of zeroed signal and zeroed
    normalization at the plan's precision (the gather tolerance of that
    precision). Bins emptied by a mask give 0."""
    d = pilatus
    print(f"case dynamic masks {np.dtype(dtype).name} {plan_mask} {precision}", flush=True)
    n = 9
    frames = np.stack([synth.frame_pilatus(s) for s in range(n)]).astype(dtype)
    masks = dynamic_masks(d, n, 31)
    plan = GpuPlan(d["poni"], d["w"]["shape"], n_radial=d["w"]["n_radial"],
                   polarization_factor=0.99, precision=precision,
                   mask=d["mask"] if plan_mask else None)
    I, S = plan.integrate_masked(frames, masks, errors=True)
    I0, S0 = plan.integrate_masked(frames, masks, errors=False)
    assert S0 is None and np.array_equal(I0, I)
    empty = 0
    for k in range(n):
        i_ref, s_ref = oracle_dynamic(d, frames[k], masks[k])
        assert np.allclose(I[k], i_ref, rtol=tol, atol=tol * 1e-3 * np.abs(i_ref).max()), k
        assert np.allclose(S[k], s_ref, rtol=tol, atol=tol * 1e-3 * np.abs(s_ref).max()), k
        z = O.denominators(O.fixed_table(d["lut"], masks[k]), d["norm"]) == 0
        assert np.all(I[k][z] == 0) and np.all(S[k][z] == 0)
        empty += int(z.sum())
    assert empty > 0   # the band masks empty some bins


@pytest.mark.parametrize("precision", ["f32_compensated", "f32_wide"])
def test_dynamic_masks_counting_table(pilatus, precision):
    """Per-frame masks on a counting (unit-weight) table of an unmasked plan:
    the tile route's masked staging against the oracle's strict gather on
    the counting table compacted by each frame's mask."""
    d = pilatus
    print(f"case dynamic masks counting {precision}", flush=True)
    w = d["w"]
    lut = oracle_lut(d["poni"], w["shape"], w["n_radial"], "counting", "2th_rad", None)
    n = 9
    frames = np.stack([synth.frame_pilatus(s) for s in range(n)]).astype(np.uint32)
    masks = dynamic_masks(d, n, 11)
    plan = GpuPlan(d["poni"], w["shape"], method="counting", n_radial=w["n_radial"],
                   polarization_factor=0.99, precision=precision)
    I, S = plan.integrate_masked(frames, masks, errors=True)
    for k in range(n):
        fixed = O.fixed_table(lut, masks[k])
        den = O.denominators(fixed, d["norm"])
        i_ref, s_ref = O.gather(fixed, den, frames[k].astype(np.float64), True)
        assert np.allclose(I[k], i_ref, rtol=2e-6, atol=2e-9 * np.abs(i_ref).max()), k
        assert np.allclose(S[k], s_ref, rtol=2e-6, atol=2e-9 * np.abs(s_ref).max()), k
        z = den == 0
        assert np.all(I[k][z] == 0) and np.all(S[k][z] == 0)


def test_dynamic_mask_equal_to_fixed_mask(pilatus):
    """A frame masked dynamically by the plan's own fixed mask gives the fixed
    path's results within the gather tolerance (fixed_table(mask) is the same
    table; the fixed path sums float32 weights)."""
    d = pilatus
    plan = GpuPlan(d["poni"], d["w"]["shape"], n_radial=d["w"]["n_radial"],
                   polarization_factor=0.99, mask=d["mask"], precision="f64")
    frames = d["frames"]
    I, S = plan.integrate_masked(frames, np.broadcast_to(d["mask"], frames.shape), errors=True)
    If, Sf, _ = plan.integrate_stack(frames, errors=True)
    assert np.allclose(I, If, rtol=2e-6, atol=0) and np.allclose(S, Sf, rtol=2e-6, atol=0)


@pytest.mark.parametrize("plan_mask", [None, "plan_mask"])
def test_dynamic_masks_batch_single_device_and_flip_bitwise(pilatus, plan_mask):
    """On the tile-kernel route (unmasked plan; 83 frames = a 64-frame chunk
    and a ragged one) and on the CSR route (plan mask; 8 frames per
    wavefront): frame k's bits do not depend on the batch it comes in, the
    device entry point gives the host entry point's bits, and flipped-row
    frames and masks on a flip-order plan give the detector-order bits."""
    from dioptrin_amd.device import DeviceArray, synchronize
    d = pilatus
    print(f"case dynamic masks bitwise {plan_mask}", flush=True)
    n = 83
    frames = np.random.default_rng(5).poisson(90.0, size=(n,) + d["w"]["shape"]).astype(
        np.uint32)
    masks = dynamic_masks(d, n, 7)
    kw = dict(n_radial=d["w"]["n_radial"], polarization_factor=0.99,
              mask=d["mask"] if plan_mask else None)
    plan = GpuPlan(d["poni"], d["w"]["shape"], **kw)
    I, S = plan.integrate_masked(frames, masks, errors=True)
    for k in (0, 3, 8, 63, 64, 82):
        i1, s1 = plan.integrate_masked(frames[k:k + 1], masks[k:k + 1], errors=True)
        assert np.array_equal(i1[0], I[k]) and np.array_equal(s1[0], S[k]), k
    fr, mk = DeviceArray.from_numpy(frames), DeviceArray.from_numpy(masks)
    oi, os_ = DeviceArray((n, plan.nbins), np.float64), DeviceArray((n, plan.nbins), np.float64)
    plan.integrate_masked_device(fr, mk, oi, os_, errors=True)
    synchronize()
    assert np.array_equal(oi.numpy(), I) and np.array_equal(os_.numpy(), S)
    flip = GpuPlan(d["poni"], d["w"]["shape"], input_order="flip_vertical", **kw)
    If, Sf = flip.integrate_masked(np.ascontiguousarray(frames[:, ::-1]),
                                   np.ascontiguousarray(masks[:, ::-1]), errors=True)
    assert np.array_equal(If, I) and np.array_equal(Sf, S)


def test_dynamic_masks_reject_unsupported(pilatus):
    """Per-frame masks take DIO_WITH_ERRORS only, on 1D plans, with masks of
    the frames' shape."""
    from dioptrin_amd import _native as N
    d = pilatus
    plan = GpuPlan(d["poni"], d["w"]["shape"], n_radial=d["w"]["n_radial"])
    frames = d["frames"][:1]
    with pytest.raises(ValueError):
        plan.integrate_masked(frames, np.zeros((2,) + d["w"]["shape"], np.uint8))
    I = np.empty((1, plan.nbins))
    m = np.zeros_like(frames, dtype=np.uint8)
    rc = N.lib().dio_integrate_masked(plan._h, frames.ctypes.data, N.DTYPES["float32"], 1,
                                      m.ctypes.data, N.WITH_AZIMUTHAL_STD, I.ctypes.data, None)
    assert rc == N.DIO_ERR_UNSUPPORTED
    g = GEOMETRIES_2D[0]
    plan2 = GpuPlan(g[1], g[2], n_radial=g[3], n_azimuthal=g[4])
    with pytest.raises(Exception):
        plan2.integrate_masked(np.zeros((1,) + g[2], np.float32),
                               np.zeros((1,) + g[2], np.uint8))


# ------------------------------------------------ azimuthal sectors (2D)

SECTOR_RANGES = [(0.0, math.pi / 2.0), (2.5, 2.5 + math.pi / 2.0), (-2.0, -1.0),
                 (0.3, 0.3 + 2.0 * math.pi)]
SECTOR_PONI = dict(pixel1=1e-4, pixel2=1e-4, distance=0.05, poni1=0.003, poni2=0.004,
                   rot1=0.01, rot2=-0.02, wavelength=1e-10)


@pytest.mark.parametrize("method", ["counting", "pixel_split", "supersampled", "pyfai_full"])
@pytest.mark.parametrize("rng", SECTOR_RANGES, ids=[f"{a:.2f}_{b:.2f}" for a, b in SECTOR_RANGES])
def test_sector_tables_bit_exact(method, rng):
    """The 60 x 80 detector with the beam inside and a slight tilt of the
    reference's azimuthal-range tests (azimuthal_range_tests.rs:8-27): 2D
    tables of 12 x 18 bins in 2theta degrees over sectors and an offset full
    circle, built on the device, equal the oracle's bit for bit."""
    shape, nr, na = (60, 80), 12, 18
    P = O.Poni(**SECTOR_PONI)
    plan = GpuPlan(SECTOR_PONI, shape, method=method, n_ss=3 if method == "supersampled" else None,
                   unit="2th_deg", n_radial=nr, n_azimuthal=na, azimuthal_range=rng)
    if method == "supersampled":
        lut, _ = O.build_lut2d_ss(P, *shape, nr, na, 3, "2th_deg", None, rng)
    elif method == "pyfai_full":
        lut, _ = O.build_lut2d_pyfai(P, *shape, nr, na, "2th_deg", None, rng)
    else:
        lut, _ = O.build_lut2d(P, *shape, nr, na, method, "2th_deg", None, rng)
    assert_same_table(plan, lut)


# ---------------------------------------------------------------- q units

Q_CASES = [
    # name, poni, shape, method, unit, nr, na, radial_range
    ("ps_qA_tilted", dict(pixel1=1e-4, pixel2=1.2e-4, distance=0.05, poni1=5e-4, poni2=7e-4,
                          rot1=0.02, rot2=-0.01, wavelength=1e-10), (23, 29), "pixel_split",
     "q_A^-1", 40, None, None),
    ("ps_qnm_beam", dict(pixel1=172e-6, pixel2=172e-6, distance=0.05, poni1=24 * 172e-6,
                         poni2=36 * 172e-6, wavelength=0.7e-10), (64, 80), "pixel_split",
     "q_nm^-1", 50, None, None),
    ("counting_qA", dict(pixel1=1.7e-4, pixel2=2.1e-4, distance=0.14, poni1=-0.003,
                         poni2=-0.002, rot1=0.007, rot2=-0.009, wavelength=1e-10), (23, 29),
     "counting", "q_A^-1", 30, None, None),
    ("ss_qnm", dict(pixel1=1e-4, pixel2=1.2e-4, distance=0.05, poni1=5e-4, poni2=7e-4,
                    rot1=0.02, rot2=-0.01, wavelength=1e-10), (11, 13), "supersampled",
     "q_nm^-1", 12, None, None),
    ("pf_qA_beam", dict(pixel1=172e-6, pixel2=172e-6, distance=0.05, poni1=24 * 172e-6,
                        poni2=36 * 172e-6, rot1=0.02, rot2=-0.01, rot3=0.05, wavelength=1e-10),
     (64, 80), "pyfai_full", "q_A^-1", 40, None, None),
    ("ps2d_qA_beam", dict(pixel1=172e-6, pixel2=172e-6, distance=0.05, poni1=24 * 172e-6,
                          poni2=36 * 172e-6, wavelength=1e-10), (64, 80), "pixel_split",
     "q_A^-1", 32, 36, None),
    ("counting2d_qnm", dict(pixel1=1e-4, pixel2=1e-4, distance=0.05, poni1=4.5e-4,
                            poni2=5.5e-4, wavelength=1e-10), (9, 11), "counting", "q_nm^-1",
     10, 8, None),
    ("ss2d_qA", dict(pixel1=1e-4, pixel2=1.2e-4, distance=0.05, poni1=5e-4, poni2=7e-4,
                     rot1=0.02, rot2=-0.01, wavelength=1e-10), (11, 13), "supersampled",
     "q_A^-1", 12, 16, None),
    ("pf2d_qnm", dict(pixel1=172e-6, pixel2=172e-6, distance=0.05, poni1=24 * 172e-6,
                      poni2=36 * 172e-6, rot1=0.02, rot2=-0.01, rot3=0.05, wavelength=1e-10),
     (64, 80), "pyfai_full", "q_nm^-1", 32, 36, None),
    ("ps_qA_range", dict(pixel1=1e-4, pixel2=1e-4, distance=0.05, poni1=4.5e-4, poni2=5.5e-4,
                         wavelength=1e-10), (9, 11), "pixel_split", "q_A^-1", 20, None,
     (0.05, 1.5)),
]


def oracle_q_table(poni, shape, method, unit, nr, na, rr):
    P = O.Poni(**poni)
    if na is None:
        if method == "supersampled":
            return O.build_lut1d_ss(P, *shape, nr, 3, unit, rr)
        if method == "pyfai_full":
            return O.build_lut1d_pyfai(P, *shape, nr, unit, rr)
        return O.build_lut1d(P, *shape, nr, method, unit, rr)
    if method == "supersampled":
        return O.build_lut2d_ss(P, *shape, nr, na, 3, unit, rr)[0]
    if method == "pyfai_full":
        return O.build_lut2d_pyfai(P, *shape, nr, na, unit, rr)[0]
    return O.build_lut2d(P, *shape, nr, na, method, unit, rr)[0]


@pytest.mark.parametrize("g", Q_CASES, ids=[g[0] for g in Q_CASES])
def test_q_unit_tables_bit_exact(g):
    """Tables in q (RadialConverter::convert, integration.rs:171-180: q =
    4 pi / (lambda 1e10 or 1e9) sin(2theta / 2)), every method, 1D and 2D:
    the device's certified sine, with glibc on the host where it is
    uncertain, gives the oracle's tables bit for bit."""
    name, poni, shape, method, unit, nr, na, rr = g
    print(f"case {name}", flush=True)
    plan = GpuPlan(poni, shape, method=method, n_ss=3 if method == "supersampled" else None,
                   unit=unit, n_radial=nr, n_azimuthal=na, radial_range=rr)
    lut = oracle_q_table(poni, shape, method, unit, nr, na, rr)
    assert_same_table(plan, lut)
    if na is None:
        assert np.array_equal(plan.radial, lut.bins())


def test_q_unit_full_size_table_and_gather():
    """The batch1d detector (3269 x 3110) at 2000 bins in q (1/nm): the
    table bit-exact and a frame's intensity within the gather tolerance."""
    w = synth.WORKLOADS["eiger9m_batch1d"]
    poni = synth.workload_poni("eiger9m_batch1d")
    print("case eiger9m_batch1d q_nm^-1", flush=True)
    plan = GpuPlan(poni, w["shape"], n_radial=w["n_radial"], unit="q_nm^-1",
                   precision="f32_wide")
    P = O.Poni(**poni)
    lut = O.build_lut1d(P, *w["shape"], w["n_radial"], "pixel_split", "q_nm^-1")
    assert_same_table(plan, lut)
    frame = synth.frames_device("eiger9m_batch1d", 1, seed=3).numpy()[0]
    I, _, _ = plan.integrate_stack(frame[None])
    den = O.denominators(lut, O.normalization(P, *w["shape"], None))
    i, _ = O.gather(lut, den, frame.astype(np.float64), False)
    assert np.all(np.abs(I[0] - i) <= 2e-6 * np.abs(i))


def test_q_units_need_a_wavelength():
    poni = dict(pixel1=1e-4, pixel2=1e-4, distance=0.05, poni1=4.5e-4, poni2=5.5e-4,
                wavelength=0.0)
    with pytest.raises(ValueError, match="wavelength"):
        GpuPlan(poni, (9, 11), unit="q_A^-1", n_radial=10)
