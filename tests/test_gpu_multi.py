"""Sums before division, explicit grids, supplied variances and
multi-detector merging on the GPU.

Ported from the reference's tests (src/multi/tests.rs):
  unequal_geometries_match_independent_counting_oracle (138-231): two
    detectors of different geometry, masks, polarization and monitor
    factors, relative and absolute normalization, against an independent
    per-pixel binning;
  identical_detectors_preserve_intensity_and_reduce_std (96-136);
  explicit_variances_follow_squared_weights_and_detector_normalization
    (885-998);
  split_detector_merges_to_the_whole_detector_pattern_in_1d (1002-1052);
  pixels_outside_the_radial_range_do_not_leak_into_edge_bins (1054-1111).
These run at FP64 precision with the reference's tolerances (2e-11, 1e-9,
1e-12 relative). Raw sums on an explicit sector grid are checked against the
oracle's table and sums at f32_compensated (2e-6 of the magnitude) and
FP64 (1e-12); supplied variances equal to max(x, 0) give the Poisson
path's bits.
"""
import math

import numpy as np
import pytest

import oracle as O
from dioptrin_amd import (GpuPlan, IntegrationGrid2d, Integrator, MultiIntegrator)

pytestmark = pytest.mark.gpu

PI, TAU = math.pi, 2.0 * math.pi
METHODS = [("counting", None), ("pixel_split", None), ("supersampled", 4)]


def geometry(**kw):
    p = dict(pixel1=0.0001, pixel2=0.00012, distance=0.1, poni1=0.0007, poni2=0.0011,
             wavelength=1e-10)
    p.update(kw)
    return p


def ai(poni, method, n_ss=None, radial_range=(0.0, 0.1), **kw):
    return Integrator(poni, method=method, n_ss=n_ss, radial_range=radial_range,
                      precision="f64", **kw)


def close(a, b, tol=2e-11):
    a, b = np.ravel(a), np.ravel(b)
    assert a.shape == b.shape
    scale = np.maximum(np.maximum(np.abs(a), np.abs(b)), 1.0)
    bad = np.abs(a - b) > tol * scale
    assert not bad.any(), (np.flatnonzero(bad)[:5], a[bad][:5], b[bad][:5])


def test_unequal_geometries_match_independent_counting_oracle():
    g1 = geometry()
    g2 = geometry(pixel1=0.0002, pixel2=0.00012 * 1.7, distance=0.14, rot1=0.012, rot2=-0.008,
                  rot3=0.13)
    shapes = [(7, 11), (5, 9)]
    data = [np.full(shapes[0], 13.0), np.full(shapes[1], 47.0)]
    masks = [(np.arange(r * c) % 7 == 0).astype(np.uint8).reshape(r, c) for r, c in shapes]
    for norm in ("relative", "absolute"):
        its = []
        for i, (g, m) in enumerate(zip((g1, g2), masks)):
            it = ai(g, "counting", polarization_factor=0.3 + i * 0.2)
            it.set_mask(m)
            its.append(it)
        out = MultiIntegrator(its).integrate2d(data, 8, 13, errors=True, normalization=norm,
                                               normalization_factors=[2.0, 3.0], raw=True)
        signal, denom, variance = np.zeros(104), np.zeros(104), np.zeros(104)
        for i, (g, (r, c)) in enumerate(zip((g1, g2), shapes)):
            P = O.Poni(**g)
            tth, chi = O.centres(P, r, c)
            sa = O.solid_angle(P, r, c)
            pol = O.polarization(0.3 + i * 0.2, tth, chi)
            scale = (i + 2.0) * (g["pixel1"] * g["pixel2"] / g["distance"] ** 2
                                 if norm == "absolute" else 1.0)
            x = data[i].ravel()
            for p in range(r * c):
                if masks[i].ravel()[p]:
                    continue
                # Independent half-open bins, not the builder's rule.
                rb = math.floor(tth[p] / 0.1 * 8.0)
                ab = math.floor(((chi[p] + PI) % TAU) / TAU * 13.0)
                if rb >= 8 or ab >= 13:
                    continue
                b = ab * 8 + rb
                signal[b] += x[p]
                variance[b] += x[p]
                denom[b] += sa[p] * pol[p] * scale
        close(out.raw["sum_signal"], signal)
        close(out.raw["sum_norm"], denom)
        close(out.raw["sum_variance"], variance)
        close(out.intensity, np.divide(signal, denom, out=np.zeros(104), where=denom != 0))
        close(out.errors, np.divide(np.sqrt(variance), denom, out=np.zeros(104),
                                    where=denom != 0))
        assert (out.raw["count"] == 0).any()
        assert out.intensity.shape == (13, 8)


@pytest.mark.parametrize("method,n_ss", METHODS)
def test_identical_detectors_preserve_intensity_and_reduce_std(method, n_ss):
    g = geometry()
    data = np.full((9, 12), 81.0)
    one = MultiIntegrator([ai(g, method, n_ss)])
    two = MultiIntegrator([ai(g, method, n_ss), ai(g, method, n_ss)])
    a = one.integrate2d([data], 9, 16, errors=True, raw=True)
    b = two.integrate2d([data, data], 9, 16, errors=True, raw=True)
    close(a.intensity, b.intensity)
    close(a.errors / math.sqrt(2.0), b.errors)
    close(2.0 * a.raw["count"], b.raw["count"])


@pytest.mark.parametrize("method,n_ss", METHODS)
def test_single_detector_equals_integrate2d_on_the_same_grid(method, n_ss):
    g = geometry(poni1=-0.001, poni2=-0.001)
    data = (10.0 + np.arange(13 * 19) % 23).reshape(13, 19)
    it = ai(g, method, n_ss, radial_range=(0.0, 0.12))
    out = MultiIntegrator([it]).integrate2d([data], 11, 17, errors=True)
    ref = it.integrate2d(data, 11, 17, errors=True, azimuthal_range=(-PI, PI))
    # Same sums, one division: identical intensity; sigma sqrt(v)/n against
    # sqrt(v/n^2) within an ulp or two.
    assert np.array_equal(out.intensity, ref.intensity)
    close(out.errors, ref.errors, 1e-15)


@pytest.mark.parametrize("method,n_ss", METHODS)
def test_explicit_variances_follow_squared_weights(method, n_ss):
    g = geometry()
    data = np.full((9, 12), 25.0)
    var = np.full((9, 12), 25.0)
    multi = MultiIntegrator([ai(g, method, n_ss), ai(g, method, n_ss)])
    kw = dict(errors=True, normalization="absolute")
    poisson = multi.integrate2d([data, data], 17, 11, **kw)
    explicit = multi.integrate2d([data, data], 17, 11, variances=[var, var], **kw)
    close(explicit.intensity, poisson.intensity)
    close(explicit.errors, poisson.errors)
    # A gain of -3: intensity times -3, std times 3.
    changed = multi.integrate2d([np.full((9, 12), -75.0)] * 2, 17, 11,
                                variances=[np.full((9, 12), 225.0)] * 2, **kw)
    close(changed.intensity, -3.0 * poisson.intensity)
    close(changed.errors, 3.0 * poisson.errors)
    bad = var.copy()
    bad[0, 2] = -1.0
    with pytest.raises(ValueError):
        multi.integrate2d([data, data], 17, 11, variances=[var, bad], **kw)
    bad[0, 2] = np.nan
    with pytest.raises(ValueError):
        multi.integrate2d([data, data], 17, 11, variances=[var, bad], **kw)
    with pytest.raises(ValueError):
        multi.integrate2d([data, data], 17, 11, variances=[var[:3], var], **kw)
    with pytest.raises(ValueError):
        multi.integrate2d([data, data], 17, 11, variances=[var, var], errors=False)


@pytest.mark.parametrize("method,n_ss", METHODS)
def test_split_detector_merges_to_the_whole_detector_pattern_in_1d(method, n_ss):
    rows, cols, cut = 20, 20, 3
    whole = geometry(poni1=-0.001, poni2=-0.001)
    bottom = dict(whole, poni1=whole["poni1"] - cut * whole["pixel1"])
    i = np.arange(rows * cols)
    data = (10.0 + (i * 7 % 31) + 40.0 * (i // cols < cut)).reshape(rows, cols)
    rr = (0.012, 0.046)
    exp = ai(whole, method, n_ss, radial_range=rr).integrate1d(data, 40, errors=True)
    multi = MultiIntegrator([ai(whole, method, n_ss, radial_range=rr),
                             ai(bottom, method, n_ss, radial_range=rr)])
    merged = multi.integrate1d([data[:cut], data[cut:]], 40)
    merged_e = multi.integrate1d([data[:cut], data[cut:]], 40, errors=True)
    assert np.count_nonzero(exp.intensity > 0) > 20
    close(merged.intensity, exp.intensity, 1e-9)
    close(merged_e.intensity, exp.intensity, 1e-9)
    close(merged_e.errors, exp.errors, 1e-9)
    # batch1d: two frame sets, each the merged pattern of its own images.
    sets = [[data[:cut], data[cut:]], [2.0 * data[:cut], 2.0 * data[cut:]]]
    res = multi.batch1d(sets, 40, errors=True)
    assert np.array_equal(res[0].intensity, merged_e.intensity)
    close(res[1].intensity, 2.0 * merged_e.intensity, 1e-12)
    with pytest.raises(ValueError, match="Frame set 0"):
        multi.batch1d([[data]], 40)


@pytest.mark.parametrize("method,n_ss", METHODS + [("pyfai_full", None)])
def test_pixels_outside_the_radial_range_do_not_leak(method, n_ss):
    rows, cols, n = 20, 20, 12
    g = geometry(poni1=-0.001, poni2=-0.001)
    outside = dict(g, poni1=-0.006)
    lo, hi = 0.018, 0.034
    P = O.Poni(**g)
    tth, _ = O.corners(P, rows, cols)
    t = tth.reshape(rows + 1, cols + 1)
    corners = np.stack([t[:-1, :-1], t[:-1, 1:], t[1:, :-1], t[1:, 1:]]).reshape(4, -1)
    out_of_range = np.all(corners < lo, axis=0) | np.all(corners > hi, axis=0)
    assert out_of_range.sum() > 50 and (~out_of_range).sum() > 50
    to, _ = O.corners(O.Poni(**outside), rows, cols)
    assert np.all(to > hi)
    clean = (10.0 + np.arange(rows * cols) % 13).reshape(rows, cols)
    dirty = np.where(out_of_range.reshape(rows, cols), 1e6, clean)
    bright = np.full((rows, cols), 1e6)
    ranged = lambda geo: ai(geo, method, n_ss, radial_range=(lo, hi))
    exp = ranged(g).integrate1d(clean, n, errors=True)
    assert np.all(exp.intensity > 0)
    single = ranged(g).integrate1d(dirty, n, errors=True)
    merged = MultiIntegrator([ranged(g), ranged(outside)]).integrate1d([clean, bright], n,
                                                                       errors=True)
    for got, want in [(single.intensity, exp.intensity), (single.errors, exp.errors),
                      (merged.intensity, exp.intensity), (merged.errors, exp.errors)]:
        assert np.all(np.abs(got - want) <= 1e-12 * np.abs(want))


PILATUS_SECTOR = dict(pixel1=172e-6, pixel2=172e-6, distance=0.12, poni1=60 * 172e-6,
                      poni2=50 * 172e-6, rot1=0.05, rot2=-0.03, wavelength=1e-10)


@pytest.mark.parametrize("precision,tol", [("f32_compensated", 2e-6), ("f64", 1e-12)])
@pytest.mark.parametrize("method", ["counting", "pixel_split", "pyfai_full"])
def test_raw_on_sector_grid_matches_oracle(method, precision, tol):
    shape = (120, 100)
    rng = np.random.default_rng(5)
    img = rng.poisson(300.0, size=shape).astype(np.float64)
    var = rng.uniform(0.0, 50.0, size=shape)
    grid = IntegrationGrid2d(30, 8, (0.01, 0.12), (0.3, 0.3 + PI / 2))
    it = Integrator(PILATUS_SECTOR, method=method, precision=precision, polarization_factor=0.9)
    r = it.integrate2d_raw_on_grid(img, grid, errors=True)
    rv = it.integrate2d_raw_on_grid_with_variances(img, var, grid)
    P = O.Poni(**PILATUS_SECTOR)
    build = O.build_lut2d_pyfai if method == "pyfai_full" else \
        (lambda *a, **k: O.build_lut2d(*a, method=method, **k))
    lut, _ = build(P, *shape, 30, 8, radial_range=(0.01, 0.12),
                   azimuthal_range=(0.3, 0.3 + PI / 2))
    plan = it.grid_plan(shape, grid)
    rp, col, _, _ = plan.export_csr()
    assert np.array_equal(rp, lut.row_ptr) and np.array_equal(col, lut.col)
    w = np.ones(lut.nnz) if lut.w is None else lut.w
    sq = O.Lut(lut.row_ptr, lut.col, w * w, lut.hist_min, lut.hist_step)
    sums = lambda L, x: O.denominators(L, np.ascontiguousarray(x, np.float64).ravel())
    # FP64 denominators of the device normalization, bit for bit.
    assert np.array_equal(r.sum_norm, sums(lut, plan.normalization()))
    assert np.array_equal(r.count, sums(lut, np.ones(img.size)))
    s_ref = sums(lut, img)
    v_ref = sums(sq, np.maximum(img, 0.0))
    vv_ref = sums(sq, var)
    mag = sums(lut, np.abs(img))
    assert np.all(np.abs(r.sum_signal - s_ref) <= tol * mag)
    assert np.all(np.abs(r.sum_variance - v_ref) <= tol * v_ref)
    assert np.array_equal(rv.sum_signal, r.sum_signal)
    assert np.all(np.abs(rv.sum_variance - vv_ref) <= tol * vv_ref)
    assert np.count_nonzero(s_ref) > 100 and np.count_nonzero(s_ref == 0) > 0
    # The same table as integrate2d on that azimuthal range with the radial range.
    it2 = Integrator(PILATUS_SECTOR, method=method, precision=precision,
                     polarization_factor=0.9, radial_range=(0.01, 0.12))
    plan = it2._plan(shape, 30, 8, (0.3, 0.3 + PI / 2))
    assert np.array_equal(plan.integrate_raw(img[None], errors=True)["sum_signal"][0],
                          r.sum_signal)


@pytest.mark.parametrize("dtype", [np.uint32, np.float32])
def test_supplied_poisson_variances_give_the_poisson_bits(dtype):
    shape = (104, 96)
    P = dict(pixel1=75e-6, pixel2=75e-6, distance=0.03, poni1=52 * 75e-6, poni2=47 * 75e-6,
             rot1=0.02, wavelength=0.5e-10)
    rng = np.random.default_rng(8)
    frames = rng.poisson(150.0, size=(5, *shape)).astype(dtype)
    for nr, na in [(120, None), (40, 18)]:
        plan = GpuPlan(P, shape, n_radial=nr, n_azimuthal=na, precision="f32_compensated")
        I, S, _ = plan.integrate_stack(frames, errors=True)
        var = np.maximum(frames.astype(np.float32), 0.0)
        Iv, Sv = plan.integrate_with_variances(frames, var)
        assert np.array_equal(Iv, I) and np.array_equal(Sv, S)
        raw = plan.integrate_raw(frames, errors=True)
        rs, rvar = plan.integrate_with_variances(frames, var, raw=True)
        assert np.array_equal(rs, raw["sum_signal"]) and np.array_equal(rvar, raw["sum_variance"])
    bad = var.copy()
    bad[1, 3, 3] = -1.0
    with pytest.raises(ValueError):
        plan.integrate_with_variances(frames, bad)
