"""Pins the FP64 oracle against the reference's golden vectors, its in-source
known answers and closed forms (CPU only).

Sources (reference repository paths):
  geometry vectors   test_data/test_reference/*_pyFAI.txt, src/geometry.rs:789-920
  counting rules     src/lut/counting_csr.rs:261-319
  pixel splitting    src/lut/pixel_split.rs:1014-1190
  azimuthal std      src/lut/azimuthal.rs:134-251
  gather             src/lut/prepared.rs:985-1121 (normalize_bin), fast_tests.rs
"""
import math
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "pyfai_v1")
V1 = O.Poni(0.000172, 0.000172, 0.208651380603, 0.0872948482846, 0.0799601126306,
            -0.0184422457059, 0.00413760084465, -2.77645988275e-08, 4.066e-11)


def _gold(name):
    return np.loadtxt(os.path.join(GOLD, name))


# ------------------------------------------------------------------ geometry

def test_positions_match_pyfai_vectors():
    d1, d2 = np.meshgrid(np.arange(10) + 0.5, np.arange(20) + 0.5, indexing="ij")
    x, y, z = O.positions(V1, d1.ravel(), d2.ravel())
    for name, v in (("x", x), ("y", y), ("z", z)):
        assert np.abs(v - _gold(f"{name}_pyFAI.txt")).max() < 1e-10


def test_angles_and_corrections_match_pyfai_vectors():
    tth, chi = O.centres(V1, 10, 20)
    assert np.abs(tth - _gold("tth_pyFAI.txt")).max() < 1e-6
    assert np.abs(chi - _gold("azimuth_pyFAI.txt")).max() < 1e-6
    sa = O.solid_angle(V1, 10, 20)
    assert np.abs(sa - _gold("solid_angle_pyFAI.txt")).max() < 1e-6
    pol = O.polarization(0.99, tth, chi)
    assert np.abs(pol - _gold("polarization_pyFAI.txt")).max() < 1e-6


def test_poni_reference_positions():
    # geometry.rs:749-760: d1_cart_poni[0] == -0.0319 exactly.
    p = O.Poni(0.0002, 0.0003, 0.0, 0.032, 0.034)
    assert 0.5 * p.pixel1 - p.poni1 == -0.0319
    d = _gold("d1_cart_poni.txt")
    ours = ((np.repeat(np.arange(10), 20) + 0.5) * p.pixel1 - p.poni1)
    assert np.abs(ours - d).max() < 1e-10


def test_on_axis_tth_closed_form():
    # Untilted detector: tth of a point at radius r is atan(r / d).
    p = O.Poni(1e-4, 1e-4, 0.1, 0.0, 0.0)
    d1 = np.array([3.0, 0.0, 10.0, 7.5])
    d2 = np.array([4.0, 10.0, 0.0, 2.5])
    tth, chi = O.tth_chi(p, d1, d2)
    r = np.hypot(d1 * 1e-4, d2 * 1e-4)
    assert np.allclose(tth, np.arctan(r / 0.1), rtol=0, atol=1e-15)
    assert np.allclose(chi, np.arctan2(d1, d2), rtol=0, atol=1e-15)


# ------------------------------------------------------------------ counting

@pytest.mark.parametrize("hi", [3.0, 30.8, 60.0, 1234.5])
def test_counting_automatic_range_keeps_every_pixel(hi):
    radial = 1.0 + (hi - 1.0) * np.arange(5000) / 4999.0
    for n in (7, 100, 1000, 1500, 3000):
        lut = O.counting_from_radial(radial, n)
        assert lut.nnz == radial.size
        assert 4999 in lut.col[lut.row_ptr[n - 1]:]


@pytest.mark.parametrize("hi", [3.0, 30.8, 60.0, 1234.5])
def test_counting_explicit_range_includes_both_limits(hi):
    radial = np.array([1.0, hi, hi + hi * 1e-12, 1.0 - 1e-12])
    for n in (7, 100, 1000, 1500, 3000):
        lut = O.counting_from_radial(radial, n, (1.0, hi))
        assert list(lut.col) == [0, 1]
        assert lut.row_ptr[1] == 1 and lut.row_ptr[n - 1] == 1


def test_counting_explicit_range_drops_pixels_outside():
    radial = np.array([0.5, 1.5, 2.5, 3.5, 1.0, 3.0, -4.0, np.inf])
    lut = O.counting_from_radial(radial, 2, (1.0, 3.0))
    assert list(lut.row_ptr) == [0, 2, 4]
    assert list(lut.col) == [1, 4, 2, 5]


# ----------------------------------------------------------- pixel splitting

def radial_strip(n, offset, distance):
    return O.Poni(1e-6, 1e-4, distance, 0.5e-6, -offset * 1e-4, wavelength=1e-10)


def pixel_extent(p, dim1, dim2, pixel):
    t, _ = O.corners(p, dim1, dim2)
    w = dim2 + 1
    b = pixel // dim2 * w + pixel % dim2
    v = t[[b, b + 1, b + w + 1, b + w]]
    return v.min(), v.max()


def pixel_totals(lut, npix):
    return np.bincount(lut.col, weights=lut.w, minlength=npix)


@pytest.mark.parametrize("offset,distance", [(300.0, 1.0), (3000.0, 10.0)])
def test_split_fractions_at_any_radius(offset, distance):
    p = radial_strip(1, offset, distance)
    lo, hi = pixel_extent(p, 1, 1, 0)
    step = hi - lo
    lut = O.ps_fractions(p, 1, 1, 2, lo - 0.7 * step, step)
    assert list(np.diff(lut.row_ptr)) == [1, 1]
    w0, w1 = lut.w
    assert abs(w0 - 0.3) < 1e-3 and abs(w1 - 0.7) < 1e-3
    assert abs(w0 + w1 - 1.0) < 1e-12


def test_fully_included_pixels_have_unit_total_weight():
    p = O.Poni(1.7e-4, 2.1e-4, 0.14, -0.003, -0.002, 0.007, -0.009, 0.0, 1e-10)
    t, _ = O.corners(p, 23, 29)
    lo, hi = t.min(), t.max() + 10 * np.finfo(float).eps
    lut = O.ps_fractions(p, 23, 29, 40, lo, (hi - lo) / 40)
    assert np.abs(pixel_totals(lut, 23 * 29) - 1.0).max() < 1e-9


def test_range_clipped_pixel_keeps_in_range_fraction():
    p = radial_strip(1, 300.0, 1.0)
    lo, hi = pixel_extent(p, 1, 1, 0)
    mx = lo + 0.4 * (hi - lo)
    mn = lo - (mx - lo)
    lut = O.ps_fractions(p, 1, 1, 4, mn, (mx - mn) / 4)
    assert abs(lut.w.sum() - 0.4) < 1e-3
    lut = O.build_lut1d(p, 1, 1, 4, radial_range=(mn, mx))
    assert np.all(lut.col == 0)
    assert lut.row_ptr[4] - lut.row_ptr[3] == 1
    assert abs(lut.w.sum() - 0.4) < 1e-3


def test_bin_weights_follow_pixel_fractions():
    p = radial_strip(5, 10.0, 0.1)
    lo1, hi1 = pixel_extent(p, 1, 5, 1)
    lo3, hi3 = pixel_extent(p, 1, 5, 3)
    lut = O.build_lut1d(p, 1, 5, 1, radial_range=(0.5 * (lo1 + hi1), 0.5 * (lo3 + hi3)))
    assert list(lut.col) == [1, 2, 3]
    assert abs(lut.w[1] - 1.0) < 1e-12
    for v, e in zip(lut.w / lut.w.sum(), (0.25, 0.5, 0.25)):
        assert abs(v - e) < 0.01


def test_beam_pixel_uses_detector_plane_fractions():
    p = O.Poni(1e-4, 1e-4, 0.1, 4.3e-4, 5.6e-4, wavelength=1e-10)
    t, _ = O.corners(p, 9, 11)
    hi = t.max() + 10 * np.finfo(float).eps
    n = 200
    step = hi / n
    lut = O.ps_fractions(p, 9, 11, n, 0.0, step)
    assert np.abs(pixel_totals(lut, 99) - 1.0).max() < 1e-9
    radius = 0.1 * math.tan(step)
    expected = math.pi * radius * radius / 1e-8
    assert lut.row_ptr[1] == 1
    assert lut.col[0] == 4 * 11 + 5
    assert abs(lut.w[0] / expected - 1.0) < 0.05


@pytest.mark.parametrize("poni1,poni2", [(4.5e-4, 5.5e-4), (4.3e-4, 5.6e-4)])
def test_automatic_range_conserves_beam_pixel(poni1, poni2):
    p = O.Poni(1e-4, 1e-4, 0.05, poni1, poni2, wavelength=1e-10)
    lut = O.build_lut1d(p, 9, 11, 200, unit="2th_deg")
    tot = pixel_totals(lut, 99)
    assert abs(tot[4 * 11 + 5] - 1.0) < 1e-12
    assert np.abs(tot - 1.0).max() < 1e-9


def test_explicit_range_keeps_outer_part_of_beam_pixel():
    p = O.Poni(1e-4, 1e-4, 0.05, 4.5e-4, 5.5e-4, wavelength=1e-10)
    tth, _ = O.centres(p, 9, 11)
    radial = np.degrees(tth)
    mn = math.degrees(math.atan(0.5e-4 / 0.05))
    lut = O.build_lut1d(p, 9, 11, 200, unit="2th_deg", radial_range=(mn, radial.max() * 2))
    tot = pixel_totals(lut, 99)
    assert abs(tot[4 * 11 + 5] - (1 - math.pi / 4)) < 2e-3


def _strip_area(poly, b0, b1):
    """Area of a convex polygon inside b0 <= x <= b1, by exact integration of
    its vertical chord length (piecewise linear in x)."""
    xs = sorted({float(x) for x, _ in poly} | {b0, b1})
    xs = [x for x in xs if b0 <= x <= b1 and min(poly[:, 0]) <= x <= max(poly[:, 0])]

    def chord(x):
        ys = []
        n = len(poly)
        for i in range(n):
            (x0, y0), (x1, y1) = poly[i], poly[(i + 1) % n]
            if (x0 - x) * (x1 - x) <= 0 and x0 != x1:
                ys.append(y0 + (x - x0) / (x1 - x0) * (y1 - y0))
            elif x0 == x1 == x:
                ys += [y0, y1]
        return (max(ys) - min(ys)) if ys else 0.0

    return sum(0.5 * (chord(a) + chord(b)) * (b - a) for a, b in zip(xs, xs[1:]))


def test_split_fractions_equal_exact_strip_areas():
    # Independent method: the fraction of a pixel in a bin is the integral of
    # the pixel's chord length over the bin's radial interval.
    p = O.Poni(1e-4, 1e-4, 0.05, -0.002, -0.001, wavelength=1e-10)
    dim1, dim2, n = 12, 15, 25
    t, a = O.corners(p, dim1, dim2)
    lut = O.build_lut1d(p, dim1, dim2, n)
    _, chi_c = O.centres(p, dim1, dim2)
    w = dim2 + 1
    rows = np.repeat(np.arange(n), np.diff(lut.row_ptr).astype(int))
    checked = 0
    for k in range(lut.nnz):
        pix = int(lut.col[k])
        b = pix // dim2 * w + pix % dim2
        idx = [b, b + 1, b + w + 1, b + w]
        poly = np.array([[t[i], a[i]] for i in idx])
        poly[:, 1] = np.where(poly[:, 1] - chi_c[pix] > math.pi, poly[:, 1] - 2 * math.pi,
                              poly[:, 1])
        total = _strip_area(poly, -1.0, 10.0)
        b0 = lut.hist_min + rows[k] * lut.hist_step
        frac = _strip_area(poly, b0, b0 + lut.hist_step) / total
        assert abs(frac - lut.w[k]) < 1e-9, (k, frac, lut.w[k])
        checked += 1
    assert checked == lut.nnz > 0


# -------------------------------------------------------------- gather / std

def _one_bin(w=None):
    return O.Lut(np.array([0, 3], np.uint64), np.array([0, 1, 2], np.uint64),
                 None if w is None else np.array(w, float), 0.0, 1.0)


def test_std_corrected_weighted_population():
    lut = _one_bin([0.5, 0.25, 1.0])
    i, s, std = O.azimuthal_std(lut, None, [2.0, 4.0, 2.0], [4.0, 16.0, 16.0], errors=True)
    assert i[0] == 5.5
    assert abs(std[0] - math.sqrt(6.75)) < 1e-12
    assert abs(s[0] - math.sqrt(18.0) / 4.0) < 1e-12


def test_std_stable_at_large_offset():
    lut = _one_bin()
    _, _, std = O.azimuthal_std(lut, None, [1.0] * 3, [1e12 - 1, 1e12, 1e12 + 1])
    assert abs(std[0] - math.sqrt(2 / 3)) < 1e-12
    _, _, std = O.azimuthal_std(lut, None, [1.0] * 3, [6.0, 12.0, 24.0])
    assert abs(std[0] - math.sqrt(56.0)) < 1e-12


def test_std_masks_invalid_and_empty():
    lut = _one_bin()
    i, _, std = O.azimuthal_std(lut, [0, 1, 0], [2.0] * 3, [10.0, 1e6, np.nan])
    assert i[0] == 5.0 and std[0] == 0.0
    i, s, std = O.azimuthal_std(lut, [1, 1, 1], [1.0] * 3, [10.0] * 3, errors=True)
    assert i[0] == 0.0 and s[0] == 0.0 and np.isnan(std[0])
    _, _, std = O.azimuthal_std(lut, None, [0.0, -1.0, np.inf], [10.0] * 3)
    assert np.isnan(std[0])


def test_std_constant_signal_has_zero_spread():
    i, _, std = O.azimuthal_std(_one_bin([0.2, 0.7, 1.0]), None, [1.0, 2.0, 4.0],
                                [20.0, 40.0, 80.0])
    assert i[0] == 20.0 and std[0] == 0.0


def test_gather_closed_form_and_zero_denominator():
    lut = O.Lut(np.array([0, 2, 2, 3], np.uint64), np.array([0, 1, 2], np.uint64),
                np.array([0.5, 0.25, 1.0]), 0.0, 1.0)
    norm = np.array([2.0, 4.0, 0.0])
    den = O.denominators(lut, norm)
    assert list(den) == [0.5 * 2 + 0.25 * 4, 0.0, 0.0]
    i, s = O.gather(lut, den, np.array([4.0, -8.0, 3.0]))
    assert i[0] == (0.5 * 4 + 0.25 * -8) / 2.0
    assert s[0] == math.sqrt(0.25 * 4 / 4.0)
    assert i[1] == 0.0 and i[2] == 0.0 and s[2] == 0.0


def test_masked_table_keeps_order():
    lut = O.Lut(np.array([0, 3, 5], np.uint64), np.array([4, 1, 2, 0, 3], np.uint64),
                np.arange(5, dtype=float), 0.0, 1.0)
    f = O.fixed_table(lut, np.array([0, 1, 0, 1, 0], np.uint8))
    assert list(f.row_ptr) == [0, 2, 3]
    assert list(f.col) == [4, 2, 0]
    assert list(f.w) == [0.0, 2.0, 3.0]


# ---------------------------------------------------------------- 2D tables

BEAM = O.Poni(172e-6, 172e-6, 0.05, 24 * 172e-6, 36 * 172e-6, wavelength=1e-10)


def test_2d_full_circle_conserves_every_pixel():
    # Explicit edges covering all pixels and the full circle: every pixel's
    # fractions add up to one (shared_grid.rs:307-354 conservation).
    t, _ = O.corners(BEAM, 64, 80)
    lut, _ = O.build_lut2d(BEAM, 64, 80, 24, 36, "pixel_split",
                           radial_range=(0.0, t.max() * 1.01),
                           azimuthal_range=(-math.pi, math.pi))
    tot = np.bincount(lut.col, weights=lut.w, minlength=64 * 80)
    assert np.abs(tot - 1.0).max() < 1e-9


def test_2d_summed_over_azimuth_equals_1d():
    # Same radial edges (corner range + 10 eps) and the full circle: summing a
    # pixel's 2D fractions over azimuth gives its 1D fraction per radial bin.
    p = O.Poni(1e-4, 1.2e-4, 0.05, -0.002, -0.001, 0.01, -0.02, 0.0, 1e-10)
    dim1, dim2, nr = 15, 17, 20
    l1 = O.build_lut1d(p, dim1, dim2, nr)
    l2, edges = O.build_lut2d(p, dim1, dim2, nr, 72, "pixel_split",
                              azimuthal_range=(-math.pi, math.pi))
    assert edges[0] == l1.hist_min
    r1 = np.repeat(np.arange(nr), np.diff(l1.row_ptr).astype(int))
    rows2 = np.repeat(np.arange(nr * 72), np.diff(l2.row_ptr).astype(int)) % nr
    a = np.zeros((dim1 * dim2, nr))
    b = np.zeros((dim1 * dim2, nr))
    np.add.at(a, (l1.col.astype(int), r1), l1.w)
    np.add.at(b, (l2.col.astype(int), rows2), l2.w)
    assert np.abs(a - b).max() < 1e-9


def test_2d_counting_matches_brute_force_bins():
    p = O.Poni(1e-4, 1e-4, 0.05, 4.5e-4, 5.5e-4, wavelength=1e-10)
    lut, (r0, r1, a0, a1) = O.build_lut2d(p, 9, 11, 10, 8, "counting")
    tth, chi = O.centres(p, 9, 11)
    assert r0 == tth.min() and r1 == tth.max() + 10 * np.finfo(float).eps
    assert a0 == chi.min() and a1 == chi.max() + 10 * np.finfo(float).eps
    ir = np.minimum(((tth - r0) / ((r1 - r0) / 10)).astype(int), 9)
    ia = np.minimum(((chi - a0) / ((a1 - a0) / 8)).astype(int), 7)
    expect = ia * 10 + ir
    rows = np.repeat(np.arange(80), np.diff(lut.row_ptr).astype(int))
    got = np.empty(99, int)
    got[lut.col.astype(int)] = rows
    assert lut.nnz == 99 and np.array_equal(got, expect)


# ------------------------------------------------------------- supersampled
# supersampled.rs:1545-1552 test geometry and 1737-1790 known answers.

SS_GEO = O.Poni(1e-4, 1e-4, 0.1, 0.0, 0.0, wavelength=1e-10)


def test_ss_weights_1d_are_fractions_of_the_pixel():
    lut = O.build_lut1d_ss(SS_GEO, 10, 10, 7, 4)
    tot = np.zeros(100)
    for c, w in zip(lut.col, lut.w):  # sequential sums, as the reference test adds
        tot[c] += w
    assert np.all(tot == 1.0)


def test_ss_explicit_range_drops_subpixels_outside_it():
    n_ss, n_points, lo, hi = 4, 6, 0.004, 0.009
    lut = O.build_lut1d_ss(SS_GEO, 10, 10, n_points, n_ss, radial_range=(lo, hi))
    expected = np.zeros(100)
    expected_bins = np.zeros(n_points, int)
    for pixel in range(100):
        row, col = divmod(pixel, 10)
        d1 = np.array([row + (i + 0.5) / n_ss for i in range(n_ss) for _ in range(n_ss)])
        d2 = np.array([((col * n_ss + j) + 0.5) / n_ss for _ in range(n_ss) for j in range(n_ss)])
        r, _ = O.tth_chi(SS_GEO, d1, d2)
        for v in r:
            if lo <= v <= hi:
                expected[pixel] += 1.0 / (n_ss * n_ss)
                expected_bins[int(((v - lo) / (hi - lo)) * n_points)] += 1
    tot = np.zeros(100)
    for c, w in zip(lut.col, lut.w):
        tot[c] += w
    assert np.array_equal(tot, expected)
    assert 1.0 in tot and 0.0 in tot and np.any((tot > 0) & (tot < 1))
    for b in (0, n_points - 1):
        assert lut.w[lut.row_ptr[b]:lut.row_ptr[b + 1]].sum() == expected_bins[b] / n_ss ** 2


def test_ss_2d_totals_and_1d_consistency():
    p = O.Poni(1e-4, 1.2e-4, 0.05, 5e-4, 7e-4, 0.02, -0.01, 0.0, 1e-10)
    l2, (r0, r1, a0, a1) = O.build_lut2d_ss(p, 11, 13, 12, 16, 3)
    tot = np.bincount(l2.col, weights=l2.w, minlength=143)
    assert np.abs(tot - 1.0).max() < 1e-12   # clamped default grid keeps every subpixel
    l1 = O.build_lut1d_ss(p, 11, 13, 12, 3)
    assert l1.hist_min == r0 and abs(l1.hist_step - (r1 - r0) / 12) <= 1e-18


# ------------------------------------------------------------ pyFAI split

def test_pyfai_split1d_matches_pyfai_golden_weights():
    # pyfai_1d.rs:391-413: pyFAI 2026.9.0 fullSplit1D_engine on one pixel.
    got = O.pyfai_split1d([(0.1, 0.2), (1.2, 0.3), (1.8, 1.5), (0.4, 1.7)], 0.0, 2.0, 4)
    expected = [0.21462747684427036, 0.4066985688044358, 0.32228298342690204,
                0.05639097092439181]
    assert [b for b, _ in got] == [0, 1, 2, 3]
    for (_, w), e in zip(got, expected):
        assert abs(w - e) < 1e-7


def test_pyfai_1d_conserves_pixels():
    p = O.Poni(172e-6, 172e-6, 0.18, 1043 * 172e-6 / 2 * 0.05, 981 * 172e-6 / 2 * 0.05,
               0.05, -0.03, 0.10, 0.3344e-10)
    lut = O.build_lut1d_pyfai(p, 40, 50, 300)
    tot = np.bincount(lut.col, weights=lut.w, minlength=2000)
    assert np.abs(tot - 1.0).max() < 1e-9


def test_pyfai_2d_totals():
    p = O.Poni(172e-6, 172e-6, 0.05, 24 * 172e-6, 36 * 172e-6, 0.02, -0.01, 0.05, 1e-10)
    lut, (r0, r1, a0, a1) = O.build_lut2d_pyfai(p, 64, 80, 32, 36)
    assert (a0, a1) == (-math.pi, math.pi) and r0 >= 0.0
    tot = np.bincount(lut.col, weights=lut.w, minlength=64 * 80)
    # Split pixels are renormalized to one; the sampled beam pixel keeps only
    # its samples above the first radial edge (2D grids have no low clamp).
    off = np.abs(tot - 1.0) > 1e-9
    assert off.sum() <= 1 and np.all(tot[off] < 1.0)
