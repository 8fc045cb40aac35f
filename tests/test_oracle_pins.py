"""Pins of the oracle's 2D and pyFAI-compatible splitting that the reference
checks only through its own tests, plus brute-force checks of where the 2D
tables put each pixel along the azimuth.

Reference tests ported (the cases and expectations are the reference's):
  pyfai_1d.rs:360-388   wrap_tau is rem_euclid bit for bit
  pyfai_1d.rs:415-424   partial radial pixels renormalize, disjoint sectors
                        exclude
  pyfai_1d.rs:429-434   the seam test uses pyFAI's operator precedence
  azimuthal_range_tests.rs:160-191   counting puts each pixel in the bin
                        holding its azimuth
  azimuthal_range_tests.rs:193-237   sectors exclude pixels outside their
                        edges (counting, pixel split, supersampled)
  azimuthal_range_tests.rs:390-447   index rounding stays in the last bin
  azimuthal_range_tests.rs:1043-1070 shared grids keep values whose index
                        rounds to n (the grid arithmetic)

Brute force, independent of the oracle's clipping code:
  pixel split 2D  each weight = area of (pixel quad in (2theta, chi),
                  unwrapped around its centre) n bin rectangle / quad area,
                  summed over the +-2pi copies, by exact strip integration
  pyFAI 2D        pyFAI's raster rule (edge areas per radial strip stacked
                  into the azimuthal rows from the bottom, which the reference
                  states differs from exact clipping), to float32 accuracy
  supersampled 2D each weight = (subpixels whose (2theta, chi) falls in the
                  bin) / n_ss^2, counted one subpixel at a time
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

TAU = 2.0 * math.pi
PI = math.pi

# azimuthal_range_tests.rs:8-27: beam inside the detector, slight tilt.
ROWS, COLS, NR, NA = 60, 80, 12, 18
PONI = dict(pixel1=1e-4, pixel2=1e-4, distance=0.05, poni1=0.003, poni2=0.004, rot1=0.01,
            rot2=-0.02, wavelength=1e-10)
P = O.Poni(**PONI)


# ------------------------------------------------------------ pyfai_1d.rs

def rem_euclid_exact(x: float) -> float:
    """Rust's x.rem_euclid(TAU) from its definition: r = x % TAU (fmod,
    whose result x - trunc(x / TAU) * TAU is exact), then r + TAU rounded
    once when r < 0; computed in rational arithmetic."""
    if math.isnan(x) or math.isinf(x):
        return math.nan
    if x == 0.0:
        return x
    fx, ft = Fraction(x), Fraction(TAU)
    q = fx / ft
    n = math.floor(q) if q >= 0 else -math.floor(-q)   # trunc
    r = fx - n * ft
    if r == 0:
        return math.copysign(0.0, x)
    r = float(r)                                         # exact (fmod)
    return float(Fraction(r) + ft) if r < 0 else r


def test_wrap_tau_is_rem_euclid_bit_for_bit():
    values = [0.0, -0.0, TAU, -TAU, PI, -PI, math.nextafter(TAU, 0.0), math.nextafter(TAU, 10.0),
              2.2250738585072014e-308, -2.2250738585072014e-308, 1e300, -1e300, math.inf,
              -math.inf, math.nan]
    state = 0x9E3779B97F4A7C15
    mask = (1 << 64) - 1
    for _ in range(100_000):
        state ^= (state << 13) & mask
        state ^= state >> 7
        state ^= (state << 17) & mask
        values.append((state >> 11) / float(1 << 53) * 8.0 * TAU - 4.0 * TAU)
    for x in values:
        got, want = O.wrap_tau(x), rem_euclid_exact(x)
        if math.isnan(want):
            assert math.isnan(got), x
        else:
            assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64), x


def test_pyfai_partial_radial_pixels_renormalize_and_disjoint_sectors_exclude():
    poly = [(-0.5, 0.1), (-0.5, 0.8), (0.75, 0.8), (0.75, 0.1)]
    w = O.pyfai_split1d_sector(poly, 0.0, 2.0, 4, 0.2, 0.4, sector=True)
    assert w and abs(sum(x for _, x in w) - 1.0) < 1e-12
    assert all(math.isfinite(x) and x > 0 for _, x in w)
    assert O.pyfai_split1d_sector(poly, 0.0, 2.0, 4, 1.0, 2.0, sector=True) == []


def test_pyfai_seam_test_uses_operator_precedence():
    """0.5 * X - Y is negative for this folded quadrilateral (no unwrapping),
    while the signed area 0.5 * (X - Y) is positive, which would unwrap two
    corners above 2pi - 1 and pull the pixel into the sector."""
    poly = [(1.0, -0.5), (1.2, -0.4), (2.0, 0.25), (2.2, 0.6)]
    assert O.pyfai_split1d_sector(poly, 0.0, 3.0, 10, 1.0, 2.0, sector=True) == []


# ---------------------------------------------------------------- 2D grids

def test_index_rounding_stays_in_the_last_bin():
    n = NA
    found = None
    for k in range(1, 2000):
        lo = -PI + k * 1e-4
        hi = lo + 2.0
        step = (hi - lo) / n
        v = hi
        for _ in range(4):
            v = math.nextafter(v, -math.inf)
            if int((v - lo) / step) >= n:
                found = (lo, hi, v)
                break
        if found:
            break
    lo, hi, v = found
    R = O.GRID_INDEX_ROUNDING
    assert O.grid2d_bin(NR, n, (0.0, 1.0), (lo, hi), 0, 0.5, v) is None
    assert O.grid2d_bin(NR, n, (0.0, 1.0), (lo, hi), R, 0.5, v) == (n - 1) * NR + NR // 2
    assert O.grid2d_bin(NR, n, (0.0, 1.0), (lo, hi), R, 0.5, hi) is None
    assert O.grid2d_bin(NR, n, (0.0, 1.0), (lo, hi), R, 0.5, lo - 1e-12) is None
    found = None
    for k in range(1, 2000):
        r1 = 30.0 + k * 1e-3
        step = r1 / NR
        v = r1
        for _ in range(4):
            v = math.nextafter(v, -math.inf)
            if int(v / step) >= NR:
                found = (r1, v)
                break
        if found:
            break
    r1, v = found
    assert O.grid2d_bin(NR, n, (0.0, r1), (0.0, 1.0), 0, v, 0.5) is None
    assert O.grid2d_bin(NR, n, (0.0, r1), (0.0, 1.0), R, v, 0.5) == 9 * NR + NR - 1
    below = math.nextafter(-PI, -math.inf)
    assert O.grid2d_bin(NR, n, (0.0, 1.0), (-PI, PI), R, 0.5, below) == (n - 1) * NR + NR // 2


def test_shared_grids_keep_values_whose_index_rounds_to_n():
    R = O.GRID_INDEX_ROUNDING
    for (lo, hi), n in (((0.5, 25.0), 1500), ((0.0, 8.0), 7), ((1.0, 30.8), 100)):
        v = math.nextafter(hi, -math.inf)
        assert int((v - lo) / ((hi - lo) / n)) >= n
        assert O.grid2d_bin(n, 4, (lo, hi), (-PI, PI), 0, v, 0.1) is None
        assert O.grid2d_bin(n, 4, (lo, hi), (-PI, PI), R, v, 0.1) == 2 * n + n - 1
        assert O.grid2d_bin(n, 4, (lo, hi), (-PI, PI), R, hi, 0.1) is None
    for az in (-1e-16, -4e-16):
        assert O.grid2d_bin(NR, NA, (0.0, 1.0), (0.0, TAU), 0, 0.5, az) is None
        assert O.grid2d_bin(NR, NA, (0.0, 1.0), (0.0, TAU), R, 0.5, az) == \
            (NA - 1) * NR + NR // 2
    for az in (-1e-16, 1.0, 2.0, -PI):
        assert O.grid2d_bin(NR, NA, (0.0, 1.0), (0.0, 1.0), R, 0.5, az) is None


def unwrapped_azimuth(rng):
    c = 0.5 * (rng[0] + rng[1])
    _, chi = O.centres(P, ROWS, COLS)
    return c - PI + np.array([O.wrap_tau(a - c + PI) for a in chi])


def table_2d(method, rng):
    if method == "supersampled":
        lut, _ = O.build_lut2d_ss(P, ROWS, COLS, NR, NA, 3, "2th_deg", None, rng)
    else:
        lut, _ = O.build_lut2d(P, ROWS, COLS, NR, NA, method, "2th_deg", None, rng)
    return lut


def gather_raw(lut, x):
    """Per bin sum w * x (denominators of 1)."""
    i, _ = O.gather(lut, np.ones(lut.nbins), np.asarray(x, np.float64), False)
    return i


RANGES = [(0.0, PI / 2.0), (2.5, 2.5 + PI / 2.0), (-PI, PI), (0.3, 0.3 + TAU)]


@pytest.mark.parametrize("rng", RANGES, ids=[f"{a:.2f}_{b:.2f}" for a, b in RANGES])
def test_counting_puts_each_pixel_in_the_bin_holding_its_azimuth(rng):
    lut = table_2d("counting", rng)
    az = unwrapped_azimuth(rng)
    norm = gather_raw(lut, np.ones(ROWS * COLS))
    sig = gather_raw(lut, az)
    step = (rng[1] - rng[0]) / NA
    full = rng[1] - rng[0] > 6.0
    filled = 0
    for ia in range(NA):
        lo = rng[0] + ia * step
        any_ = False
        for ir in range(NR):
            k = ia * NR + ir
            if norm[k] == 0.0 or (full and ia in (0, NA - 1)):
                continue
            any_ = True
            mean = sig[k] / norm[k]
            assert lo - 1e-9 <= mean <= lo + step + 1e-9, (rng, ia, mean)
        filled += any_
    assert filled >= NA - 2


@pytest.mark.parametrize("method", ["counting", "pixel_split", "supersampled"])
def test_sectors_exclude_pixels_outside_their_edges(method):
    tth, _ = O.centres(P, ROWS, COLS)
    radius_px = PONI["distance"] * np.tan(tth) / PONI["pixel1"]
    image = 10.0 + ((np.arange(ROWS * COLS) * 37) % 101).astype(np.float64)
    for rng in ((0.0, PI / 2.0), (2.5, 2.5 + PI / 2.0), (-2.0, -1.0)):
        lut = table_2d(method, rng)
        az = unwrapped_azimuth(rng)
        far = ((radius_px > 15.0) & ((az < rng[0] - 0.1) | (az > rng[1] + 0.1))).astype(float)
        assert far.sum() > 100.0
        assert np.all(gather_raw(lut, far) == 0.0), (method, rng)
        assert np.count_nonzero(gather_raw(lut, image) > 0) > 40


# ------------------------------------------------- brute-force chi placement

def poly_area(poly):
    s = 0.0
    for i in range(len(poly)):
        x0, y0 = poly[i]
        x1, y1 = poly[(i + 1) % len(poly)]
        s += x0 * y1 - x1 * y0
    return 0.5 * s


def clip_area(poly, x0, x1, y0, y1):
    """Area of a convex polygon inside [x0, x1] x [y0, y1] by strip
    integration: between consecutive breakpoints (vertices, rectangle edges
    and crossings of y0 / y1) the clipped vertical extent is linear in x, so
    the trapezoid rule is exact."""
    n = len(poly)
    xs = {x0, x1} | {p[0] for p in poly}
    for i in range(n):
        (ax, ay), (bx, by) = poly[i], poly[(i + 1) % n]
        for yc in (y0, y1):
            if (ay - yc) * (by - yc) < 0:
                xs.add(ax + (yc - ay) * (bx - ax) / (by - ay))
    pxmin, pxmax = min(p[0] for p in poly), max(p[0] for p in poly)
    xs = sorted(x for x in xs if max(x0, pxmin) <= x <= min(x1, pxmax))

    def extent(x):
        ys = []
        for i in range(n):
            (ax, ay), (bx, by) = poly[i], poly[(i + 1) % n]
            if ax == bx:
                if x == ax:
                    ys += [ay, by]
                continue
            if min(ax, bx) <= x <= max(ax, bx):
                ys.append(ay + (x - ax) * (by - ay) / (bx - ax))
        if not ys:
            return 0.0
        return max(0.0, min(max(ys), y1) - max(min(ys), y0))

    area = 0.0
    for a, b in zip(xs, xs[1:]):
        area += (b - a) * 0.5 * (extent(a) + extent(b))
    return area


def pixel_quad(tc, ac, tch, px):
    """(2theta, chi) of pixel px's corners in the winding (i00, i01, i11,
    i10), chi unwrapped to within pi of the pixel-centre chi."""
    r, c = divmod(px, COLS)
    w = COLS + 1
    idx = [r * w + c, r * w + c + 1, (r + 1) * w + c + 1, (r + 1) * w + c]
    return [(tc[i], tch + math.remainder(ac[i] - tch, TAU)) for i in idx]


def expected_weights(quad, edges, nr, na, renormalize):
    r0, r1, a0, a1 = edges
    dr, da = (r1 - r0) / nr, (a1 - a0) / na
    total = abs(poly_area(quad))
    out = {}
    for k in (-1, 0, 1):
        q = [(x, y + k * TAU) for x, y in quad]
        for ia in range(na):
            for ir in range(nr):
                a = clip_area(q, r0 + ir * dr, r0 + (ir + 1) * dr, a0 + ia * da,
                              a0 + (ia + 1) * da)
                if a > 0:
                    out[ia * nr + ir] = out.get(ia * nr + ir, 0.0) + a / total
    if renormalize:
        s = sum(out.values())
        out = {b: v / s for b, v in out.items()}
    return out


def table_entries(lut, px):
    rows = np.repeat(np.arange(lut.nbins), np.diff(lut.row_ptr).astype(np.int64))
    sel = lut.col == px
    return dict(zip(rows[sel].tolist(), lut.w[sel].tolist()))


def sample_pixels(tth_c, chi_c, n, seed):
    """Pixels away from the beam and from the chi seam, corners included."""
    rng = np.random.default_rng(seed)
    ok = (tth_c > 0.02) & (np.abs(np.abs(chi_c) - PI) > 0.2)
    return rng.choice(np.flatnonzero(ok), size=n, replace=False)


def test_2d_pixel_split_weights_are_exact_areas():
    """Every weight of 40 pixels of the 2D pixel-split table (12 x 18 bins
    over the full circle, 2theta in radians) against the brute-force areas."""
    lut, edges = O.build_lut2d(P, ROWS, COLS, NR, NA)
    tc, ac = O.corners(P, ROWS, COLS)
    tth_c, chi_c = O.centres(P, ROWS, COLS)
    for px in sample_pixels(tth_c, chi_c, 40, 5):
        quad = pixel_quad(tc, ac, chi_c[px], int(px))
        want = expected_weights(quad, edges, NR, NA, False)
        got = table_entries(lut, int(px))
        assert set(got) == {b for b, v in want.items() if v > 1e-9}, px
        for b, v in got.items():
            assert abs(v - want[b]) <= 1e-9, (px, b, v, want[b])


def pyfai_raster_weights(quad, edges, nr, na):
    """pyFAI's fullSplit2D histogram rule as the reference states it
    (pyfai_split.rs:1-4, 140-170): corners in bin units relative to the
    pixel's lowest (radial, azimuthal) cell; each non-vertical edge's signed
    area over a radial strip [left, right] of one column is width * h, h the
    edge's height at the strip centre, and it is stacked into the rows from
    the bottom, row a receiving width * clip(h - a, 0, 1); the cells' signed
    totals, kept where positive and inside the grid, renormalized to 1."""
    r0, r1, a0, a1 = edges
    dr, da = (r1 - r0) / nr, (a1 - a0) / na
    pts = [((min(max(x, r0), r1) - r0) / dr, (min(max(y, a0), a1) - a0) / da)
           for x, y in quad]
    ir = math.floor(min(p[0] for p in pts))
    ia = math.floor(min(p[1] for p in pts))
    pts = [(x - ir, y - ia) for x, y in pts]
    cells = {}
    for k in range(4):
        (x0, y0), (x1, y1) = pts[k], pts[(k + 1) % 4]
        if x0 == x1:
            continue
        slope = (y1 - y0) / (x1 - x0)
        icpt = y1 - slope * x1
        sgn = 1.0 if x1 > x0 else -1.0
        lo, hi = min(x0, x1), max(x0, x1)
        for col in range(math.floor(lo), math.ceil(hi)):
            left, right = max(lo, col), min(hi, col + 1)
            h = 0.5 * slope * (right + left) + icpt
            area = sgn * (right - left) * h
            for row in range(0, math.ceil(abs(h)) + 1):
                part = (right - left) * min(max(abs(h) - row, 0.0), 1.0)
                if part > 0:
                    cells[(col, row)] = cells.get((col, row), 0.0) + math.copysign(part, area)
    # The quad's winding sets the sign of every cell; pyFAI divides by the
    # signed total, so orient the cells by the sign of their sum.
    orient = 1.0 if sum(cells.values()) >= 0 else -1.0
    out = {(ia + a) * nr + ir + c: orient * v for (c, a), v in cells.items()
           if orient * v > 0 and ir + c < nr and ia + a < na}
    total = sum(out.values())
    return {b: v / total for b, v in out.items()}


def test_2d_pyfai_weights_follow_the_raster_rule():
    """Every weight of 40 pixels of the pyFAI-compatible 2D table against the
    raster rule evaluated in float64 (the table's line buffers are float32:
    tolerance 5e-5 on weights of order 1)."""
    lut, edges = O.build_lut2d_pyfai(P, ROWS, COLS, NR, NA)
    tc, ac = O.corners(P, ROWS, COLS)
    tth_c, chi_c = O.centres(P, ROWS, COLS)
    for px in sample_pixels(tth_c, chi_c, 40, 5):
        quad = pixel_quad(tc, ac, chi_c[px], int(px))
        want = pyfai_raster_weights(quad, edges, NR, NA)
        got = table_entries(lut, int(px))
        assert set(got) == {b for b, v in want.items() if v > 1e-6}, (px, got, want)
        for b, v in got.items():
            assert abs(v - want[b]) <= 5e-5, (px, b, v, want[b])


def test_2d_supersampled_weights_count_subpixels():
    """Each weight of 40 pixels of the supersampled 2D table (n_ss = 3) is
    the count of the pixel's subpixels whose (2theta, chi) falls in the bin,
    over 9: subpixel centres row + (i + 0.5) * (1/3), ((col * 3 + j) + 0.5)
    * (1/3) (supersampled.rs:310-311, 904-905), bins by the grid's rule with
    the default clamp flags."""
    n = 3
    lut, edges = O.build_lut2d_ss(P, ROWS, COLS, NR, NA, n)
    inv = 1.0 / n
    tth_c, chi_c = O.centres(P, ROWS, COLS)
    flags = O.GRID_INDEX_ROUNDING | O.GRID_RADIAL_OUTSIDE | O.GRID_AZIMUTHAL_OUTSIDE
    for px in sample_pixels(tth_c, chi_c, 40, 6):
        r, c = divmod(int(px), COLS)
        d1 = np.array([r + (i + 0.5) * inv for i in range(n) for j in range(n)])
        d2 = np.array([((c * n + j) + 0.5) * inv for i in range(n) for j in range(n)])
        t, a = O.tth_chi(P, d1, d2)
        counts = {}
        for tt, aa in zip(t, a):
            b = O.grid2d_bin(NR, NA, edges[:2], edges[2:], flags, tt, aa)
            if b is not None:
                counts[b] = counts.get(b, 0) + 1
        want = {b: k / (n * n) for b, k in counts.items()}
        assert table_entries(lut, int(px)) == want, px
