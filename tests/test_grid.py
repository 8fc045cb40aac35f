"""Explicit shared 2D grids and azimuthal brackets (host logic, no GPU).

Known answers of the reference's own tests:
  src/integration/azimuthal_range_tests.rs:1119-1255 (brackets: 72 rows of
  2.5 degrees, nine 10-degree rows over a quadrant, rejected tilings, far
  starts corrected from the first centre, single-precision turns);
  src/integration/grid.rs:48-105 (IntegrationGrid2d::new validation and the
  correction of starts outside +-2 pi).
"""
import math

import numpy as np
import pytest

from dioptrin_amd.grid import (IntegrationGrid2d, azimuthal_bracket_bins,
                               resolve_azimuthal_bins)

PI, TAU = math.pi, 2.0 * math.pi


def test_brackets_centre_rows_on_the_requested_directions():
    bracket = math.radians(2.5)
    n, (lo, hi) = azimuthal_bracket_bins((0.0, TAU), bracket)
    assert n == 72
    g = IntegrationGrid2d(5, n, (0.0, 1.0), (lo, hi))
    c = g.azimuthal_centers()
    assert np.all(np.abs(c - 2.0 * bracket * np.arange(72)) < 1e-12)
    assert abs(c[18] - PI / 2) < 1e-12


def test_brackets_describe_sectors_and_reject_uneven_tilings():
    n, (lo, hi) = azimuthal_bracket_bins((0.0, PI / 2), math.radians(5.0))
    assert n == 9
    assert abs(lo + math.radians(5.0)) < 1e-12
    assert abs(hi - math.radians(85.0)) < 1e-12
    for rng, bracket in [((0.0, TAU), math.radians(3.5)), ((0.0, TAU), 0.0),
                         ((0.0, TAU), -0.1), ((0.0, TAU), math.nan), ((0.0, TAU), TAU),
                         ((1.0, 1.0), 0.1), ((0.0, math.inf), 0.1), ((0.0, 2 * TAU), 0.1 * TAU)]:
        with pytest.raises(ValueError):
            azimuthal_bracket_bins(rng, bracket)


def test_brackets_accept_f32_and_correct_distant_centres():
    bracket = math.radians(2.5)
    f32 = lambda x: float(np.float32(x))
    assert azimuthal_bracket_bins((0.0, TAU), f32(math.radians(2.5)))[0] == 72
    first_centre = lambda rng: azimuthal_bracket_bins(rng, bracket)[1][0] + bracket
    assert abs(first_centre((TAU, 2 * TAU)) - TAU) < 1e-12
    assert abs(first_centre((-PI, PI)) + PI) < 1e-12
    for rows in (1.0, 2.0, 100.0, 1000.0):
        for start, expected in [(-TAU + 0.01, 0.01), (-1.5 * PI, 0.5 * PI),
                                (-PI - 1e-9, PI - 1e-9), (TAU + 1e-9, 1e-9)]:
            half = PI / rows
            n, (lo, hi) = azimuthal_bracket_bins((start, start + TAU), half)
            assert n == rows
            assert abs(lo + half - expected) < 1e-9, (start, rows)
            # The grid must not shift the edges again.
            assert IntegrationGrid2d(1, n, (0.0, 1.0), (lo, hi)).azimuthal_range == (lo, hi)
    for end in (PI + 4e-13, f32(PI)):
        n, (lo, hi) = azimuthal_bracket_bins((-PI, end), PI)
        assert n == 1
        assert abs(lo + PI + PI) < 1e-12 and hi - lo <= TAU
        IntegrationGrid2d(1, n, (0.0, 1.0), (lo, hi))
    assert azimuthal_bracket_bins((0.0, f32(TAU)), f32(math.radians(2.5)))[0] == 72
    with pytest.raises(ValueError):
        azimuthal_bracket_bins((0.0, TAU), PI / 1_000_000.5)
    assert abs(first_centre((TAU + 2 * bracket, 2 * TAU + 2 * bracket)) - 2 * bracket) < 1e-12
    assert abs(first_centre((-TAU, 0.0))) < 1e-12
    assert abs(first_centre((-3 * TAU, -2 * TAU))) < 1e-12


def test_resolve_needs_exactly_one_of_count_and_bracket():
    assert resolve_azimuthal_bins(36, None, None) == (36, None)
    n, edges = resolve_azimuthal_bins(None, None, math.radians(5.0))
    assert n == 36 and abs(edges[0] + math.radians(5.0)) < 1e-12
    with pytest.raises(ValueError, match="not both"):
        resolve_azimuthal_bins(36, None, 0.1)
    with pytest.raises(ValueError, match="either"):
        resolve_azimuthal_bins(None, None, None)


def test_grid_validation_and_start_correction():
    g = IntegrationGrid2d(10, 4, (0.0, 0.5), (-PI, PI))
    assert g.shape == (4, 10)
    # lo + (i + 0.5) * step (grid.rs:224-227).
    assert np.array_equal(g.radial_centers(), 0.0 + (np.arange(10) + 0.5) * (0.5 / 10))
    for args in [(0, 4, (0.0, 1.0), (-PI, PI)), (3, 0, (0.0, 1.0), (-PI, PI)),
                 (3, 4, (1.0, 1.0), (-PI, PI)), (3, 4, (0.0, math.nan), (-PI, PI)),
                 (3, 4, (-0.1, 1.0), (-PI, PI)), (3, 4, (0.0, 1.0), (-PI, PI + 0.1)),
                 (3, 4, (1e300, 1e300 * (1 + 1e-16)), (-PI, PI)),
                 (3, 4, (0.0, 1.0), (0.0, -1.0))]:
        with pytest.raises(ValueError):
            IntegrationGrid2d(*args)
    # Starts within +-2 pi are kept; others move to [-pi, pi) with the span.
    assert IntegrationGrid2d(1, 2, (0.0, 1.0), (TAU - 1.0, TAU)).azimuthal_range == (TAU - 1.0,
                                                                                    TAU)
    g = IntegrationGrid2d(1, 2, (0.0, 1.0), (-3 * PI, -PI))
    assert g.azimuthal_range[0] == pytest.approx(-PI) and \
        g.azimuthal_range[1] - g.azimuthal_range[0] == pytest.approx(2 * PI)
    g = IntegrationGrid2d(1, 2, (0.0, 1.0), (7.0, 8.0))
    assert g.azimuthal_range == (math.fmod(7.0, TAU), math.fmod(7.0, TAU) + 1.0)
    with pytest.raises(ValueError):
        g.with_pixel_split_mode("exact")
    h = g.with_pixel_split_mode("pyfai_histogram")
    assert h.pixel_split_mode == "pyfai_histogram" and h.azimuthal_range == g.azimuthal_range
    assert h != g and g == IntegrationGrid2d(1, 2, (0.0, 1.0), (7.0, 8.0))
