"""Explicit shared 2D bin edges and raw-sum results (multi-detector merging).

IntegrationGrid2d mirrors the reference's IntegrationGrid2d
(src/integration/grid.rs:20-227): half-open radial and azimuthal bins whose
edges are given, values inside the edges whose index rounds to n kept in the
last bin, azimuthal starts outside +-2 pi corrected to [-pi, pi). A plan
built on a grid is a 2D plan with both ranges explicit: the device builder
applies the same edge rules (csrc/builder.hip grid_new).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

TAU = 2.0 * math.pi
SPLIT_MODES = ("fractional", "pyfai_histogram")


def _wrap_start(x: float) -> float:
    """x.rem_euclid(2 pi), moved to [-pi, pi) (grid.rs:81-90, 279-288)."""
    w = math.fmod(x, TAU)
    if w < 0.0:
        w += TAU
    return w - TAU if w >= math.pi else w


class IntegrationGrid2d:
    """Shared 2D bin edges (IntegrationGrid2d::new, grid.rs:48-105): radial
    values in the integrator's unit, azimuthal values in radians."""

    def __init__(self, n_radial: int, n_azimuthal: int, radial_range: Tuple[float, float],
                 azimuthal_range: Tuple[float, float] = (-math.pi, math.pi)):
        n_radial, n_azimuthal = int(n_radial), int(n_azimuthal)
        if n_radial <= 0 or n_azimuthal <= 0 or n_radial * n_azimuthal > (2 ** 63 - 1) // 8:
            raise ValueError("2D bin counts must be positive and their product must fit in "
                             "memory indexing")
        rr = (float(radial_range[0]), float(radial_range[1]))
        ar = (float(azimuthal_range[0]), float(azimuthal_range[1]))
        for name, (lo, hi), count in (("radial", rr, n_radial), ("azimuthal", ar, n_azimuthal)):
            step = (hi - lo) / count
            if not (math.isfinite(lo) and math.isfinite(hi) and math.isfinite(step)) or \
                    step <= 0.0 or lo + step == lo or hi - step == hi:
                raise ValueError(f"{name} range must have finite, increasing, resolvable bin "
                                 f"edges")
        if rr[0] < 0.0:
            raise ValueError("radial range must be nonnegative")
        if ar[1] - ar[0] > TAU + 1e-12:
            raise ValueError("azimuthal span must be at most 2π (one full turn)")
        if not -TAU <= ar[0] <= TAU:
            start = _wrap_start(ar[0])
            ar = (start, start + (ar[1] - ar[0]))
        self.n_radial, self.n_azimuthal = n_radial, n_azimuthal
        self.radial_range, self.azimuthal_range = rr, ar
        self.pixel_split_mode = "fractional"

    def with_pixel_split_mode(self, mode: str) -> "IntegrationGrid2d":
        """The same edges with another splitting convention: "fractional"
        (polygon areas) or "pyfai_histogram" (pyFAI's full splitting)."""
        if mode not in SPLIT_MODES:
            raise ValueError("pixel_split_mode must be 'fractional' or 'pyfai_histogram'")
        g = IntegrationGrid2d(self.n_radial, self.n_azimuthal, self.radial_range,
                              self.azimuthal_range)
        g.pixel_split_mode = mode
        return g

    @property
    def shape(self) -> Tuple[int, int]:
        return (self.n_azimuthal, self.n_radial)

    @staticmethod
    def _centers(rng, count):
        # grid.rs:224-227: lo + (i + 0.5) * step.
        lo, hi = rng
        step = (hi - lo) / count
        return lo + (np.arange(count, dtype=np.float64) + 0.5) * step

    def radial_centers(self) -> np.ndarray:
        return self._centers(self.radial_range, self.n_radial)

    def azimuthal_centers(self) -> np.ndarray:
        return self._centers(self.azimuthal_range, self.n_azimuthal)

    def __eq__(self, other):
        return isinstance(other, IntegrationGrid2d) and \
            (self.n_radial, self.n_azimuthal, self.radial_range, self.azimuthal_range,
             self.pixel_split_mode) == (other.n_radial, other.n_azimuthal, other.radial_range,
                                        other.azimuthal_range, other.pixel_split_mode)

    def __hash__(self):
        return hash((self.n_radial, self.n_azimuthal, self.radial_range, self.azimuthal_range,
                     self.pixel_split_mode))

    def __repr__(self):
        return (f"IntegrationGrid2d({self.n_radial}, {self.n_azimuthal}, {self.radial_range}, "
                f"{self.azimuthal_range}, pixel_split_mode={self.pixel_split_mode!r})")


def azimuthal_bracket_bins(center_range: Tuple[float, float], half_width: float):
    """(n rows, (first edge, last edge)) of contiguous azimuthal rows centred
    from center_range[0] in steps of 2 * half_width, end exclusive
    (azimuthal_bracket_bins, grid.rs:229-290)."""
    first, end = float(center_range[0]), float(center_range[1])
    span = end - first
    if not (math.isfinite(first) and math.isfinite(end) and math.isfinite(span)) or span <= 0.0:
        raise ValueError("azimuthal centre range must be finite and increasing")
    half_width = float(half_width)
    if not math.isfinite(half_width) or half_width <= 0.0:
        raise ValueError("azimuthal bracket must be finite and positive")
    if span > TAU * (1.0 + 1e-6):
        raise ValueError("azimuthal centres must span at most 2π (one full turn)")
    span = min(span, TAU)
    rows = span / (2.0 * half_width)
    # Rust's f64::round rounds halves away from zero.
    n = math.floor(rows + 0.5) if rows >= 0 else -math.floor(-rows + 0.5)
    if not math.isfinite(n) or n < 1.0 or abs(rows - n) > min(1e-6 * n, 1e-3):
        raise ValueError(f"azimuthal bracket of ±{half_width} rad does not divide the span of "
                         f"{span} rad into a whole number of rows ({rows} rows)")
    half = span / (2.0 * n)
    if not -math.pi <= first <= TAU:
        first = _wrap_start(first)
    return int(n), (first - half, first - half + span)


def resolve_azimuthal_bins(n_azimuthal: Optional[int], azimuthal_range, azimuthal_bracket):
    """Either n_azimuthal or azimuthal_bracket (python/integrator.rs:43-63)."""
    if n_azimuthal is not None and azimuthal_bracket is None:
        return int(n_azimuthal), azimuthal_range
    if n_azimuthal is None and azimuthal_bracket is not None:
        centres = azimuthal_range if azimuthal_range is not None else (0.0, TAU)
        n, edges = azimuthal_bracket_bins(centres, azimuthal_bracket)
        return n, edges
    if n_azimuthal is not None:
        raise ValueError("give either n_azimuthal or azimuthal_bracket, not both")
    raise ValueError("give either n_azimuthal or azimuthal_bracket")


@dataclass
class IntermediateResult1d:
    """Sums before division of one 1D integration (lut/integrate.rs:371-378)."""
    bins: np.ndarray
    sum_signal: np.ndarray
    sum_norm: np.ndarray
    sum_variance: Optional[np.ndarray]
    count: np.ndarray


@dataclass
class IntermediateResult2d:
    """Sums before division of one 2D integration, flat over az * n_radial + r
    (lut/integrate.rs:380-388)."""
    sum_signal: np.ndarray
    sum_norm: np.ndarray
    sum_variance: Optional[np.ndarray]
    count: np.ndarray
    n_radial: int
    n_azimuthal: int
