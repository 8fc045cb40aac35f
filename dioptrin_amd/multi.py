"""Multi-detector integration (the reference's MultiIntegrator,
src/multi.rs:88-516 and src/python/multi_integrator.rs:1-317).

Every detector keeps its own Integrator (geometry, method, mask,
polarization) and plan; all share one bin grid: the common radial range and
unit, and in 2D the same explicit azimuthal edges. Per detector the GPU forms
the sums before division; dio_multi_integrate adds them on the device in
detector order and divides once.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import List, Optional, Sequence

import numpy as np

from . import _native as N
from .grid import SPLIT_MODES, IntegrationGrid2d, resolve_azimuthal_bins
from .integrator import _NP_DTYPES, Integrator, IntegrationResult
from .tiff import read_tiff

_EPS = np.finfo(np.float64).eps
_WIDEN = Integrator._WIDEN


def _validate(integrators: Sequence[Integrator]) -> None:
    """validate_integrators (multi.rs:100-130)."""
    if not integrators:
        raise ValueError("MultiIntegrator requires at least one integrator")
    first = integrators[0]
    if first.radial_range is None:
        raise ValueError("All integrators must have a radial_range set for MultiIntegrator")
    for i, it in enumerate(integrators):
        if it.unit != first.unit:
            raise ValueError(f"All integrators must have the same radial_unit (integrator {i} "
                             f"differs)")
        if it.radial_range is None:
            raise ValueError(f"All integrators must have a radial_range set (integrator {i} "
                             f"missing)")
        if abs(it.radial_range[0] - first.radial_range[0]) > _EPS or \
                abs(it.radial_range[1] - first.radial_range[1]) > _EPS:
            raise ValueError(f"All integrators must have the same radial_range (integrator {i} "
                             f"differs)")


class MultiIntegrator:
    """Merges one image per detector into one pattern by summing raw signal,
    normalization and variance across detectors before dividing."""

    def __init__(self, integrators: Sequence[Integrator], *,
                 pixel_split_mode: str = "fractional"):
        if pixel_split_mode not in SPLIT_MODES:
            raise ValueError("pixel_split_mode must be 'fractional' or 'pyfai_histogram'")
        integrators = list(integrators)
        _validate(integrators)
        if len({it.device for it in integrators}) != 1:
            raise ValueError("all integrators must use one device")
        self.integrators = integrators
        self.pixel_split_mode = pixel_split_mode

    # -- inputs
    def _load(self, images, context: str = "") -> List[np.ndarray]:
        """One image per detector: TIFF paths (read, flipped as the reference
        reads them) or 2D arrays, each checked against its detector's mask
        (load_images, python/multi_integrator.rs:246-291)."""
        if isinstance(images, (str, bytes, bytearray, os.PathLike)):
            raise TypeError("images must be a list with one image per detector, not a single "
                            "path")
        images = list(images)
        n = len(self.integrators)
        if len(images) != n:
            raise ValueError(f"Expected {n} images (one per detector), got {len(images)}")
        out = []
        for i, (im, it) in enumerate(zip(images, self.integrators)):
            if isinstance(im, (str, bytes, bytearray, os.PathLike)):
                try:
                    arr = read_tiff(os.fsdecode(im))
                except (OSError, ValueError) as e:
                    raise ValueError(f"{context}detector {i}: failed to load TIFF: {e}") from e
                arr = np.ascontiguousarray(arr.astype(_WIDEN.get(arr.dtype, arr.dtype),
                                                      copy=False))
            else:
                arr = np.asarray(im)
                if arr.ndim != 2:
                    raise ValueError(f"{context}detector {i}: image must be 2D")
                if arr.dtype not in _NP_DTYPES:
                    raise TypeError(f"{context}detector {i}: unsupported dtype {arr.dtype}; "
                                    f"expected float64, float32, uint16, uint32 or int32")
                if not arr.flags["C_CONTIGUOUS"]:
                    raise ValueError(f"{context}detector {i}: array must be contiguous")
            if arr.size == 0:
                raise ValueError(f"{context}detector {i}: image is empty")
            if it.mask is not None and arr.shape != it.mask.shape:
                raise ValueError(f"{context}detector {i} image shape {arr.shape} does not "
                                 f"match mask shape {it.mask.shape}")
            out.append(arr)
        return out

    # -- device merge
    @staticmethod
    def _merge(plans, frames, variances=None, factors=None, errors=False, raw=False):
        """dio_multi_integrate over stacks (n_frames, dim1, dim2) per detector."""
        nd = len(plans)
        nf = frames[0].shape[0]
        nbins = plans[0].nbins
        P = (C.c_void_p * nd)(*[p._h for p in plans])
        F = (C.c_void_p * nd)(*[f.ctypes.data for f in frames])
        D = (C.c_int * nd)(*[N.DTYPES[_NP_DTYPES[f.dtype]] for f in frames])
        V = VD = None
        if variances is not None:
            V = (C.c_void_p * nd)(*[v.ctypes.data for v in variances])
            VD = (C.c_int * nd)(*[N.DTYPES[_NP_DTYPES[v.dtype]] for v in variances])
        FA = None if factors is None else (C.c_double * nd)(*factors)
        errors = errors or variances is not None
        I = np.empty((nf, nbins))
        S = np.empty((nf, nbins)) if errors else None
        rs = None
        if raw:
            rs = dict(sum_signal=np.empty((nf, nbins)), sum_norm=np.empty(nbins),
                      sum_variance=np.empty((nf, nbins)) if errors else None,
                      count=np.empty(nbins))
            R = N.RawSums(*[None if rs[k] is None else rs[k].ctypes.data
                            for k in ("sum_signal", "sum_norm", "sum_variance", "count")])
        N.check(N.lib().dio_multi_integrate(
            nd, P, F, D, V, VD, FA, nf, N.WITH_ERRORS if errors else 0, I.ctypes.data,
            None if S is None else S.ctypes.data, C.byref(R) if raw else None))
        return I, S, rs

    # -- 1D
    def _plans1d(self, shapes, n_points):
        return [it._plan(shape, n_points) for it, shape in zip(self.integrators, shapes)]

    def integrate1d(self, images, n_points: int, errors: bool = False) -> IntegrationResult:
        """MultiIntegrator.integrate1d (multi.rs:346-418): intensity =
        sum_d sum_signal / sum_d sum_norm; sigma = sqrt(max(sum_d var /
        (sum_d norm)^2, 0)); 0 where the merged norm is 0."""
        if n_points <= 0:
            raise ValueError("n_points must be positive")
        imgs = self._load(images)
        plans = self._plans1d([a.shape for a in imgs], n_points)
        I, S, _ = self._merge(plans, [a[None] for a in imgs], errors=errors)
        return IntegrationResult(plans[0].radial, I[0], None if S is None else S[0])

    def batch1d(self, images, n_points: int, errors: bool = False) -> list:
        """One merged pattern per frame set [[det0, det1, ...], ...]
        (python/multi_integrator.rs:190-225); frame sets of the same shapes
        and dtypes go to the device together."""
        if n_points <= 0:
            raise ValueError("n_points must be positive")
        sets = []
        n_det = len(self.integrators)
        for k, s in enumerate(images):
            if isinstance(s, (str, bytes, bytearray, os.PathLike)) or len(s) != n_det:
                got = 1 if isinstance(s, (str, bytes, bytearray, os.PathLike)) else len(s)
                raise ValueError(f"Frame set {k} has {got} images, expected {n_det} (one per "
                                 f"detector)")
            sets.append(self._load(s, f"frame set {k}, "))
        out: list = [None] * len(sets)
        groups: dict = {}
        for k, imgs in enumerate(sets):
            groups.setdefault(tuple((a.shape, a.dtype.str) for a in imgs), []).append(k)
        for key, members in groups.items():
            plans = self._plans1d([shape for shape, _ in key], n_points)
            stacks = [np.stack([sets[k][d] for k in members]) for d in range(n_det)]
            I, S, _ = self._merge(plans, stacks, errors=errors)
            for j, k in enumerate(members):
                out[k] = IntegrationResult(plans[0].radial, I[j], None if S is None else S[j])
        return out

    # -- 2D
    def grid(self, n_radial: int, n_azimuthal: int, azimuthal_range=None) -> IntegrationGrid2d:
        """The shared grid of integrate2d (multi.rs:216-225): the common radial
        range, azimuthal edges (default -pi, pi), this merger's splitting
        convention."""
        ar = azimuthal_range if azimuthal_range is not None else (-math.pi, math.pi)
        return IntegrationGrid2d(n_radial, n_azimuthal, self.integrators[0].radial_range,
                                 ar).with_pixel_split_mode(self.pixel_split_mode)

    def integrate2d(self, images, n_radial: int, n_azimuthal: Optional[int] = None,
                    errors: bool = False, azimuthal_range=None, *,
                    normalization: str = "relative", normalization_factors=None,
                    azimuthal_bracket: Optional[float] = None,
                    variances=None, raw: bool = False) -> IntegrationResult:
        """MultiIntegrator.integrate2d (multi.rs:198-316,
        python/multi_integrator.rs:39-102): per detector raw sums on the
        shared grid, merged with sum_norm scaled by each detector's factor
        (monitor normalization_factors[d], times pixel1 * pixel2 /
        distance^2 with normalization="absolute"); intensity = S / N,
        sigma = sqrt(max(V, 0)) / N, 0 where N == 0; intensity and errors
        shaped (n_azimuthal, n_radial). variances (one plane per detector,
        finite and nonnegative) replace the Poisson variance
        (integrate2d_from_slices_with_variances, multi.rs:174-196); raw=True
        also returns the merged sums in `raw`."""
        if n_radial <= 0:
            raise ValueError("n_radial must be positive")
        n_az, az_range = resolve_azimuthal_bins(n_azimuthal, azimuthal_range, azimuthal_bracket)
        if n_az <= 0:
            raise ValueError("n_azimuthal must be positive")
        if normalization not in ("relative", "absolute"):
            raise ValueError("normalization must be 'relative' or 'absolute'")
        imgs = self._load(images)
        grid = self.grid(n_radial, n_az, az_range)
        if any(it.radial_range != grid.radial_range for it in self.integrators):
            raise ValueError("All integrators must have identical finite radial_range edges for "
                             "2D integration")
        nd = len(self.integrators)
        if variances is not None:
            if not errors:
                raise ValueError("explicit variance integration requires errors and one "
                                 "variance plane per detector")
            variances = list(variances)
            if len(variances) != nd:
                raise ValueError("explicit variance integration requires errors and one "
                                 "variance plane per detector")
            vs = []
            for i, (v, a) in enumerate(zip(variances, imgs)):
                v = np.ascontiguousarray(v)
                if v.shape != a.shape or v.dtype not in _NP_DTYPES:
                    raise ValueError(f"detector {i}: pixel variances must match the image and "
                                     f"be finite and nonnegative")
                vs.append(v[None])
            variances = vs
        if normalization_factors is not None:
            normalization_factors = [float(f) for f in normalization_factors]
            if len(normalization_factors) != nd or any(
                    not math.isfinite(f) or f <= 0.0 for f in normalization_factors):
                raise ValueError("normalization_factors must contain one finite positive "
                                 "monitor value per detector")
        factors = []
        for i, it in enumerate(self.integrators):
            monitor = 1.0 if normalization_factors is None else normalization_factors[i]
            p = it.poni
            sac = 1.0 if normalization == "relative" else \
                (p["pixel1"] / p["distance"]) * (p["pixel2"] / p["distance"])
            f = monitor * sac
            if not math.isfinite(f) or f <= 0.0:
                raise ValueError(f"detector {i}: normalization factor underflows or overflows")
            factors.append(f)
        plans = [it.grid_plan(a.shape, grid) for it, a in zip(self.integrators, imgs)]
        I, S, rs = self._merge(plans, [a[None] for a in imgs], variances, factors, errors, raw)
        shape = (n_az, n_radial)
        merged = None
        if raw:
            merged = {k: (v[0] if k in ("sum_signal", "sum_variance") and v is not None else v)
                      for k, v in rs.items()}
        return IntegrationResult(grid.radial_centers(), I[0].reshape(shape),
                                 None if S is None else S[0].reshape(shape), None,
                                 grid.azimuthal_centers(), n_radial, n_az, merged)
