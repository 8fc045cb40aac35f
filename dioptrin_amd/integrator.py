"""Python front end mirroring the reference's Integrator / Cuda1dPlan API.

Reference interfaces (paths relative to the reference repository root):
  Integrator(...)                dioptrin.pyi:337-460, src/python/integrator.rs
  Integrator.integrate1d         dioptrin.pyi:597-636, src/integration/methods.rs:24-190
  Integrator.integrate2d         dioptrin.pyi:660-733
  Integrator.batch1d / batch2d   dioptrin.pyi:845-1063, src/python/integrator.rs:1867, 2423
  Integrator.prepare_cuda1d/2d   dioptrin.pyi:490-560, src/integration/cuda.rs:1089-1157
  Cuda1dPlan.integrate1d/batch1d dioptrin.pyi:1358-1433

Every computation runs in libdioptrin_hip.so on the GPU; this module only
marshals arguments, caches plans and wraps results.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence, Union

import numpy as np

from . import _native as N
from .grid import IntegrationGrid2d, IntermediateResult1d, IntermediateResult2d
from .poni import load_poni  # noqa: F401  (part of the package API)
from .tiff import read_tiff

_NP_DTYPES = {np.dtype(np.uint16): "uint16", np.dtype(np.uint32): "uint32",
              np.dtype(np.int32): "int32", np.dtype(np.float32): "float32",
              np.dtype(np.float64): "float64"}


@dataclass
class IntegrationResult:
    """Result of one frame (python/results.rs: radial, intensity, errors,
    azimuthal_std; 2D adds azimuthal and the grid shape)."""
    radial: Optional[np.ndarray]
    intensity: np.ndarray
    errors: Optional[np.ndarray] = None
    azimuthal_std: Optional[np.ndarray] = None
    azimuthal: Optional[np.ndarray] = None
    n_radial: Optional[int] = None
    n_azimuthal: Optional[int] = None
    raw: Optional[dict] = None  # merged sums before division (MultiIntegrator, raw=True)


@dataclass
class BatchResult:
    """One entry of a batch (python BatchResult: result or error per item)."""
    filename: str
    result: Optional[IntegrationResult] = None
    error: Optional[str] = None

    def is_ok(self) -> bool:
        return self.error is None

    def is_err(self) -> bool:
        return self.error is not None


class _Context:
    _cache: dict = {}

    @classmethod
    def get(cls, device: int):
        if device not in cls._cache:
            h = C.c_void_p()
            N.check(N.lib().dio_open(device, C.byref(h)))
            cls._cache[device] = h
        return cls._cache[device]


def device_info(device: int = 0) -> dict:
    """Device description (cuda_info counterpart, cuda.rs:309-326)."""
    info = N.DeviceInfo()
    N.check(N.lib().dio_device_info_get(_Context.get(device), C.byref(info)))
    return {"name": info.name.decode(), "arch": info.arch.decode(),
            "compute_units": info.compute_units, "xcds": info.xcds,
            "total_memory_bytes": info.total_memory, "runtime_version": info.runtime_version}


def plan_args(poni: dict, shape, method="pixel_split", n_ss=None, unit="2th_rad",
              n_radial=1000, n_azimuthal=None, radial_range=None, azimuthal_range=None,
              polarization_factor=None, mask=None, precision="f32_compensated",
              input_order="detector"):
    """The (dio_poni, dio_plan_desc, mask array) triple naming one plan. The
    mask array must stay alive while the descriptor is in use. input_order
    "flip_vertical" (InputOrder::FlipVertical, fixed.rs:20-24) reads frames
    whose rows arrive bottom-up."""
    if method not in N.METHODS:
        raise ValueError(f"unknown method {method!r}")
    if unit not in N.UNITS:
        raise ValueError(f"unknown unit {unit!r}")
    if precision not in N.PRECISIONS:
        raise ValueError(f"unknown precision {precision!r}; expected one of "
                         f"{sorted(N.PRECISIONS)}")
    if input_order not in N.ORDERS:
        raise ValueError(f"unknown input order {input_order!r}; expected one of "
                         f"{sorted(N.ORDERS)}")
    dim1, dim2 = int(shape[0]), int(shape[1])
    p = N.Poni(*(float(poni.get(k, 0.0)) for k in (
        "pixel1", "pixel2", "distance", "poni1", "poni2", "rot1", "rot2", "rot3",
        "wavelength")))
    d = N.PlanDesc()
    d.dim1, d.dim2 = dim1, dim2
    d.method = N.METHODS[method]
    d.n_ss = int(n_ss or (5 if method == "supersampled" else 0))
    d.unit = N.UNITS[unit]
    d.ndim = 1 if n_azimuthal is None else 2
    d.n_radial = int(n_radial)
    d.n_azimuthal = int(n_azimuthal or 1)
    if radial_range is not None:
        d.has_radial_range = 1
        d.radial_min, d.radial_max = map(float, radial_range)
    if azimuthal_range is not None:
        d.has_azimuthal_range = 1
        d.azimuthal_min, d.azimuthal_max = map(float, azimuthal_range)
    if polarization_factor is not None:
        d.has_polarization = 1
        d.polarization_factor = float(polarization_factor)
    m = None
    if mask is not None:
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        if m.shape != (dim1, dim2):
            raise ValueError(f"mask shape {m.shape} does not match image shape {(dim1, dim2)}")
        d.mask = m.ctypes.data
    d.precision = N.PRECISIONS[precision]
    d.input_order = N.ORDERS[input_order]
    return p, d, m


class GpuPlan:
    """A lookup table built and kept on one GPU (Cuda1dPlan / Cuda2dPlan
    counterpart). A snapshot: later Integrator changes do not affect it."""

    def __init__(self, poni: dict, shape, method="pixel_split", n_ss=None, unit="2th_rad",
                 n_radial=1000, n_azimuthal=None, radial_range=None, azimuthal_range=None,
                 polarization_factor=None, mask=None, precision="f32_compensated", device=0,
                 input_order="detector"):
        self.shape = (int(shape[0]), int(shape[1]))
        self.input_order = input_order
        self.ndim = 1 if n_azimuthal is None else 2
        self.precision = precision
        self.device = device
        p, d, self._mask = plan_args(poni, shape, method, n_ss, unit, n_radial, n_azimuthal,
                                     radial_range, azimuthal_range, polarization_factor, mask,
                                     precision, input_order)
        h = C.c_void_p()
        N.check(N.lib().dio_plan_create(_Context.get(device), C.byref(p), C.byref(d), C.byref(h)))
        self._h = h
        self._mask = None
        info = N.PlanInfo()
        N.check(N.lib().dio_plan_info_get(h, C.byref(info)))
        self.info = {f: getattr(info, f) for f, _ in N.PlanInfo._fields_}
        self.n_radial = info.n_radial
        self.n_azimuthal = info.n_azimuthal if self.ndim == 2 else None
        self.nbins = info.nbins
        self.radial = np.empty(self.n_radial)
        N.check(N.lib().dio_plan_radial(h, self.radial.ctypes.data))
        self.azimuthal = None
        if self.ndim == 2:
            self.azimuthal = np.empty(self.n_azimuthal)
            N.check(N.lib().dio_plan_azimuthal(h, self.azimuthal.ctypes.data))

    # -- properties mirroring Cuda1dPlan
    @property
    def n_points(self):
        return self.n_radial

    @property
    def table_entries(self):
        return self.info["nnz_masked"]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().dio_plan_destroy(h)
            self._h = None

    # -- table access for parity checks
    def export_csr(self, masked=False):
        nnz = self.info["nnz_masked"] if masked else self.info["nnz"]
        rp = np.empty(self.nbins + 1, np.uint32)
        col = np.empty(max(nnz, 1), np.uint32)
        w64 = np.empty(max(nnz, 1), np.float64)
        w32 = np.empty(max(nnz, 1), np.float32)
        N.check(N.lib().dio_plan_export_csr(self._h, int(masked), rp.ctypes.data, col.ctypes.data,
                                            w64.ctypes.data, w32.ctypes.data, max(nnz, 1)))
        return rp, col[:nnz], w64[:nnz], w32[:nnz]

    def denominators(self):
        out = np.empty(self.nbins)
        N.check(N.lib().dio_plan_denominators(self._h, out.ctypes.data))
        return out

    def normalization(self):
        out = np.empty(self.shape[0] * self.shape[1])
        N.check(N.lib().dio_plan_normalization(self._h, out.ctypes.data))
        return out

    def counts(self):
        """Per bin: the sum of the execution table's weights (count)."""
        out = np.empty(self.nbins)
        N.check(N.lib().dio_plan_counts(self._h, out.ctypes.data))
        return out

    def integrate_raw(self, stack, errors=False) -> dict:
        """Sums before division for multi-detector merging (integrate1d_raw /
        IntermediateResult1d, lut/integrate.rs:372-500): per frame and bin
        sum_signal = sum w*x and, with errors, sum_variance = sum w^2*max(x, 0);
        per bin sum_norm (the denominators) and count."""
        stack = self._stack(stack)
        n = stack.shape[0]
        flags = N.RAW_SUMS | (N.WITH_ERRORS if errors else 0)
        sig = np.empty((n, self.nbins))
        var = np.empty((n, self.nbins)) if errors else None
        if n:
            N.check(N.lib().dio_integrate(self._h, stack.ctypes.data, N.DTYPES[_NP_DTYPES[
                stack.dtype]], n, flags, sig.ctypes.data,
                None if var is None else var.ctypes.data, None))
        return {"bins": self.radial, "sum_signal": sig, "sum_norm": self.denominators(),
                "sum_variance": var, "count": self.counts()}

    def integrate_with_variances(self, stack, variances, raw=False):
        """Intensity and sigma from supplied pixel variances
        (dio_integrate_variances; Prepared1d::integrate_with_variances,
        lut/prepared.rs:737-766): per bin sum w*x / den and
        sqrt(sum w^2*v) / den, or with raw=True the sums sum w*x and
        sum w^2*v. variances: one finite, nonnegative plane per frame."""
        stack = self._stack(stack)
        var = np.asarray(variances)
        if var.ndim == 2:
            var = var[None]
        var = np.ascontiguousarray(var)
        if var.shape != stack.shape:
            raise ValueError(f"variances have shape {var.shape}, frames {stack.shape}")
        if var.dtype not in _NP_DTYPES:
            raise TypeError(f"unsupported variance dtype {var.dtype}")
        n = stack.shape[0]
        I = np.empty((n, self.nbins))
        S = np.empty((n, self.nbins))
        if n:
            N.check(N.lib().dio_integrate_variances(
                self._h, stack.ctypes.data, N.DTYPES[_NP_DTYPES[stack.dtype]], var.ctypes.data,
                N.DTYPES[_NP_DTYPES[var.dtype]], n, N.RAW_SUMS if raw else 0, I.ctypes.data,
                S.ctypes.data))
        return I, S

    # -- Bitshuffle + LZ4 chunks decoded on the device
    def integrate_bslz4(self, chunks, dtype="uint32", errors=False, raw=False):
        """(I, S) of raw HDF5 Bitshuffle+LZ4 chunks, one per frame, decoded
        and integrated on the GPU (Cuda1dPlan::integrate_bslz4,
        cuda.rs:550-564): the bits of integrate_stack on the decoded frames."""
        from . import bslz4
        return bslz4.integrate(self, chunks, dtype, errors, raw)

    def batch1d_bslz4(self, chunks, dtype="uint32", *, errors=False) -> list:
        """One IntegrationResult per chunk (Cuda1dPlan.batch1d_bslz4)."""
        I, S = self.integrate_bslz4(chunks, dtype, errors)
        return [self._result(I[k], None if S is None else S[k], None) for k in range(I.shape[0])]

    batch2d_bslz4 = batch1d_bslz4

    def batch2d_bslz4_stack(self, chunks, dtype="uint32", *, errors=False):
        """(n_frames, n_azimuthal, n_radial) intensity (and errors) of a 2D
        plan (Cuda2dPlan.batch2d_bslz4_stack)."""
        if self.ndim != 2:
            raise ValueError("batch2d_bslz4_stack needs a 2D plan")
        I, S = self.integrate_bslz4(chunks, dtype, errors)
        shape = (I.shape[0], self.n_azimuthal, self.n_radial)
        return (I.reshape(shape), S.reshape(shape)) if errors else I.reshape(shape)

    # -- integration
    def _stack(self, images) -> np.ndarray:
        if isinstance(images, np.ndarray) and images.ndim == 3:
            arr = images
        else:
            images = list(images)
            if not images:
                return np.empty((0,) + self.shape, np.float32)
            dts = {np.asarray(im).dtype for im in images}
            if len(dts) != 1:
                raise TypeError("all images of one call must have the same dtype")
            arr = np.stack([np.asarray(im) for im in images])
        if arr.dtype not in _NP_DTYPES:
            raise TypeError(f"unsupported dtype {arr.dtype}; expected float64, float32, "
                            f"uint16, uint32 or int32")
        if arr.shape[1:] != self.shape:
            raise ValueError(f"frames have shape {arr.shape[1:]}, but the plan expects "
                             f"{self.shape}")
        return np.ascontiguousarray(arr)

    def _check_device(self, frames, out_i, out_s=None, out_d=None, errors=False,
                      azimuthal_std=False, masks=None, outputs=True) -> str:
        """Host-side checks of device buffers before the C ABI trusts their
        sizes: frames (n, dim1, dim2) of a supported dtype, outputs float64
        (n, nbins), sigma present with errors and the spread with
        azimuthal_std, masks uint8 of the frames' shape. Returns the dtype
        name."""
        dt = np.dtype(frames.dtype)
        if dt not in _NP_DTYPES:
            raise TypeError(f"unsupported dtype {dt}; expected float64, float32, uint16, "
                            f"uint32 or int32")
        if len(frames.shape) != 3 or tuple(frames.shape[1:]) != self.shape:
            raise ValueError(f"frames have shape {tuple(frames.shape)}, but the plan expects "
                             f"(n,) + {self.shape}")
        n = frames.shape[0]
        want = (n, self.nbins)
        for name, buf, needed in (("intensity", out_i, outputs), ("sigma", out_s, errors),
                                  ("azimuthal std", out_d, azimuthal_std)):
            if buf is None:
                if needed:
                    raise ValueError(f"{name} output buffer required")
                continue
            if np.dtype(buf.dtype) != np.float64 or tuple(buf.shape) != want:
                raise ValueError(f"{name} output must be float64 of shape {want}, got "
                                 f"{np.dtype(buf.dtype)} {tuple(buf.shape)}")
        if masks is not None and (tuple(masks.shape) != tuple(frames.shape)
                                  or np.dtype(masks.dtype) != np.uint8):
            raise ValueError("masks must be uint8 of the frames' shape")
        return _NP_DTYPES[dt]

    def integrate_stack(self, stack: np.ndarray, errors=False, azimuthal_std=False):
        """Raw arrays for a (n, dim1, dim2) stack: intensity, sigma, std."""
        stack = self._stack(stack)
        n = stack.shape[0]
        flags = (N.WITH_ERRORS if errors else 0) | (N.WITH_AZIMUTHAL_STD if azimuthal_std else 0)
        I = np.empty((n, self.nbins))
        S = np.empty((n, self.nbins)) if errors else None
        D = np.empty((n, self.nbins)) if azimuthal_std else None
        if n:
            N.check(N.lib().dio_integrate(self._h, stack.ctypes.data, N.DTYPES[_NP_DTYPES[
                stack.dtype]], n, flags, I.ctypes.data, None if S is None else S.ctypes.data,
                None if D is None else D.ctypes.data))
        return I, S, D

    def integrate_masked(self, stack, masks, errors=False):
        """Per-frame (dynamic) masks, Dynamic1dPlan::integrate
        (integration/dynamic.rs:143-186): frame k is integrated on the
        unmasked table with masks[k] (nonzero = masked) instead of the plan's
        mask. Returns intensity and sigma, (n, nbins) each (sigma None
        without errors)."""
        stack = self._stack(stack)
        n = stack.shape[0]
        masks = np.ascontiguousarray(masks, dtype=np.uint8)
        if masks.shape != stack.shape:
            raise ValueError(f"masks of shape {masks.shape} for frames {stack.shape}")
        I = np.empty((n, self.nbins))
        S = np.empty((n, self.nbins)) if errors else None
        N.check(N.lib().dio_integrate_masked(
            self._h, stack.ctypes.data, N.DTYPES[_NP_DTYPES[stack.dtype]], n, masks.ctypes.data,
            N.WITH_ERRORS if errors else 0, I.ctypes.data, None if S is None else S.ctypes.data))
        return I, S

    def integrate_masked_device(self, frames, masks, out_i, out_s=None, errors=False,
                                stream=None):
        """Device variant of integrate_masked: frames (n, dim1, dim2) and masks
        (n, dim1, dim2) uint8 DeviceArrays, (n, nbins) float64 outputs;
        enqueued on `stream` (None = the plan's stream)."""
        dt = self._check_device(frames, out_i, out_s, None, errors, False, masks)
        N.check(N.lib().dio_integrate_masked_device(
            self._h, frames.ptr, N.DTYPES[dt], frames.shape[0], masks.ptr,
            N.WITH_ERRORS if errors else 0, out_i.ptr, None if out_s is None else out_s.ptr,
            stream))

    def _result(self, i, s, d):
        if self.ndim == 2:
            shape = (self.n_azimuthal, self.n_radial)
            return IntegrationResult(self.radial, i.reshape(shape),
                                     None if s is None else s.reshape(shape), None,
                                     self.azimuthal, self.n_radial, self.n_azimuthal)
        return IntegrationResult(self.radial, i, s, d)

    def integrate1d(self, image, *, errors=False, azimuthal_std=False) -> IntegrationResult:
        I, S, D = self.integrate_stack(np.asarray(image)[None], errors, azimuthal_std)
        return self._result(I[0], None if S is None else S[0], None if D is None else D[0])

    integrate2d = integrate1d

    def batch1d(self, images, *, errors=False, azimuthal_std=False) -> list:
        I, S, D = self.integrate_stack(images, errors, azimuthal_std)
        return [self._result(I[k], None if S is None else S[k], None if D is None else D[k])
                for k in range(I.shape[0])]

    batch2d = batch1d

    # -- device-resident path (DeviceArray frames and outputs)
    def integrate_device(self, frames, out_i, out_s=None, out_d=None, errors=False,
                         azimuthal_std=False, stream=None):
        """Enqueues integration of device frames (DeviceArray (n, dim1, dim2))
        into device outputs ((n, nbins) float64 DeviceArrays) on `stream`
        (a hipStream_t handle; None = the plan's stream)."""
        dt = self._check_device(frames, out_i, out_s, out_d, errors, azimuthal_std)
        flags = (N.WITH_ERRORS if errors else 0) | (N.WITH_AZIMUTHAL_STD if azimuthal_std else 0)
        N.check(N.lib().dio_integrate_device(
            self._h, frames.ptr, N.DTYPES[dt], frames.shape[0], flags, out_i.ptr,
            None if out_s is None else out_s.ptr, None if out_d is None else out_d.ptr, stream))

    def integrate_device_noalloc(self, frames, out_i, out_s=None, out_d=None, errors=False,
                                 azimuthal_std=False, raw=False, stream=None):
        """integrate_device through dio_integrate_device_noalloc, which
        allocates no device memory (its partials use a fixed 64 MiB array
        per device); same results bit for bit. Calls that overlap on one
        device race."""
        dt = self._check_device(frames, out_i, out_s, out_d, errors, azimuthal_std)
        flags = ((N.WITH_ERRORS if errors else 0) | (N.WITH_AZIMUTHAL_STD if azimuthal_std else 0)
                 | (N.RAW_SUMS if raw else 0))
        N.check(N.lib().dio_integrate_device_noalloc(
            self._h, frames.ptr, N.DTYPES[dt], frames.shape[0], flags, out_i.ptr,
            None if out_s is None else out_s.ptr, None if out_d is None else out_d.ptr, stream))

    def autotune(self, dtype="uint32", n_frames=64, errors=False):
        """Picks and stores (frames per pass, frames per lane); returns them."""
        t, g = C.c_uint32(), C.c_uint32()
        N.check(N.lib().dio_plan_autotune(self._h, N.DTYPES[dtype], n_frames,
                                          N.WITH_ERRORS if errors else 0, C.byref(t),
                                          C.byref(g)))
        return t.value, g.value

    def set_tile(self, frames_per_tile: int, frames_per_lane: int = 0):
        N.check(N.lib().dio_plan_set_tile(self._h, frames_per_tile, frames_per_lane))

    def time_gather(self, frames, errors=False, reps=20) -> float:
        """Average ms of one launch of the segment kernel over one tile."""
        dt = self._check_device(frames, None, outputs=False)
        ms = C.c_double()
        N.check(N.lib().dio_plan_time_gather(self._h, frames.ptr, N.DTYPES[dt],
                                             frames.shape[0], N.WITH_ERRORS if errors else 0,
                                             None, None, reps, C.byref(ms)))
        return ms.value

    def time_passes(self, frames, out_i, out_s=None, errors=False) -> list:
        """(duration in ms, frames) of each gather-kernel launch of one
        device-resident integration of the batch, timed with HIP events on
        the plan's stream (a launch covers one pass, or two passes sharing
        the table)."""
        dt = self._check_device(frames, out_i, out_s, None, errors)
        cap = frames.shape[0]  # at least one frame per launch
        buf = (C.c_float * cap)()
        nfr = (C.c_uint32 * cap)()
        n = C.c_uint32()
        N.check(N.lib().dio_plan_time_passes(
            self._h, frames.ptr, N.DTYPES[dt], frames.shape[0], N.WITH_ERRORS if errors else 0,
            out_i.ptr, None if out_s is None else out_s.ptr, buf, nfr, cap, C.byref(n)))
        k = min(cap, n.value)
        return list(zip(list(buf)[:k], list(nfr)[:k]))

    def capture_graph(self, frames, out_i, out_s=None, out_d=None, errors=False,
                      azimuthal_std=False):
        return GpuGraph(self, frames, out_i, out_s, out_d, errors, azimuthal_std)


class GpuGraph:
    """hipGraph of one device-resident integration, replayed by launch()."""

    def __init__(self, plan: GpuPlan, frames, out_i, out_s, out_d, errors, azimuthal_std):
        dt = plan._check_device(frames, out_i, out_s, out_d, errors, azimuthal_std)
        flags = (N.WITH_ERRORS if errors else 0) | (N.WITH_AZIMUTHAL_STD if azimuthal_std else 0)
        h = C.c_void_p()
        N.check(N.lib().dio_graph_create(
            plan._h, frames.ptr, N.DTYPES[dt], frames.shape[0], flags, out_i.ptr,
            None if out_s is None else out_s.ptr, None if out_d is None else out_d.ptr,
            C.byref(h)))
        self._h, self.plan = h, plan
        self._keep = (frames, out_i, out_s, out_d)

    def launch(self, stream=None):
        N.check(N.lib().dio_graph_launch(self._h, stream))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().dio_graph_destroy(h)
            self._h = None


class Integrator:
    """Radial integration configured like the reference Integrator.

    Every integration builds (once per shape and key, like the reference's
    single-entry caches, cache.rs:42-96) a GpuPlan and runs on the GPU.
    """

    def __init__(self, poni_path: Union[str, dict], method: Optional[str] = None,
                 polarization_factor: Optional[float] = None, n_ss: Optional[int] = None,
                 unit: Optional[str] = None, radial_range=None, *,
                 precision: str = "f32_compensated", device: int = 0):
        self.poni = dict(poni_path) if isinstance(poni_path, dict) else load_poni(poni_path)
        self.method = method or "pixel_split"
        if self.method not in N.METHODS:
            raise ValueError(f"unknown method {self.method!r}")
        if n_ss is not None and n_ss <= 0:
            raise ValueError("n_ss must be positive")
        self.n_ss = n_ss if n_ss is not None else 5
        self.unit = unit or "2th_rad"
        if self.unit not in N.UNITS:
            raise ValueError(f"unknown unit {self.unit!r}")
        if polarization_factor is not None and not -1.0 <= polarization_factor <= 1.0:
            raise ValueError("polarization_factor must lie in [-1, 1]")
        self.polarization_factor = polarization_factor
        if radial_range is not None:
            lo, hi = map(float, radial_range)
            if not (np.isfinite(lo) and np.isfinite(hi) and lo < hi):
                raise ValueError("radial_range must be finite with min < max")
            radial_range = (lo, hi)
        self.radial_range = radial_range
        self.precision = precision
        self.device = device
        self.mask = None
        self._plans: dict = {}

    @classmethod
    def from_poni_dict(cls, poni_dict: dict, *args, **kwargs):
        return cls(poni_dict, *args, **kwargs)

    # -- setters mirroring the reference (each invalidates cached plans)
    def set_method(self, method: str, n_ss: Optional[int] = None):
        if method not in N.METHODS:
            raise ValueError(f"unknown method {method!r}")
        self.method = method
        if n_ss is not None:
            self.n_ss = n_ss
        self._plans.clear()

    def set_polarization_factor(self, factor: Optional[float]):
        self.polarization_factor = factor
        self._plans.clear()

    def set_unit(self, unit: str):
        if unit not in N.UNITS:
            raise ValueError(f"unknown unit {unit!r}")
        self.unit = unit
        self._plans.clear()

    def set_mask(self, mask: Optional[np.ndarray]):
        self.mask = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        self._plans.clear()

    def clear_mask(self):
        self.set_mask(None)

    def get_unit(self) -> str:
        return self.unit

    def _plan(self, shape, n_radial, n_azimuthal=None, azimuthal_range=None, precision=None,
              *, radial_range=None, method=None):
        rr = radial_range if radial_range is not None else self.radial_range
        method = method or self.method
        key = (tuple(shape), n_radial, n_azimuthal, azimuthal_range, precision or self.precision,
               rr, method)
        plan = self._plans.get(key)
        if plan is None:
            if self.mask is not None and self.mask.shape != tuple(shape):
                raise ValueError(f"image shape {tuple(shape)} does not match mask shape "
                                 f"{self.mask.shape}")
            plan = GpuPlan(self.poni, shape, method, self.n_ss, self.unit, n_radial,
                           n_azimuthal, rr, azimuthal_range,
                           self.polarization_factor, self.mask, precision or self.precision,
                           self.device)
            self._plans = {key: plan}  # single entry, like the reference caches
        return plan

    def grid_plan(self, shape, grid: IntegrationGrid2d, precision=None) -> "GpuPlan":
        """The plan of a shared grid (Integrator::shared_grid_key,
        grid.rs:406-428): the grid's edges, clamped index rounding and, for
        pixel splitting, the grid's splitting convention (pyFAI's full
        splitting for pyfai_full integrators whatever the grid says,
        grid.rs:366-371)."""
        method = self.method
        if method == "pyfai_full" or (method == "pixel_split" and
                                      grid.pixel_split_mode == "pyfai_histogram"):
            method = "pyfai_full"
        return self._plan(tuple(shape), grid.n_radial, grid.n_azimuthal, grid.azimuthal_range,
                          precision, radial_range=grid.radial_range, method=method)

    @staticmethod
    def _image2d(image) -> np.ndarray:
        a = np.asarray(image)
        if a.ndim != 2:
            raise ValueError(f"image must be 2D, got {a.ndim} dimensions")
        if a.size == 0:
            raise ValueError("image is empty")
        return a

    # -- sums before division (multi-detector merging)
    def integrate1d_raw(self, image, n_points: int) -> IntermediateResult1d:
        """integrate1d_raw (integration/methods.rs:318-345): sum_signal,
        sum_norm (the denominators) and count per bin, without dividing."""
        return self._raw1d(image, n_points, False)

    def integrate1d_raw_with_variance(self, image, n_points: int) -> IntermediateResult1d:
        """integrate1d_raw_with_variance: adds sum_variance =
        sum w^2 * max(x, 0) (Poisson)."""
        return self._raw1d(image, n_points, True)

    def _raw1d(self, image, n_points, variance):
        if n_points <= 0:
            raise ValueError("n_points must be positive")
        image = self._image2d(image)
        r = self._plan(image.shape, n_points).integrate_raw(image[None], errors=variance)
        return IntermediateResult1d(r["bins"], r["sum_signal"][0], r["sum_norm"],
                                    None if r["sum_variance"] is None else r["sum_variance"][0],
                                    r["count"])

    def integrate2d_raw_on_grid(self, image, grid: IntegrationGrid2d,
                                errors: bool = False) -> IntermediateResult2d:
        """integrate2d_raw_on_grid_from_slice (grid.rs:327-341, 358-392): raw
        sums on explicit shared edges; errors adds the Poisson sum_variance."""
        image = self._image2d(image)
        plan = self.grid_plan(image.shape, grid)
        r = plan.integrate_raw(image[None], errors=errors)
        return IntermediateResult2d(r["sum_signal"][0], r["sum_norm"],
                                    None if r["sum_variance"] is None else r["sum_variance"][0],
                                    r["count"], grid.n_radial, grid.n_azimuthal)

    def integrate2d_raw_on_grid_with_variances(self, image, variance,
                                               grid: IntegrationGrid2d) -> IntermediateResult2d:
        """integrate2d_raw_on_grid_with_variances_from_slice (grid.rs:343-356):
        sum_variance = sum w^2 * v of the supplied (finite, nonnegative)
        variances, which must match the image."""
        image = self._image2d(image)
        variance = np.asarray(variance)
        if variance.shape != image.shape:
            raise ValueError("pixel variances must match the image and be finite and "
                             "nonnegative")
        plan = self.grid_plan(image.shape, grid)
        sig, var = plan.integrate_with_variances(image[None], variance[None], raw=True)
        return IntermediateResult2d(sig[0], plan.denominators(), var[0], plan.counts(),
                                    grid.n_radial, grid.n_azimuthal)

    def prepare_gpu1d(self, shape, n_points: int, *, precision: str = "f32_compensated",
                      device: Optional[int] = None, input_order: str = "detector") -> GpuPlan:
        """prepare_cuda1d counterpart: an independent plan snapshot
        (input_order as Cuda1dOptions::with_input_order, cuda.rs:174-181)."""
        return GpuPlan(self.poni, shape, self.method, self.n_ss, self.unit, n_points, None,
                       self.radial_range, None, self.polarization_factor, self.mask, precision,
                       self.device if device is None else device, input_order)

    def prepare_gpu2d(self, shape, n_radial: int, n_azimuthal: int, *, azimuthal_range=None,
                      precision: str = "f32_compensated", device: Optional[int] = None,
                      input_order: str = "detector") -> GpuPlan:
        return GpuPlan(self.poni, shape, self.method, self.n_ss, self.unit, n_radial,
                       n_azimuthal, self.radial_range, azimuthal_range, self.polarization_factor,
                       self.mask, precision, self.device if device is None else device,
                       input_order)

    def integrate1d(self, image: np.ndarray, n_points: int, errors: bool = False, *,
                    azimuthal_std: bool = False) -> IntegrationResult:
        if n_points <= 0:
            raise ValueError("n_points must be positive")
        image = np.asarray(image)
        return self._plan(image.shape, n_points).integrate1d(image, errors=errors,
                                                             azimuthal_std=azimuthal_std)

    def integrate1d_with_errors(self, image, n_points):
        return self.integrate1d(image, n_points, True)

    def integrate1d_with_azimuthal_std(self, image, n_points, errors=False):
        return self.integrate1d(image, n_points, errors, azimuthal_std=True)

    def integrate2d(self, image: np.ndarray, n_radial: int, n_azimuthal: int,
                    errors: bool = False, azimuthal_range=None) -> IntegrationResult:
        image = np.asarray(image)
        return self._plan(image.shape, n_radial, n_azimuthal, azimuthal_range).integrate1d(
            image, errors=errors)

    def batch1d(self, images, n_points: int, errors: bool = False, num_workers=None,
                num_threads=None, *, azimuthal_std: bool = False) -> list:
        """batch1d (dioptrin.pyi:845-904, python/integrator.rs:1867-2210):
        a list of 2D arrays or of TIFF paths (or one 3D stack) -> one
        BatchResult per image, in input order, named "array[i]" or by path.
        Arrays that are empty, not C-contiguous or of another shape than the
        mask raise ValueError, and unsupported dtypes TypeError, for the
        whole call; a file that cannot be read or does not match the mask
        reports its error in its own BatchResult. num_workers/num_threads
        size the reference's CPU pools and are accepted for compatibility."""
        if n_points <= 0:
            raise ValueError("n_points must be positive")
        return self._batch(images, lambda shape: self._plan(shape, n_points), errors,
                           azimuthal_std)

    def batch2d(self, images, n_radial: int, n_azimuthal: int, errors: bool = False,
                azimuthal_range=None, num_workers=None, num_threads=None) -> list:
        """batch2d (python/integrator.rs:2423): batch1d's inputs and error
        rules, 2D results."""
        if n_radial <= 0 or n_azimuthal <= 0:
            raise ValueError("n_radial and n_azimuthal must be positive")
        return self._batch(images, lambda shape: self._plan(shape, n_radial, n_azimuthal,
                                                           azimuthal_range), errors, False)

    # Sample types the device reads directly, and exact widenings of the
    # others a TIFF may hold (the reference converts every sample to f64).
    _WIDEN = {np.dtype(np.uint8): np.uint16, np.dtype(np.int8): np.int32,
              np.dtype(np.int16): np.int32, np.dtype(np.uint64): np.float64,
              np.dtype(np.int64): np.float64}

    def _batch(self, images, plan_for, errors, azimuthal_std):
        if isinstance(images, (str, bytes, bytearray, os.PathLike)):
            raise TypeError("images must be a list of paths or arrays; wrap one path in a list")
        if isinstance(images, np.ndarray):
            if images.ndim != 3:
                raise TypeError("a single array must be a 3D stack (n_frames, dim1, dim2)")
            items = [images[i] for i in range(images.shape[0])]
        else:
            items = list(images)
        is_path = [isinstance(x, (str, os.PathLike)) for x in items]
        if any(is_path) and not all(is_path):
            raise TypeError("images must be all paths or all arrays")
        out: list = [None] * len(items)
        frames = []  # (index, name, array)
        for i, im in enumerate(items):
            if is_path[i]:
                name = os.fspath(im)
                try:
                    arr = read_tiff(name)
                except (OSError, ValueError) as e:
                    out[i] = BatchResult(name, None, str(e))
                    continue
                arr = arr.astype(self._WIDEN.get(arr.dtype, arr.dtype), copy=False)
                if self.mask is not None and arr.shape != self.mask.shape:
                    out[i] = BatchResult(name, None, f"image shape {arr.shape} does not match "
                                                     f"mask shape {self.mask.shape}")
                    continue
            else:
                if not isinstance(im, np.ndarray) or im.ndim != 2:
                    raise TypeError(f"images[{i}] must be a 2D numpy array")
                if im.dtype not in _NP_DTYPES:
                    raise TypeError(f"images[{i}] has unsupported dtype {im.dtype}; expected "
                                    f"float64, float32, uint16, uint32 or int32")
                if im.size == 0:
                    raise ValueError(f"images[{i}] is empty")
                if not im.flags["C_CONTIGUOUS"]:
                    raise ValueError(f"Array {i} must be contiguous")
                if self.mask is not None and im.shape != self.mask.shape:
                    raise ValueError(f"images[{i}] shape {im.shape} does not match mask "
                                     f"shape {self.mask.shape}")
                name, arr = f"array[{i}]", im
            frames.append((i, name, arr))
        groups: dict = {}
        for i, name, arr in frames:
            groups.setdefault((arr.shape, arr.dtype.str), []).append((i, name, arr))
        for (shape, _), members in groups.items():
            plan = plan_for(shape)
            res = plan.batch1d([a for _, _, a in members], errors=errors,
                               azimuthal_std=azimuthal_std)
            for (i, name, _), r in zip(members, res):
                out[i] = BatchResult(name, r, None)
        return out
