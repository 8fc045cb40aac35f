"""ctypes front end of the FP64 CPU oracle (test infrastructure only).

Only tests, the smoke check and the benchmark's CPU-baseline leg import this
module; the product package never does. See oracle.h for the reference
functions each entry restates.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

UNITS = {"2th_rad": 0, "2th_deg": 1, "q_A^-1": 2, "q_nm^-1": 3}


class _Poni(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "pixel1", "pixel2", "distance", "poni1", "poni2", "rot1", "rot2", "rot3", "wavelength")]


class _Lut(C.Structure):
    _fields_ = [
        ("nbins", C.c_size_t), ("nnz", C.c_size_t),
        ("row_ptr", C.POINTER(C.c_uint64)), ("col", C.POINTER(C.c_uint64)),
        ("w", C.POINTER(C.c_double)), ("hist_min", C.c_double), ("hist_step", C.c_double)]


def build():
    """Compile liboracle.so with its Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _setup(_lib)
    return _lib


def _setup(L):
    dp = C.POINTER(C.c_double)
    L.or_build_lut1d_ps.restype = C.c_int
    L.or_build_lut1d_counting.restype = C.c_int
    L.or_csc_new.restype = C.c_void_p
    L.or_csc_new.argtypes = [C.POINTER(_Lut), C.c_size_t, C.c_void_p, dp]
    L.or_csc_free.argtypes = [C.c_void_p]
    L.or_csc_batch_u32.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, dp]


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class Poni:
    pixel1: float
    pixel2: float
    distance: float
    poni1: float
    poni2: float
    rot1: float = 0.0
    rot2: float = 0.0
    rot3: float = 0.0
    wavelength: float = 0.0

    def c(self) -> _Poni:
        return _Poni(self.pixel1, self.pixel2, self.distance, self.poni1, self.poni2,
                     self.rot1, self.rot2, self.rot3, self.wavelength)


@dataclass
class Lut:
    row_ptr: np.ndarray
    col: np.ndarray
    w: Optional[np.ndarray]
    hist_min: float
    hist_step: float

    @property
    def nbins(self):
        return len(self.row_ptr) - 1

    @property
    def nnz(self):
        return len(self.col)

    def bins(self):
        out = np.empty(self.nbins)
        lib().or_bins(C.c_size_t(self.nbins), C.c_double(self.hist_min),
                      C.c_double(self.hist_step), _d(out))
        return out

    def _c(self):
        self._keep = (np.ascontiguousarray(self.row_ptr, np.uint64),
                      np.ascontiguousarray(self.col, np.uint64),
                      None if self.w is None else np.ascontiguousarray(self.w, np.float64))
        rp, co, w = self._keep
        return _Lut(self.nbins, self.nnz, rp.ctypes.data_as(C.POINTER(C.c_uint64)),
                    co.ctypes.data_as(C.POINTER(C.c_uint64)),
                    None if w is None else _d(w), self.hist_min, self.hist_step)


def _take(c: _Lut) -> Lut:
    n, nnz = c.nbins, c.nnz
    rp = np.ctypeslib.as_array(c.row_ptr, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(c.col, shape=(max(nnz, 1),))[:nnz].copy()
    w = None
    if c.w:
        w = np.ctypeslib.as_array(c.w, shape=(max(nnz, 1),))[:nnz].copy()
    out = Lut(rp, col, w, c.hist_min, c.hist_step)
    lib().or_lut_free(C.byref(c))
    return out


def tth_chi(p: Poni, d1, d2):
    d1 = np.ascontiguousarray(d1, np.float64)
    d2 = np.ascontiguousarray(d2, np.float64)
    t = np.empty_like(d1)
    a = np.empty_like(d1)
    lib().or_tth_chi(C.byref(p.c()), _d(d1), _d(d2), C.c_size_t(d1.size), _d(t), _d(a))
    return t, a


def positions(p: Poni, d1, d2):
    d1 = np.ascontiguousarray(d1, np.float64)
    d2 = np.ascontiguousarray(d2, np.float64)
    x, y, z = (np.empty_like(d1) for _ in range(3))
    lib().or_positions(C.byref(p.c()), _d(d1), _d(d2), C.c_size_t(d1.size), _d(x), _d(y), _d(z))
    return x, y, z


def corners(p: Poni, dim1, dim2):
    n = (dim1 + 1) * (dim2 + 1)
    t, a = np.empty(n), np.empty(n)
    lib().or_corners(C.byref(p.c()), C.c_size_t(dim1), C.c_size_t(dim2), _d(t), _d(a))
    return t, a


def centres(p: Poni, dim1, dim2):
    n = dim1 * dim2
    t, a = np.empty(n), np.empty(n)
    lib().or_centres(C.byref(p.c()), C.c_size_t(dim1), C.c_size_t(dim2), _d(t), _d(a))
    return t, a


def solid_angle(p: Poni, dim1, dim2):
    out = np.empty(dim1 * dim2)
    lib().or_solid_angle(C.byref(p.c()), C.c_size_t(dim1), C.c_size_t(dim2), _d(out))
    return out


def polarization(factor, tth, chi):
    tth = np.ascontiguousarray(tth, np.float64)
    chi = np.ascontiguousarray(chi, np.float64)
    out = np.empty_like(tth)
    lib().or_polarization(C.c_double(factor), _d(tth), _d(chi), C.c_size_t(tth.size), _d(out))
    return out


def normalization(p: Poni, dim1, dim2, polarization_factor=None):
    out = np.empty(dim1 * dim2)
    lib().or_normalization(C.byref(p.c()), C.c_size_t(dim1), C.c_size_t(dim2),
                           C.c_int(polarization_factor is not None),
                           C.c_double(polarization_factor or 0.0), _d(out))
    return out


def build_lut1d(p: Poni, dim1, dim2, n_points, method="pixel_split", unit="2th_rad",
                radial_range=None) -> Lut:
    c = _Lut()
    has = radial_range is not None
    lo, hi = radial_range if has else (0.0, 0.0)
    fn = {"pixel_split": lib().or_build_lut1d_ps,
          "counting": lib().or_build_lut1d_counting}[method]
    rc = fn(C.byref(p.c()), C.c_size_t(dim1), C.c_size_t(dim2), C.c_size_t(n_points),
            C.c_int(UNITS[unit]), C.c_int(has), C.c_double(lo), C.c_double(hi), C.byref(c))
    if rc != 0:
        raise ValueError("invalid table request")
    return _take(c)


def fixed_table(lut: Lut, mask) -> Lut:
    c = _Lut()
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    lib().or_fixed_table(C.byref(lut._c()), None if m is None else m.ctypes.data_as(C.c_void_p),
                         C.byref(c))
    return _take(c)


def denominators(lut: Lut, norm) -> np.ndarray:
    norm = np.ascontiguousarray(norm, np.float64)
    den = np.empty(lut.nbins)
    lib().or_denominators(C.byref(lut._c()), _d(norm), _d(den))
    return den


def gather(lut: Lut, den, signal, with_var=True):
    den = np.ascontiguousarray(den, np.float64)
    signal = np.ascontiguousarray(signal, np.float64).ravel()
    inten, sig = np.empty(lut.nbins), np.empty(lut.nbins)
    lib().or_gather(C.byref(lut._c()), _d(den), _d(signal), C.c_int(int(with_var)), _d(inten),
                    _d(sig))
    return inten, (sig if with_var else None)


def azimuthal_std(lut: Lut, mask, norm, signal, errors=False):
    norm = np.ascontiguousarray(norm, np.float64)
    signal = np.ascontiguousarray(signal, np.float64).ravel()
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    inten, sig, std = np.empty(lut.nbins), np.empty(lut.nbins), np.empty(lut.nbins)
    lib().or_azimuthal_std(C.byref(lut._c()), None if m is None else m.ctypes.data_as(C.c_void_p),
                           _d(norm), _d(signal), C.c_int(int(errors)), _d(inten),
                           _d(sig) if errors else None, _d(std))
    return inten, (sig if errors else None), std


class Integrator:
    """Oracle counterpart of the reference Integrator for fixed masks:
    build table, compact by mask, strict denominators, strict gather."""

    def __init__(self, poni: Poni, method="pixel_split", polarization_factor=None, unit="2th_rad",
                 radial_range=None, mask=None):
        self.poni, self.method, self.pol = poni, method, polarization_factor
        self.unit, self.radial_range, self.mask = unit, radial_range, mask
        self._cache = {}

    def _prepared(self, shape, n_points):
        key = (shape, n_points)
        if key not in self._cache:
            dim1, dim2 = shape
            lut = build_lut1d(self.poni, dim1, dim2, n_points, self.method, self.unit,
                              self.radial_range)
            norm = normalization(self.poni, dim1, dim2, self.pol)
            fixed = fixed_table(lut, self.mask) if self.mask is not None else lut
            den = denominators(fixed, norm)
            self._cache[key] = (lut, fixed, norm, den)
        return self._cache[key]

    def integrate1d(self, frame, n_points, errors=False):
        lut, fixed, _, den = self._prepared(frame.shape, n_points)
        i, s = gather(fixed, den, np.asarray(frame, np.float64), errors)
        return lut.bins(), i, s

    def integrate1d_with_azimuthal_std(self, frame, n_points, errors=False):
        lut, _, norm, _ = self._prepared(frame.shape, n_points)
        i, s, std = azimuthal_std(lut, self.mask, norm, np.asarray(frame, np.float64), errors)
        return lut.bins(), i, s, std


class CscBaseline:
    """Timed CPU baseline: restated default reference batch1d (CSC, 4-frame tiles)."""

    def __init__(self, lut: Lut, npix: int, mask, norm):
        self._lutc = lut._c()
        self._m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        self._norm = np.ascontiguousarray(norm, np.float64)
        self.nbins = lut.nbins
        self.h = lib().or_csc_new(C.byref(self._lutc), C.c_size_t(npix),
                                  None if self._m is None else self._m.ctypes.data_as(C.c_void_p),
                                  _d(self._norm))

    def batch_u32(self, frames: np.ndarray, threads: int) -> np.ndarray:
        frames = np.ascontiguousarray(frames, np.uint32)
        out = np.empty((frames.shape[0], self.nbins))
        lib().or_csc_batch_u32(C.c_void_p(self.h), frames.ctypes.data_as(C.c_void_p),
                               C.c_size_t(frames.shape[0]), C.c_int(threads), _d(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_csc_free(C.c_void_p(self.h))
            self.h = None


def counting_from_radial(radial, n_points, radial_range=None) -> Lut:
    r = np.ascontiguousarray(radial, np.float64)
    c = _Lut()
    has = radial_range is not None
    lo, hi = radial_range if has else (0.0, 0.0)
    rc = lib().or_counting_from_radial(_d(r), C.c_size_t(r.size), C.c_size_t(n_points),
                                       C.c_int(has), C.c_double(lo), C.c_double(hi), C.byref(c))
    if rc != 0:
        raise ValueError("invalid table request")
    return _take(c)


def ps_fractions(p: Poni, dim1, dim2, n_points, min_, step, clamp_low=False,
                 unit="2th_rad") -> Lut:
    c = _Lut()
    rc = lib().or_ps_fractions(C.byref(p.c()), C.c_size_t(dim1), C.c_size_t(dim2),
                               C.c_size_t(n_points), C.c_int(UNITS[unit]), C.c_double(min_),
                               C.c_double(step), C.c_int(int(clamp_low)), C.byref(c))
    if rc != 0:
        raise ValueError("invalid table request")
    return _take(c)


def build_lut2d(p: Poni, dim1, dim2, n_radial, n_azimuthal, method="pixel_split",
                unit="2th_rad", radial_range=None, azimuthal_range=None):
    """Default 2D table (counting or pixel_split); returns (Lut, (r0, r1, a0, a1))."""
    c = _Lut()
    edges = np.empty(4)
    rr = radial_range or (0.0, 0.0)
    ar = azimuthal_range or (0.0, 0.0)
    rc = lib().or_build_lut2d(C.byref(p.c()), C.c_size_t(dim1), C.c_size_t(dim2),
                              C.c_int({"counting": 0, "pixel_split": 1}[method]),
                              C.c_int(UNITS[unit]), C.c_size_t(n_radial), C.c_size_t(n_azimuthal),
                              C.c_int(radial_range is not None), C.c_double(rr[0]),
                              C.c_double(rr[1]), C.c_int(azimuthal_range is not None),
                              C.c_double(ar[0]), C.c_double(ar[1]), C.byref(c), _d(edges))
    if rc != 0:
        raise ValueError("invalid 2D table request")
    return _take(c), tuple(edges)


def build_lut1d_ss(p: Poni, dim1, dim2, n_points, n_ss, unit="2th_rad", radial_range=None) -> Lut:
    c = _Lut()
    has = radial_range is not None
    lo, hi = radial_range if has else (0.0, 0.0)
    rc = lib().or_build_lut1d_ss(C.byref(p.c()), C.c_size_t(dim1), C.c_size_t(dim2),
                                 C.c_size_t(n_points), C.c_size_t(n_ss), C.c_int(UNITS[unit]),
                                 C.c_int(has), C.c_double(lo), C.c_double(hi), C.byref(c))
    if rc != 0:
        raise ValueError("invalid table request")
    return _take(c)


def build_lut2d_ss(p: Poni, dim1, dim2, n_radial, n_azimuthal, n_ss, unit="2th_rad",
                   radial_range=None, azimuthal_range=None):
    c = _Lut()
    edges = np.empty(4)
    rr = radial_range or (0.0, 0.0)
    ar = azimuthal_range or (0.0, 0.0)
    rc = lib().or_build_lut2d_ss(C.byref(p.c()), C.c_size_t(dim1), C.c_size_t(dim2),
                                 C.c_size_t(n_radial), C.c_size_t(n_azimuthal), C.c_size_t(n_ss),
                                 C.c_int(UNITS[unit]), C.c_int(radial_range is not None),
                                 C.c_double(rr[0]), C.c_double(rr[1]),
                                 C.c_int(azimuthal_range is not None), C.c_double(ar[0]),
                                 C.c_double(ar[1]), C.byref(c), _d(edges))
    if rc != 0:
        raise ValueError("invalid 2D table request")
    return _take(c), tuple(edges)


def build_lut1d_pyfai(p: Poni, dim1, dim2, n_points, unit="2th_rad", radial_range=None) -> Lut:
    c = _Lut()
    has = radial_range is not None
    lo, hi = radial_range if has else (0.0, 0.0)
    rc = lib().or_build_lut1d_pyfai(C.byref(p.c()), C.c_size_t(dim1), C.c_size_t(dim2),
                                    C.c_size_t(n_points), C.c_int(UNITS[unit]), C.c_int(has),
                                    C.c_double(lo), C.c_double(hi), C.byref(c))
    if rc not in (0, 3):
        raise ValueError("invalid table request")
    lut = _take(c)
    if rc == 3:
        raise ValueError("pyFAI-compatible splitting exceeds the 64-million-entry bound")
    return lut


def build_lut2d_pyfai(p: Poni, dim1, dim2, n_radial, n_azimuthal, unit="2th_rad",
                      radial_range=None, azimuthal_range=None):
    c = _Lut()
    edges = np.empty(4)
    rr = radial_range or (0.0, 0.0)
    ar = azimuthal_range or (0.0, 0.0)
    rc = lib().or_build_lut2d_pyfai(C.byref(p.c()), C.c_size_t(dim1), C.c_size_t(dim2),
                                    C.c_int(UNITS[unit]), C.c_size_t(n_radial),
                                    C.c_size_t(n_azimuthal), C.c_int(radial_range is not None),
                                    C.c_double(rr[0]), C.c_double(rr[1]),
                                    C.c_int(azimuthal_range is not None), C.c_double(ar[0]),
                                    C.c_double(ar[1]), C.byref(c), _d(edges))
    if rc != 0:
        raise ValueError("invalid 2D table request")
    return _take(c), tuple(edges)


def pyfai_split1d(poly, r0, r1, nbins):
    p = np.ascontiguousarray(poly, np.float64).ravel()
    b = np.empty(64, np.uint64)
    w = np.empty(64)
    n = lib().or_pyfai_split1d(_d(p), C.c_double(r0), C.c_double(r1), C.c_size_t(nbins),
                               b.ctypes.data_as(C.c_void_p), _d(w), C.c_size_t(64))
    return list(zip(b[:n].tolist(), w[:n].tolist()))


def pyfai_split1d_sector(poly, r0, r1, nbins, a0, a1, sector=True, sign=1.0):
    """split1d (pyfai_1d.rs:259-354) on one polygon with an azimuthal range
    [a0, a1], sector exclusion and corner sign."""
    p = np.ascontiguousarray(poly, np.float64).ravel()
    b = np.empty(64, np.uint64)
    w = np.empty(64)
    L = lib()
    L.or_pyfai_split1d_sector.restype = C.c_int
    n = L.or_pyfai_split1d_sector(_d(p), C.c_double(r0), C.c_double(r1), C.c_size_t(nbins),
                                  C.c_double(a0), C.c_double(a1), C.c_int(int(sector)),
                                  C.c_double(sign), b.ctypes.data_as(C.c_void_p), _d(w),
                                  C.c_size_t(64))
    return list(zip(b[:n].tolist(), w[:n].tolist()))


def wrap_tau(x: float) -> float:
    """pyfai_1d.rs:180-186: x.rem_euclid(TAU), values in [0, TAU) unchanged."""
    L = lib()
    L.or_wrap_tau.restype = C.c_double
    L.or_wrap_tau.argtypes = [C.c_double]
    return L.or_wrap_tau(x)


GRID_RADIAL_END, GRID_INDEX_ROUNDING, GRID_RADIAL_OUTSIDE, GRID_AZIMUTHAL_OUTSIDE = 1, 2, 4, 8


def grid2d_bin(nr, na, radial_range, azimuthal_range, flags, radial, az):
    """IntegrationGrid2d::bin (grid.rs:176-221): the bin index, None for no
    bin; raises ValueError for an invalid grid."""
    L = lib()
    L.or_grid2d_bin.restype = C.c_longlong
    b = L.or_grid2d_bin(C.c_size_t(nr), C.c_size_t(na), C.c_double(radial_range[0]),
                        C.c_double(radial_range[1]), C.c_double(azimuthal_range[0]),
                        C.c_double(azimuthal_range[1]), C.c_int(flags), C.c_double(radial),
                        C.c_double(az))
    if b == -2:
        raise ValueError("invalid grid")
    return None if b < 0 else int(b)


# ------------------------------------------------------------ Bitshuffle + LZ4

def lz4_decode_block(block: bytes, out_cap: int):
    """Decoded bytes of one LZ4 block, or None when it is malformed."""
    out = np.empty(max(out_cap, 1), np.uint8)
    L = lib()
    L.or_lz4_decode_block.restype = C.c_long
    n = L.or_lz4_decode_block(C.c_char_p(bytes(block)), C.c_size_t(len(block)),
                              out.ctypes.data_as(C.c_void_p), C.c_size_t(out_cap))
    return None if n < 0 else out[:n].tobytes()


def bitshuffle(arr: np.ndarray) -> bytes:
    a = np.ascontiguousarray(arr)
    out = np.empty(a.nbytes, np.uint8)
    lib().or_bitshuffle(a.ctypes.data_as(C.c_void_p), C.c_size_t(a.size),
                        C.c_size_t(a.itemsize), out.ctypes.data_as(C.c_void_p))
    return out.tobytes()


def bitunshuffle(data: bytes, dtype, n: int) -> np.ndarray:
    out = np.empty(n, dtype)
    lib().or_bitunshuffle(C.c_char_p(bytes(data)), C.c_size_t(n), C.c_size_t(out.itemsize),
                          out.ctypes.data_as(C.c_void_p))
    return out


def bslz4_decode(chunk: bytes, pixels: int, dtype):
    """(status, elements): status 0 ok, 1 header, 2 truncated, 3 bad block."""
    out = np.zeros(pixels, dtype)
    rc = lib().or_bslz4_decode(C.c_char_p(bytes(chunk)), C.c_size_t(len(chunk)),
                               C.c_size_t(pixels), C.c_size_t(out.itemsize),
                               out.ctypes.data_as(C.c_void_p))
    return rc, out


def bslz4_encode(frame: np.ndarray, block_bytes: int = 8192, mode: int = 0) -> bytes:
    """The reference's test encoders: mode 0 literal-only LZ4, 1 byte runs."""
    a = np.ascontiguousarray(frame).ravel()
    cap = 64 + (a.nbytes // block_bytes + 3) * (block_bytes + block_bytes // 255 + 24) + a.nbytes
    out = np.empty(cap, np.uint8)
    L = lib()
    L.or_bslz4_encode.restype = C.c_size_t
    n = L.or_bslz4_encode(a.ctypes.data_as(C.c_void_p), C.c_size_t(a.size),
                          C.c_size_t(a.itemsize), C.c_size_t(block_bytes), C.c_int(mode),
                          out.ctypes.data_as(C.c_void_p), C.c_size_t(cap))
    assert n > 0
    return out[:n].tobytes()
