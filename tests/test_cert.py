"""The device builder's certified atan2 and sine (dioptrin_amd/csrc/
atan2_cert.hpp, sin_cert.hpp, compiled for the host by oracle/Makefile)
against glibc, the libm the reference's f64::atan2 and f64::sin call
(geometry.rs:684-692, integration.rs:171-180): every value a certificate
accepts equals libm bit for bit, on random arguments and on those of
detector geometry (2theta = atan2(sqrt(x^2 + y^2), z), chi = atan2(x, y),
q = q_scale * sin(2theta / 2)); rejected values are the minority the host
recomputes; exact special cases match IEEE."""
import ctypes as C
import math
import os

import numpy as np
import pytest

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                   "libcert.so")


def run(y, x):
    lib = C.CDLL(LIB)
    y = np.ascontiguousarray(y, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    n = y.size
    cert, libm = np.empty(n), np.empty(n)
    flag = np.empty(n, np.uint8)
    dp = C.POINTER(C.c_double)
    lib.or_atan2_cert(y.ctypes.data_as(dp), x.ctypes.data_as(dp), C.c_size_t(n),
                      cert.ctypes.data_as(dp), libm.ctypes.data_as(dp),
                      flag.ctypes.data_as(C.c_void_p))
    return cert, libm, flag.astype(bool)


def check(y, x, max_flagged):
    cert, libm, flag = run(y, x)
    ok = ~flag
    bad = np.flatnonzero(cert[ok].view(np.uint64) != libm[ok].view(np.uint64))
    assert bad.size == 0, (y[ok][bad[:5]], x[ok][bad[:5]])
    assert flag.mean() <= max_flagged, flag.mean()
    return flag.mean()


def test_random_quadrants_and_magnitudes():
    rng = np.random.default_rng(20261204)
    n = 2_000_000
    mag = 10.0 ** rng.uniform(-6, 6, size=(2, n))
    y = mag[0] * rng.choice([-1.0, 1.0], n)
    x = mag[1] * rng.choice([-1.0, 1.0], n)
    check(y, x, 0.2)


def test_detector_geometry_arguments():
    """Corner coordinates of a 3269 x 3110 detector of 75 um pixels at
    0.25 m and a tilted 1043 x 981 one of 172 um pixels at 0.18 m."""
    rng = np.random.default_rng(7)
    for pixel, d1, d2, dist in ((75e-6, 3269, 3110, 0.25), (172e-6, 1043, 981, 0.18)):
        i = rng.integers(0, d1 + 1, 1_000_000).astype(np.float64)
        j = rng.integers(0, d2 + 1, 1_000_000).astype(np.float64)
        xs = i * pixel - d1 * pixel / 2
        ys = j * pixel - d2 * pixel / 2
        z = np.full_like(xs, dist) + rng.normal(0.0, 1e-3, xs.size)
        check(np.sqrt(ys * ys + xs * xs), z, 0.2)     # 2theta
        check(xs, ys, 0.2)                            # chi


def test_special_cases():
    pi = math.pi
    cases = [(0.0, 1.0, 0.0), (-0.0, 1.0, -0.0), (0.0, -1.0, pi), (-0.0, -1.0, -pi),
             (0.0, 0.0, 0.0), (0.0, -0.0, pi), (1.0, 0.0, pi / 2), (-1.0, 0.0, -pi / 2),
             (2.0, -0.0, pi / 2)]
    y = np.array([c[0] for c in cases])
    x = np.array([c[1] for c in cases])
    cert, libm, flag = run(y, x)
    assert not flag.any()
    for k, (_, _, want) in enumerate(cases):
        assert cert[k] == want and math.copysign(1.0, cert[k]) == math.copysign(1.0, want)
        assert cert[k] == libm[k]
    _, _, flag = run(np.array([np.inf, np.nan, 1.0]), np.array([1.0, 1.0, np.inf]))
    assert flag.all()   # non-finite arguments go to the host libm


def run_sin(x):
    lib = C.CDLL(LIB)
    x = np.ascontiguousarray(x, np.float64)
    n = x.size
    cert, libm = np.empty(n), np.empty(n)
    flag = np.empty(n, np.uint8)
    dp = C.POINTER(C.c_double)
    lib.or_sin_cert(x.ctypes.data_as(dp), C.c_size_t(n), cert.ctypes.data_as(dp),
                    libm.ctypes.data_as(dp), flag.ctypes.data_as(C.c_void_p))
    return cert, libm, flag.astype(bool)


def check_sin(x, max_flagged):
    cert, libm, flag = run_sin(x)
    ok = ~flag
    bad = np.flatnonzero(cert[ok].view(np.uint64) != libm[ok].view(np.uint64))
    assert bad.size == 0, x[ok][bad[:5]]
    assert flag.mean() <= max_flagged, flag.mean()


def test_sin_random_and_half_angles():
    """Half angles 2theta / 2 over [0, pi/2] uniform, small angles down to
    2^-20 (pixels next to the beam), and the half angles of the corner
    2theta of a 3269 x 3110 detector at 0.25 m."""
    rng = np.random.default_rng(20261205)
    check_sin(rng.uniform(0.0, math.pi / 2, 3_000_000), 0.2)
    check_sin(np.ldexp(rng.uniform(0.5, 1.0, 1_000_000), -rng.integers(1, 20, 1_000_000)), 0.2)
    i = rng.integers(0, 3270, 1_000_000) * 75e-6 - 3269 * 75e-6 / 2
    j = rng.integers(0, 3111, 1_000_000) * 75e-6 - 3110 * 75e-6 / 2
    check_sin(0.5 * np.arctan2(np.sqrt(i * i + j * j), 0.25), 0.2)


def test_sin_special_cases():
    cert, libm, flag = run_sin(np.array([0.0, 1e-300, 2.0 ** -21, -0.5, 3.0, np.nan, np.inf]))
    assert not flag[0] and cert[0] == 0.0 and math.copysign(1.0, cert[0]) == 1.0
    assert flag[1:].all()   # outside (2^-20, 2]: the host decides
