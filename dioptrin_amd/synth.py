"""Seeded synthetic inputs for the benchmark workloads (no datasets needed).

Workloads (geometry in metres / radians, PONI at the detector centre):
  pilatus1m_ps1d   1043 x 981, 172 um, d 0.20 m, lambda 0.3344 A, polarization
                   0.99, 1250 bins pixel splitting; float32 frames: flat 100 +
                   Gaussian Debye-Scherrer rings (FWHM 0.05 deg 2theta) at the
                   fcc a = 5.41 A reflections, amplitudes U[200, 5000], Poisson
                   sampled; mask = 2 % random pixels + 7-pixel horizontal gaps
                   every 195 rows.
  eiger9m_batch1d  3269 x 3110, 75 um, d 0.25 m, lambda 0.4 A, 2000 bins pixel
                   splitting; uint32 frames, per-frame background U[10, 500]
                   plus the ring pattern, Poisson sampled with a per-frame seed.
  eiger4m_ss2d     2167 x 2070, 75 um, d 0.15 m, rot1 0.02, rot2 -0.01, lambda
                   0.5 A, supersampling n_ss = 3, 1000 x 360 bins; uint32 rings
                   x (1 + 0.3 cos 2 chi) with spotty rings.
  pilatus1m_tilted 1043 x 981, 172 um, d 0.18 m, rot (0.05, -0.03, 0.10), 3000
                   bins and 500 x 360 bins, float32 frames.
"""
from __future__ import annotations

import math

import numpy as np

FCC_A = 5.41e-10


def poni_centre(shape, pixel, distance, wavelength, rot=(0.0, 0.0, 0.0)):
    dim1, dim2 = shape
    return {"pixel1": pixel, "pixel2": pixel, "distance": distance,
            "poni1": dim1 * pixel / 2.0, "poni2": dim2 * pixel / 2.0,
            "rot1": rot[0], "rot2": rot[1], "rot3": rot[2], "wavelength": wavelength}


WORKLOADS = {
    "pilatus1m_ps1d": dict(shape=(1043, 981), pixel=172e-6, distance=0.20, wavelength=0.3344e-10,
                           rot=(0.0, 0.0, 0.0), method="pixel_split", n_radial=1250,
                           polarization=0.99, dtype="float32"),
    "eiger4m_ss2d": dict(shape=(2167, 2070), pixel=75e-6, distance=0.15, wavelength=0.5e-10,
                         rot=(0.02, -0.01, 0.0), method="supersampled", n_ss=3, n_radial=1000,
                         n_azimuthal=360, polarization=None, dtype="uint32"),
    "eiger9m_batch1d": dict(shape=(3269, 3110), pixel=75e-6, distance=0.25, wavelength=0.4e-10,
                            rot=(0.0, 0.0, 0.0), method="pixel_split", n_radial=2000,
                            polarization=None, dtype="uint32", frames=256),
    "pilatus1m_tilted": dict(shape=(1043, 981), pixel=172e-6, distance=0.18,
                             wavelength=0.3344e-10, rot=(0.05, -0.03, 0.10), method="pyfai_full",
                             n_radial=3000, n_radial_2d=500, n_azimuthal=360, polarization=None,
                             dtype="float32"),
}


def workload_poni(name: str) -> dict:
    w = WORKLOADS[name]
    return poni_centre(w["shape"], w["pixel"], w["distance"], w["wavelength"], w["rot"])


def fcc_two_theta(wavelength: float, max_tth: float):
    """2theta of the allowed fcc reflections (hkl all even or all odd)."""
    out = set()
    hmax = 40
    for h in range(hmax):
        for k in range(h + 1):
            for l in range(k + 1):
                if (h, k, l) == (0, 0, 0):
                    continue
                if len({h % 2, k % 2, l % 2}) != 1:
                    continue
                s = h * h + k * k + l * l
                d = FCC_A / math.sqrt(s)
                x = wavelength / (2 * d)
                if x < 1:
                    t = 2 * math.asin(x)
                    if t <= max_tth:
                        out.add(round(t, 15))
    return sorted(out)


def pixel_angles(shape, poni: dict):
    """Approximate pixel-centre 2theta and chi for pattern synthesis (float64)."""
    dim1, dim2 = shape
    i = (np.arange(dim1) + 0.5)[:, None] * poni["pixel1"] - poni["poni1"]
    j = (np.arange(dim2) + 0.5)[None, :] * poni["pixel2"] - poni["poni2"]
    s1, c1 = math.sin(poni["rot1"]), math.cos(poni["rot1"])
    s2, c2 = math.sin(poni["rot2"]), math.cos(poni["rot2"])
    s3, c3 = math.sin(poni["rot3"]), math.cos(poni["rot3"])
    L = poni["distance"]
    x = i * c2 * c3 + j * (c3 * s1 * s2 - c1 * s3) - L * (c1 * c3 * s2 + s1 * s3)
    y = i * c2 * s3 + j * (c1 * c3 + s1 * s2 * s3) - L * (-(c3 * s1) + c1 * s2 * s3)
    z = i * s2 - j * c2 * s1 + L * c1 * c2
    return np.arctan2(np.sqrt(x * x + y * y), z), np.arctan2(x, y)


def ring_rate(shape, poni: dict, seed: int, background: float = 100.0,
              azimuthal_modulation: float = 0.0, spotty: bool = False) -> np.ndarray:
    """Expected counts: background + Gaussian rings (FWHM 0.05 deg)."""
    rng = np.random.default_rng(seed)
    tth, chi = pixel_angles(shape, poni)
    peaks = fcc_two_theta(poni["wavelength"], float(tth.max()) + 0.01)
    amps = rng.uniform(200.0, 5000.0, size=len(peaks))
    sigma = math.radians(0.05) / (2.0 * math.sqrt(2.0 * math.log(2.0)))
    rings = np.zeros(shape)
    for t0, a in zip(peaks, amps):
        d = tth - t0
        near = np.abs(d) < 6 * sigma
        rings[near] += a * np.exp(-0.5 * (d[near] / sigma) ** 2)
    if azimuthal_modulation:
        rings *= 1.0 + azimuthal_modulation * np.cos(2.0 * chi)
    if spotty:
        ring_px = rings > 1.0
        pick = ring_px & (rng.random(shape) < 0.10)
        rings[pick] *= rng.uniform(1.0, 20.0, size=int(pick.sum()))
    return background + rings


def mask_pilatus(shape, seed: int) -> np.ndarray:
    """2 % random pixels plus 7-pixel horizontal gaps every 195 rows."""
    rng = np.random.default_rng(seed + 7)
    m = (rng.random(shape) < 0.02).astype(np.uint8)
    for r0 in range(195, shape[0], 195):
        m[r0:r0 + 7, :] = 1
    return m


def frame_pilatus(seed: int, shape=(1043, 981), poni=None) -> np.ndarray:
    poni = poni or workload_poni("pilatus1m_ps1d")
    rate = ring_rate(shape, poni, seed=1234)
    rng = np.random.default_rng(seed)
    return rng.poisson(rate).astype(np.float32)


def frame_eiger4m(seed: int) -> np.ndarray:
    """uint32 frame of eiger4m_ss2d: background 100 + rings modulated by
    1 + 0.3 cos(2 chi), 10 % of ring pixels x U[1, 20] (spotty), Poisson
    sampled with `seed` (the ring amplitudes and spots are fixed)."""
    name = "eiger4m_ss2d"
    w = WORKLOADS[name]
    rate = ring_rate(w["shape"], workload_poni(name), seed=1234, azimuthal_modulation=0.3,
                     spotty=True)
    return np.random.default_rng(seed).poisson(rate).astype(np.uint32)


def frame_tilted(seed: int) -> np.ndarray:
    """float32 frame of pilatus1m_tilted: background 100 + rings of the tilted
    geometry, Poisson sampled with `seed`."""
    name = "pilatus1m_tilted"
    w = WORKLOADS[name]
    rate = ring_rate(w["shape"], workload_poni(name), seed=1234)
    return np.random.default_rng(seed).poisson(rate).astype(np.float32)


def frames_device(name: str, n: int, seed: int, device: int = 0):
    """n uint32 frames generated on the device (Poisson of the ring pattern
    plus a per-frame background U[10, 500], per-frame seeded streams)."""
    from .device import synthetic_poisson_u32
    w = WORKLOADS[name]
    poni = workload_poni(name)
    rate = ring_rate(w["shape"], poni, seed=1234, background=0.0).astype(np.float32)
    bg = np.random.default_rng(seed).uniform(10.0, 500.0, size=n)
    return synthetic_poisson_u32(rate, bg, seed, device)
