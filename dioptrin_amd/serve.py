"""Continuous-batching integration server (Python side of dio_server_*).

A Server owns a pool of stream slots, an LRU of device-resident plans and an
HBM-footprint admission budget, all inside libdioptrin_hip.so. Requests of any
detector shape, bin count and 1D/2D mode may be submitted while others run;
each is identified by the same arguments as a GpuPlan (geometry, shape,
method, bins, ranges, polarization, mask, precision), which is how the plan
cache is keyed (the reference keys its one-entry caches on the same
settings: src/integration/cache.rs:23-36).

`run_load` drives a server with Poisson arrivals and reports throughput and
p50/p99 submit-to-result latency.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass
from typing import Dict, List, Optional

import numpy as np

from . import _native as N
from .device import DeviceArray
from .integrator import _NP_DTYPES, _Context, plan_args


@dataclass
class Request:
    """One integration request: frames (n, dim1, dim2) or (dim1, dim2) as a
    host array, or a DeviceArray already in HBM, and the plan settings
    (GpuPlan arguments)."""
    poni: dict
    frames: np.ndarray
    n_radial: int = 1000
    n_azimuthal: Optional[int] = None
    method: str = "pixel_split"
    n_ss: Optional[int] = None
    unit: str = "2th_rad"
    radial_range: Optional[tuple] = None
    azimuthal_range: Optional[tuple] = None
    polarization_factor: Optional[float] = None
    mask: Optional[np.ndarray] = None
    precision: str = "f32_compensated"
    errors: bool = False
    azimuthal_std: bool = False


@dataclass
class Response:
    intensity: np.ndarray
    errors: Optional[np.ndarray]
    azimuthal_std: Optional[np.ndarray]
    latency_ms: float


class _Pending:
    __slots__ = ("req", "frames", "mask", "out", "desc", "poni", "shape")


class Server:
    """dio_server handle. n_slots requests run at once, each on its own HIP
    stream; requests whose frames + outputs do not fit hbm_budget_bytes wait
    for running ones to finish."""

    def __init__(self, n_slots: int = 4, plan_cache_entries: int = 32,
                 hbm_budget_bytes: int = 32 << 30, device: int = 0):
        cfg = N.ServerConfig(n_slots, plan_cache_entries, hbm_budget_bytes)
        h = C.c_void_p()
        N.check(N.lib().dio_server_create(_Context.get(device), C.byref(cfg), C.byref(h)))
        self._h = h
        self._pending: Dict[int, _Pending] = {}

    def submit(self, req: Request) -> int:
        on_device = isinstance(req.frames, DeviceArray)
        if on_device:
            fr = req.frames
            if len(fr.shape) != 3 or fr.shape[0] == 0:
                raise ValueError(f"device frames must be (n, dim1, dim2), got {fr.shape}")
            fptr = fr.ptr
        else:
            fr = np.asarray(req.frames)
            if fr.ndim == 2:
                fr = fr[None]
            if fr.ndim != 3 or fr.shape[0] == 0:
                raise ValueError(f"frames must be (n, dim1, dim2), got {fr.shape}")
            fr = np.ascontiguousarray(fr)
            fptr = fr.ctypes.data
        if fr.dtype not in _NP_DTYPES:
            raise ValueError(f"unsupported frame dtype {fr.dtype}")
        poni, desc, mask = plan_args(
            req.poni, fr.shape[1:], req.method, req.n_ss, req.unit, req.n_radial,
            req.n_azimuthal, req.radial_range, req.azimuthal_range, req.polarization_factor,
            req.mask, req.precision)
        nbins = req.n_radial * (req.n_azimuthal or 1)
        n = fr.shape[0]
        out_i = np.empty((n, nbins))
        out_s = np.empty((n, nbins)) if req.errors else None
        out_d = np.empty((n, nbins)) if req.azimuthal_std else None
        flags = (N.WITH_ERRORS if req.errors else 0) | \
            (N.WITH_AZIMUTHAL_STD if req.azimuthal_std else 0) | \
            (N.DEVICE_FRAMES if on_device else 0)
        t = C.c_uint64()
        dt = N.DTYPES[_NP_DTYPES[fr.dtype]]
        N.check(N.lib().dio_server_submit(
            self._h, C.byref(poni), C.byref(desc), fptr, dt, n, flags, out_i.ctypes.data, None if out_s is None else out_s.ctypes.data,
            None if out_d is None else out_d.ctypes.data, C.byref(t)))
        p = _Pending()
        p.req, p.frames, p.mask, p.out = req, fr, mask, (out_i, out_s, out_d)
        self._pending[t.value] = p
        return t.value

    def poll(self, ticket: int, timeout: float = 0.0) -> Optional[Response]:
        """The response once ready (None before); raises the request's error."""
        if ticket not in self._pending:
            raise KeyError(f"unknown ticket {ticket}")
        done = C.c_int()
        lat = C.c_double()
        status = N.lib().dio_server_poll(self._h, ticket, float(timeout), C.byref(done),
                                         C.byref(lat))
        if status == N.DIO_OK and not done.value:
            return None
        p = self._pending.pop(ticket)
        N.check(status)
        i, s, d = p.out
        return Response(i, s, d, lat.value)

    def result(self, ticket: int, timeout: float = 600.0) -> Response:
        r = self.poll(ticket, timeout)
        if r is None:
            raise TimeoutError(f"request {ticket} not done after {timeout} s")
        return r

    def stats(self) -> dict:
        st = N.ServerStats()
        N.check(N.lib().dio_server_stats_get(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in N.ServerStats._fields_}

    def close(self) -> None:
        if self._h:
            N.check(N.lib().dio_server_destroy(self._h))
            self._h = None
            self._pending.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------- load generator

@dataclass
class LoadMix:
    """Request mix: geometries (name -> (poni dict, shape, host frame pool)),
    frames per request, share of 1D requests and their bin counts, 2D grid."""
    geometries: Dict[str, tuple]
    frames_min: int = 1
    frames_max: int = 64
    p1d: float = 0.6
    bins_1d: tuple = (1000, 3000)
    grid_2d: tuple = (500, 360)
    method: str = "pixel_split"
    precision: str = "f32_compensated"
    bins_choices: int = 5  # distinct 1D bin counts per geometry (keeps the plan set finite)


def make_requests(mix: LoadMix, n: int, seed: int) -> List[Request]:
    rng = np.random.default_rng(seed)
    names = sorted(mix.geometries)
    lo, hi = mix.bins_1d
    bins_set = np.linspace(lo, hi, mix.bins_choices).round().astype(int)
    reqs = []
    for _ in range(n):
        name = names[rng.integers(len(names))]
        poni, shape, pool = mix.geometries[name]
        nf = int(rng.integers(mix.frames_min, mix.frames_max + 1))
        nf = min(nf, pool.shape[0])
        start = int(rng.integers(0, pool.shape[0] - nf + 1))
        frames = pool[start:start + nf]
        if rng.random() < mix.p1d:
            reqs.append(Request(poni, frames, n_radial=int(rng.choice(bins_set)),
                                method=mix.method, precision=mix.precision))
        else:
            reqs.append(Request(poni, frames, n_radial=mix.grid_2d[0],
                                n_azimuthal=mix.grid_2d[1], method=mix.method,
                                precision=mix.precision))
    return reqs


def run_load(server: Server, reqs: List[Request], rate: float, seed: int = 0,
             keep: bool = False) -> dict:
    """Submits reqs with exponential inter-arrival times (rate per second),
    collects every response, and returns the server's statistics plus the
    offered rate. keep=True also returns the responses in request order."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=len(reqs))
    t0 = time.perf_counter()
    due = t0
    tickets = []
    out: List[Optional[Response]] = [None] * len(reqs)
    waiting = {}
    for k, r in enumerate(reqs):
        due += gaps[k]
        while True:
            # Collect finished requests while waiting for the next arrival.
            for tk in list(waiting):
                resp = server.poll(tk, 0.0)
                if resp is not None:
                    out[waiting.pop(tk)] = resp
            now = time.perf_counter()
            if now >= due:
                break
            time.sleep(min(due - now, 0.0005))
        tk = server.submit(r)
        tickets.append(tk)
        waiting[tk] = k
    for tk, k in list(waiting.items()):
        out[k] = server.result(tk)
    wall = time.perf_counter() - t0
    st = server.stats()
    # Latency percentiles of this load (the server's own statistics cover its
    # whole lifetime).
    lat = np.sort([r.latency_ms for r in out])
    pct = lambda q: float(lat[min(len(lat) - 1, int(np.ceil(q * len(lat))) - 1)])
    st.update(p50_ms=pct(0.50), p99_ms=pct(0.99), max_ms=float(lat[-1]),
              mean_ms=float(lat.mean()))
    frames = sum(1 if len(r.frames.shape) == 2 else int(r.frames.shape[0]) for r in reqs)
    res = dict(st, offered_rate=rate, requests=len(reqs), wall_s=wall, frames_total=frames,
               frames_per_s_wall=frames / wall)
    if keep:
        res["responses"] = out
    return res
