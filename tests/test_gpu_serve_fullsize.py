"""Continuous-batching server at the serving mix's real shapes.

Pilatus 1M (1043 x 981), Eiger 4M (2167 x 2070) and Eiger 9M (3269 x 3110)
detectors, 1D requests of 1000-3000 bins and 2D requests of 500 x 360,
device-resident uint32 frames, pixel splitting:
  - every result of the concurrent mix equals the same request computed
    alone, bit for bit;
  - the Pilatus 1D and 2D requests match the FP64 oracle (per bin
    |gpu - oracle| <= 2e-6 * magnitude, relative error <= 1e-6, as the
    gather tolerance of src/integration/cuda/tests.rs:206-250);
  - a plan built while the server runs (a cold Eiger 9M table) holds back
    only its own request.
"""
import time
from dataclasses import dataclass

import numpy as np
import pytest

import oracle as O
from dioptrin_amd import GpuPlan, synth
from dioptrin_amd.device import synchronize
from dioptrin_amd.serve import Request, Server

pytestmark = pytest.mark.gpu

NAMES = ("pilatus1m_ps1d", "eiger4m_ss2d", "eiger9m_batch1d")
POOL = 8  # device frames per detector

# (detector, frames, first frame, radial bins, azimuthal bins, errors,
#  azimuthal std, precision)
SPEC = [
    ("pilatus1m_ps1d", 3, 0, 1000, None, True, False, "f32_wide"),
    ("eiger4m_ss2d", 2, 1, 500, 360, True, False, "f32_wide"),
    ("eiger9m_batch1d", 4, 2, 2000, None, True, False, "f32_wide"),
    ("pilatus1m_ps1d", 5, 3, 500, 360, False, False, "f32_wide"),
    ("eiger4m_ss2d", 2, 4, 3000, None, True, True, "f32_wide"),
    ("eiger9m_batch1d", 1, 7, 3000, None, False, False, "f32_compensated"),
    ("pilatus1m_ps1d", 8, 0, 1500, None, True, False, "f32_compensated"),
    ("eiger9m_batch1d", 2, 0, 500, 360, True, False, "f32_wide"),
    ("eiger4m_ss2d", 6, 2, 1000, None, False, False, "f32_wide"),
    ("pilatus1m_ps1d", 2, 5, 1000, None, True, False, "f32_wide"),  # plan of the first
]


@dataclass
class Pool:
    poni: dict
    shape: tuple
    dev: object       # DeviceArray (POOL, dim1, dim2) uint32
    host: np.ndarray  # the same frames on the host


@pytest.fixture(scope="module")
def pools():
    out = {}
    for k, name in enumerate(NAMES):
        dev = synth.frames_device(name, POOL, seed=31 + k)
        synchronize()
        out[name] = Pool(synth.workload_poni(name), synth.WORKLOADS[name]["shape"], dev,
                         dev.numpy())
    return out


def request(pools, spec, device=True):
    name, nf, s, nr, na, err, std, prec = spec
    p = pools[name]
    fr = p.dev[s:s + nf] if device else p.host[s:s + nf]
    return Request(p.poni, fr, n_radial=nr, n_azimuthal=na, errors=err, azimuthal_std=std,
                   precision=prec)


def same(a, b):
    return a is None and b is None or np.array_equal(np.asarray(a).view(np.uint64),
                                                     np.asarray(b).view(np.uint64))


@pytest.fixture(scope="module")
def mix(pools):
    reqs = [request(pools, sp) for sp in SPEC]
    print(f"case serve mix: {len(reqs)} requests at full detector sizes", flush=True)
    with Server(n_slots=4, plan_cache_entries=16) as srv:
        tickets = [srv.submit(r) for r in reqs]  # all in flight at once
        resps = [srv.result(t) for t in tickets]
        st = srv.stats()
    return resps, st


def test_fullsize_mix_equals_alone(pools, mix):
    resps, st = mix
    assert st["completed"] == len(SPEC) and st["failed"] == 0
    assert st["peak_inflight"] > 1, "requests never overlapped"
    assert st["plan_builds"] == len(SPEC) - 1 and st["cache_hits"] == 1
    plans = {}
    for sp, resp in zip(SPEC, resps):
        name, nf, s, nr, na, err, std, prec = sp
        key = (name, nr, na, prec)
        if key not in plans:
            p = pools[name]
            plans[key] = GpuPlan(p.poni, p.shape, n_radial=nr, n_azimuthal=na, precision=prec)
        I, S, D = plans[key].integrate_stack(pools[name].host[s:s + nf], errors=err,
                                             azimuthal_std=std)
        assert same(resp.intensity, I), sp
        assert same(resp.errors, S), sp
        assert same(resp.azimuthal_std, D), sp


def check(I, S, i_ref, s_ref, mag):
    assert np.all(np.abs(I - i_ref) <= 2e-6 * mag)
    ok = i_ref != 0
    assert np.all(I[~ok] == 0)
    rel = np.max(np.abs(I[ok] - i_ref[ok]) / np.abs(i_ref[ok]))
    assert rel <= 1e-6, rel
    if S is not None:
        assert np.all(np.abs(S - s_ref) <= 2e-6 * s_ref)
    return rel


def test_fullsize_mix_matches_oracle(pools, mix):
    resps, _ = mix
    p = pools["pilatus1m_ps1d"]
    P = O.Poni(**p.poni)
    d1, d2 = p.shape
    norm = O.normalization(P, d1, d2)
    for idx in (0, 3):
        name, nf, s, nr, na, err, std, prec = SPEC[idx]
        print(f"case serve oracle: request {idx} {name} {nr}x{na}", flush=True)
        if na is None:
            lut = O.build_lut1d(P, d1, d2, nr)
        else:
            lut, _ = O.build_lut2d(P, d1, d2, nr, na)
        den = O.denominators(lut, norm)
        resp = resps[idx]
        for f in range(nf):
            x = p.host[s + f].astype(np.float64)
            i_ref, s_ref = O.gather(lut, den, x)
            rel = check(resp.intensity[f], None if resp.errors is None else resp.errors[f],
                        i_ref, s_ref, np.abs(i_ref))
            print(f"maxrel serve request {idx} frame {f} {rel:.3e}", flush=True)


def test_cold_plan_holds_back_only_its_request(pools):
    pil, e9 = pools["pilatus1m_ps1d"], pools["eiger9m_batch1d"]
    warm = Request(pil.poni, pil.dev[0:1], n_radial=1000, precision="f32_wide")
    with Server(n_slots=4) as srv:
        srv.result(srv.submit(warm))  # its plan is now resident
        cold = srv.submit(Request(e9.poni, e9.dev[0:1], n_radial=2500, precision="f32_wide"))
        t0 = time.perf_counter()
        lat = [srv.result(t).latency_ms for t in [srv.submit(warm) for _ in range(6)]]
        t_warm = (time.perf_counter() - t0) * 1e3
        early = srv.poll(cold, 0.0)
        cold_resp = early if early is not None else srv.result(cold)
        st = srv.stats()
    print(f"serve cold plan: 6 warm requests in {t_warm:.1f} ms (max latency "
          f"{max(lat):.2f} ms), cold request {cold_resp.latency_ms:.1f} ms", flush=True)
    assert early is None, "the cold request finished before the warm ones"
    assert max(lat) < cold_resp.latency_ms
    assert st["plan_builds"] == 2 and st["completed"] == 8
