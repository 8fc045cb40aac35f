"""Continuous-batching server on the GPU: every result computed in a mix of
concurrent requests (different detector shapes, bin counts, 1D/2D, errors
and azimuthal std) equals the same request computed alone, bit for bit, and
both match the FP64 oracle within the gather tolerance; admission keeps the
in-flight footprint within the HBM budget; the plan cache evicts only idle
plans; a bad request fails alone.
"""
import numpy as np
import pytest

import oracle as O
from dioptrin_amd import GpuPlan
from dioptrin_amd.serve import LoadMix, Request, Server, make_requests, run_load

pytestmark = pytest.mark.gpu

GEOS = {
    "a": (dict(pixel1=172e-6, pixel2=172e-6, distance=0.05, poni1=30 * 172e-6,
               poni2=41 * 172e-6, wavelength=1e-10), (61, 83)),
    "b": (dict(pixel1=75e-6, pixel2=75e-6, distance=0.03, poni1=52 * 75e-6, poni2=47 * 75e-6,
               rot1=0.02, rot2=-0.01, wavelength=0.5e-10), (104, 96)),
    "c": (dict(pixel1=1e-4, pixel2=1.3e-4, distance=0.02, poni1=3e-4, poni2=4e-4, rot1=0.137,
               rot2=-0.073, rot3=0.419, wavelength=1e-10), (40, 57)),
}


def pools(seed=0, n=12):
    rng = np.random.default_rng(seed)
    out = {}
    for name, (poni, shape) in GEOS.items():
        out[name] = (poni, shape, rng.poisson(200.0, size=(n, *shape)).astype(np.uint32))
    return out


def mixed_requests(seed=1, n=24):
    rng = np.random.default_rng(seed)
    P = pools()
    reqs = []
    for k in range(n):
        name = "abc"[k % 3]
        poni, shape, pool = P[name]
        nf = int(rng.integers(1, 9))
        s = int(rng.integers(0, pool.shape[0] - nf + 1))
        fr = pool[s:s + nf]
        kind = k % 4
        if kind == 0:
            reqs.append(Request(poni, fr, n_radial=int(rng.choice([40, 75, 120])), errors=True))
        elif kind == 1:
            reqs.append(Request(poni, fr, n_radial=50, n_azimuthal=36, errors=bool(k % 2)))
        elif kind == 2:
            reqs.append(Request(poni, fr, n_radial=64, azimuthal_std=True, errors=True,
                                polarization_factor=0.99))
        else:
            reqs.append(Request(poni, fr, n_radial=90, precision="f32_wide"))
    return reqs


def alone(req):
    plan = GpuPlan(req.poni, req.frames.shape[1:], method=req.method, n_radial=req.n_radial,
                   n_azimuthal=req.n_azimuthal, polarization_factor=req.polarization_factor,
                   mask=req.mask, precision=req.precision)
    return plan.integrate_stack(req.frames, errors=req.errors, azimuthal_std=req.azimuthal_std)


def same(a, b):
    return a is None and b is None or np.array_equal(np.asarray(a).view(np.uint64),
                                                     np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("slots", [1, 4])
def test_mix_equals_alone(slots):
    reqs = mixed_requests()
    with Server(n_slots=slots, plan_cache_entries=8) as srv:
        tickets = [srv.submit(r) for r in reqs]  # all in flight at once
        resps = [srv.result(t) for t in tickets]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["peak_inflight"] <= slots
    if slots > 1:
        assert st["peak_inflight"] > 1, "requests never overlapped"
    for r, resp in zip(reqs, resps):
        I, S, D = alone(r)
        assert same(resp.intensity, I)
        assert same(resp.errors, S)
        assert same(resp.azimuthal_std, D)


def test_mix_matches_oracle():
    reqs = mixed_requests(seed=5, n=8)
    with Server(n_slots=3) as srv:
        resps = [srv.result(t) for t in [srv.submit(r) for r in reqs]]
    for r, resp in zip(reqs, resps):
        if r.n_azimuthal is not None:
            continue
        poni = O.Poni(**r.poni)
        d1, d2 = r.frames.shape[1:]
        lut = O.build_lut1d(poni, d1, d2, r.n_radial)
        norm = O.normalization(poni, d1, d2, r.polarization_factor)
        den = O.denominators(lut, norm)
        for f in range(r.frames.shape[0]):
            i, _ = O.gather(lut, den, r.frames[f].astype(np.float64))
            mag, _ = O.gather(lut, den, np.abs(r.frames[f].astype(np.float64)))
            assert np.all(np.abs(resp.intensity[f] - i) <= 2e-6 * mag + 1e-300)


def footprint(r):
    return r.frames.nbytes + r.frames.shape[0] * r.n_radial * (r.n_azimuthal or 1) * 8 * \
        (1 + r.errors + r.azimuthal_std)


def test_hbm_admission_counts_plans_and_scratch():
    reqs = mixed_requests(seed=2, n=12)
    # Resident plans and slot scratch of this mix, on a server with room for
    # everything.
    with Server(n_slots=4) as srv:
        for t in [srv.submit(r) for r in reqs]:
            srv.result(t)
        st0 = srv.stats()
    assert st0["request_bytes"] == 0 and st0["builds_pending"] == 0
    assert st0["bytes_in_use"] == st0["plan_bytes"] + st0["scratch_bytes"]
    assert st0["plans_resident"] == st0["plan_builds"] > 0 and st0["scratch_bytes"] > 0
    # Room for every plan, for the slots' scratch (a slot's scratch stays
    # below twice the largest need, and st0 holds at least that need once)
    # and for 1.2 of the largest request: admission must serialise most
    # requests.
    budget = st0["plan_bytes"] + 8 * st0["scratch_bytes"] + int(max(map(footprint, reqs)) * 1.2)
    with Server(n_slots=4, hbm_budget_bytes=budget) as srv:
        # Plans first (one frame each, one at a time), so the load below
        # runs against resident plans.
        for r in reqs:
            srv.result(srv.submit(Request(**{**r.__dict__, "frames": r.frames[:1]})))
        assert srv.stats()["plan_bytes"] == st0["plan_bytes"]
        resps = [srv.result(t) for t in [srv.submit(r) for r in reqs]]
        st = srv.stats()
    assert st["completed"] == 2 * len(reqs) and st["failed"] == 0
    assert st["peak_bytes"] <= budget
    assert st["request_bytes"] == 0 and st["inflight"] == 0
    assert st["bytes_in_use"] == st["plan_bytes"] + st["scratch_bytes"]
    for r, resp in zip(reqs[:4], resps[:4]):
        assert same(resp.intensity, alone(r)[0])


def test_idle_plans_evicted_for_room():
    # A detector large enough that its plans outweigh a request.
    shape = (512, 512)
    poni = dict(pixel1=172e-6, pixel2=172e-6, distance=0.1, poni1=250 * 172e-6,
                poni2=260 * 172e-6, wavelength=1e-10)
    fr = np.random.default_rng(3).poisson(50.0, size=(1, *shape)).astype(np.uint32)
    bins = [30, 31, 32, 30]
    # The slot scratch this sequence leaves, on a server with room for all.
    with Server(n_slots=2) as srv:
        for b in bins:
            srv.result(srv.submit(Request(poni, fr, n_radial=b)))
        scratch = srv.stats()["scratch_bytes"]
    sizes = [GpuPlan(poni, shape, n_radial=b).info["device_bytes"] for b in (30, 31, 32)]
    # Room for the largest plan, that scratch and one request, never for two
    # plans.
    budget = max(sizes) * 21 // 20 + scratch + footprint(Request(poni, fr, n_radial=32))
    print(f"plans {sizes} B, scratch {scratch} B, budget {budget} B", flush=True)
    assert budget < 2 * min(sizes) + scratch
    with Server(n_slots=2, plan_cache_entries=8, hbm_budget_bytes=budget) as srv:
        resps = [srv.result(srv.submit(Request(poni, fr, n_radial=b))) for b in bins]
        st = srv.stats()
    assert st["plan_builds"] == 4 and st["cache_hits"] == 0
    assert st["plans_resident"] == 1
    assert st["bytes_in_use"] <= budget
    for b, resp in zip(bins, resps):
        assert same(resp.intensity, alone(Request(poni, fr, n_radial=b))[0])


def test_plan_key_compares_full_mask():
    poni, shape = GEOS["a"]
    fr = np.random.default_rng(11).poisson(80.0, size=(2, *shape)).astype(np.uint32)
    m1 = np.zeros(shape, np.uint8)
    m2 = m1.copy()
    m2[30, 40] = 1  # one pixel apart
    m3 = m1.copy()  # same contents, another array
    with Server(n_slots=2) as srv:
        r1 = srv.result(srv.submit(Request(poni, fr, n_radial=40, mask=m1)))
        r2 = srv.result(srv.submit(Request(poni, fr, n_radial=40, mask=m2)))
        r3 = srv.result(srv.submit(Request(poni, fr, n_radial=40, mask=m3)))
        st = srv.stats()
    assert st["plan_builds"] == 2 and st["cache_hits"] == 1
    assert same(r1.intensity, r3.intensity)
    assert not same(r1.intensity, r2.intensity)
    assert same(r1.intensity, alone(Request(poni, fr, n_radial=40, mask=m1))[0])
    assert same(r2.intensity, alone(Request(poni, fr, n_radial=40, mask=m2))[0])


def test_plan_key_ignores_switched_off_fields(monkeypatch):
    import dioptrin_amd.serve as S
    real = S.plan_args

    def leftovers(*a, **k):
        p, d, m = real(*a, **k)
        # Values in fields that the descriptor's own settings switch off.
        d.radial_min, d.radial_max = 7.0, -3.0      # has_radial_range = 0
        d.azimuthal_min, d.azimuthal_max = 1.0, 2.0  # 1D plan
        d.n_azimuthal = 99                           # 1D plan
        d.polarization_factor = 0.25                 # has_polarization = 0
        d.n_ss = 9                                   # not supersampled
        return p, d, m

    poni, shape = GEOS["b"]
    fr = np.random.default_rng(12).poisson(80.0, size=(2, *shape)).astype(np.uint32)
    with Server(n_slots=2) as srv:
        a = srv.result(srv.submit(Request(poni, fr, n_radial=40)))
        monkeypatch.setattr(S, "plan_args", leftovers)
        b = srv.result(srv.submit(Request(poni, fr, n_radial=40)))
        st = srv.stats()
    assert st["plan_builds"] == 1 and st["cache_hits"] == 1
    assert same(a.intensity, b.intensity)


def test_plan_cache_reuse_and_eviction():
    poni, shape = GEOS["a"]
    fr = np.random.default_rng(3).poisson(50.0, size=(2, *shape)).astype(np.uint32)
    bins = [30, 31, 32, 33, 30, 31]
    with Server(n_slots=2, plan_cache_entries=2) as srv:
        resps = [srv.result(srv.submit(Request(poni, fr, n_radial=b))) for b in bins]
        st = srv.stats()
    # 30, 31 built; 32 and 33 evict them; 30 and 31 are rebuilt.
    assert st["plan_builds"] == 6 and st["cache_hits"] == 0
    with Server(n_slots=2, plan_cache_entries=4) as srv:
        for b in bins:
            srv.result(srv.submit(Request(poni, fr, n_radial=b)))
        st = srv.stats()
    assert st["plan_builds"] == 4 and st["cache_hits"] == 2
    for b, resp in zip(bins, resps):
        assert same(resp.intensity, alone(Request(poni, fr, n_radial=b))[0])


def test_bad_request_fails_alone():
    poni, shape = GEOS["b"]
    fr = np.random.default_rng(4).poisson(50.0, size=(1, *shape)).astype(np.uint32)
    with Server(n_slots=2) as srv:
        good1 = srv.submit(Request(poni, fr, n_radial=40))
        # q units need a positive wavelength (integration.rs:154-169).
        bad = srv.submit(Request(dict(poni, wavelength=0.0), fr, n_radial=40, unit="q_nm^-1"))
        std2d = srv.submit(Request(poni, fr, n_radial=20, n_azimuthal=8, azimuthal_std=True))
        good2 = srv.submit(Request(poni, fr, n_radial=41))
        assert srv.result(good1).intensity.shape == (1, 40)
        with pytest.raises(Exception):
            srv.result(bad)
        with pytest.raises(Exception):
            srv.result(std2d)
        assert srv.result(good2).intensity.shape == (1, 41)
        st = srv.stats()
    assert st["failed"] == 2 and st["completed"] == 2


def test_poisson_load_reports_latency():
    P = pools(seed=7, n=16)
    mix = LoadMix(geometries=P, frames_min=1, frames_max=16, bins_1d=(100, 300),
                  grid_2d=(50, 36))
    reqs = make_requests(mix, 30, seed=8)
    with Server(n_slots=4) as srv:
        res = run_load(srv, reqs, rate=400.0, seed=9, keep=True)
    assert res["completed"] == 30 and res["failed"] == 0
    assert 0 < res["p50_ms"] <= res["p99_ms"] <= res["max_ms"]
    assert res["frames"] == res["frames_total"]
    for r, resp in zip(reqs[:6], res["responses"][:6]):
        assert same(resp.intensity, alone(r)[0])


def test_device_frames_equal_host_frames():
    from dioptrin_amd.device import DeviceArray
    reqs = mixed_requests(seed=6, n=9)
    with Server(n_slots=3) as srv:
        host = [srv.result(srv.submit(r)) for r in reqs]
        dev_reqs = []
        for r in reqs:
            d = dict(r.__dict__)
            d["frames"] = DeviceArray.from_numpy(np.ascontiguousarray(r.frames))
            dev_reqs.append(Request(**d))
        dev = [srv.result(t) for t in [srv.submit(r) for r in dev_reqs]]
    for a, b in zip(host, dev):
        assert same(a.intensity, b.intensity)
        assert same(a.errors, b.errors)
        assert same(a.azimuthal_std, b.azimuthal_std)
