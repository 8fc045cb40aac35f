#!/usr/bin/env python3
"""Benchmark of the batched integration hot path on one MI355X.

Workload eiger9m_batch1d: 256 distinct uint32 frames of 3269 x 3110 pixels
(75 um, d = 0.25 m, lambda = 0.4 A, beam at the detector centre), 2000 bins of
pixel splitting, one device-built table shared by every frame. A step is one
batch1d pass over the 256 device-resident frames: the fused CSR x frame-tile
gather, tile picked by the autotuner, replayed from a hipGraph.

Prints one JSON line (value = frames/s of the whole job). Extra keys report
the plain per-frame GPU path, the table build, the roofline of the dominant
kernel and the CPU baseline (the restated reference CPU batch1d path timed on
a bounded sample on this host's cores).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOAD = "eiger9m_batch1d"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes_per_frame(npix, nnz, bins, tile, px_bytes=4, outputs=1):
    # frame once + table (u32 col, f32 weight, u32 row_ptr, f64 den) once per
    # `tile` frames (the frames that share one read of the table) + f64 outputs.
    return px_bytes * npix + (8 * nnz + 4 * (bins + 1) + 8 * bins) / tile + 8 * bins * outputs


def host_frame_rates(plan, frames_dev, n, errors):
    """frames/s of plan.integrate_stack on n frames held in host memory,
    pageable and page-locked (best of two runs each)."""
    from dioptrin_amd.device import PinnedArray
    host = frames_dev[:n].numpy()
    pin = PinnedArray(host.shape, host.dtype)
    pin.array[...] = host
    out = {}
    for name, src in (("pageable", host), ("pinned", pin.array)):
        best = 0.0
        for _ in range(2):
            t0 = time.perf_counter()
            plan.integrate_stack(src, errors=errors)
            best = max(best, n / (time.perf_counter() - t0))
        out[name] = round(best, 1)
    del pin
    return out


def dynamic_mask_rate(plan, frames_dev, n, errors):
    """frames/s of plan.integrate_masked_device on n device frames, each with
    its own mask (5 % of the pixels, a different set per frame); best of
    three runs. The plan has no mask of its own, so this is the tile-kernel
    route."""
    from dioptrin_amd.device import DeviceArray, synchronize
    shape = frames_dev.shape[1:]
    base = (np.random.default_rng(3).random(shape[0] * shape[1]) < 0.05).astype(np.uint8)
    masks = DeviceArray.from_numpy(np.stack([np.roll(base, 977 * k).reshape(shape)
                                             for k in range(n)]))
    oi = DeviceArray((n, plan.nbins), np.float64)
    os_ = DeviceArray((n, plan.nbins), np.float64) if errors else None
    fr = frames_dev[:n]
    plan.integrate_masked_device(fr, masks, oi, os_, errors=errors)
    synchronize()
    best = 0.0
    for _ in range(3):
        t0 = time.perf_counter()
        plan.integrate_masked_device(fr, masks, oi, os_, errors=errors)
        synchronize()
        best = max(best, n / (time.perf_counter() - t0))
    return round(best, 1)


def std_bytes_per_frame(npix, nnz, bins, frames_per_pass, px_bytes=4):
    """Algorithmic bytes per frame of the azimuthal spread on the tiled
    table: the frame once, the table (u32 column + f32 weight per entry, row
    pointers) and the float64 pixel normalization once per pass of
    frames_per_pass frames, and three float64 outputs (I, sigma, std)."""
    return (px_bytes * npix + (8 * nnz + 4 * (bins + 1) + 8 * npix) / frames_per_pass
            + 8 * bins * 3)


def azimuthal_std_rate(plan, frames_dev, n):
    """frames/s of plan.integrate_device with errors and the azimuthal std on
    n device frames (the whole call: spread kernel and bin merge); best of
    three runs."""
    from dioptrin_amd.device import DeviceArray, synchronize
    outs = [DeviceArray((n, plan.nbins), np.float64) for _ in range(3)]
    fr = frames_dev[:n]
    plan.integrate_device(fr, *outs, errors=True, azimuthal_std=True)
    synchronize()
    best = 0.0
    for _ in range(3):
        t0 = time.perf_counter()
        plan.integrate_device(fr, *outs, errors=True, azimuthal_std=True)
        synchronize()
        best = max(best, n / (time.perf_counter() - t0))
    return round(best, 1)


def uint16_rate(plan, shape, n):
    """frames/s of plan.integrate_device on n uint16 frames (random values,
    made on the host and copied to the device once), eager calls of the
    whole batch, best of three."""
    from dioptrin_amd.device import DeviceArray, synchronize
    h = np.random.default_rng(16).integers(0, 30000, size=(n,) + tuple(shape), dtype=np.uint16)
    fr = DeviceArray.from_numpy(h)
    del h
    out = DeviceArray((n, plan.nbins), np.float64)
    plan.integrate_device(fr, out)
    synchronize()
    best = 0.0
    for _ in range(3):
        t0 = time.perf_counter()
        plan.integrate_device(fr, out)
        synchronize()
        best = max(best, n / (time.perf_counter() - t0))
    return round(best, 1)


def host_cpu():
    """nproc, the model name from /proc/cpuinfo and the cores this job may
    use (OMP_NUM_THREADS when set, as the GPU host sets it to the job's share
    of its cores; else every core)."""
    n = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    share = int(os.environ.get("OMP_NUM_THREADS") or n)
    return n, model, max(1, min(n, share))


def cpu_baseline(frames_dev, poni, w, seconds=12.0):
    """Restated reference CPU batch1d (CSC, 4-frame tiles) on the cores this
    job may use, and the strict single-thread FP64 oracle gather beside it."""
    import oracle as O
    nproc, model, threads = host_cpu()
    t0 = time.time()
    lut = O.build_lut1d(O.Poni(**poni), *w["shape"], w["n_radial"])
    build_s = time.time() - t0
    norm = O.normalization(O.Poni(**poni), *w["shape"], w["polarization"])
    csc = O.CscBaseline(lut, w["shape"][0] * w["shape"][1], None, norm)
    nf = 4 * threads
    host = frames_dev[:nf].numpy().reshape(nf, -1)
    done, t0 = 0, time.time()
    while True:
        csc.batch_u32(host, threads)
        done += nf
        el = time.time() - t0
        if el >= seconds:
            break
    # Strict oracle: sequential FP64 CSR rows, one thread, a few frames.
    den = O.denominators(lut, norm)
    sx = [host[k].astype(np.float64) for k in range(4)]
    ns, t1 = 0, time.time()
    while ns < 4 or time.time() - t1 < 2.0:
        O.gather(lut, den, sx[ns % 4], False)
        ns += 1
    strict = ns / (time.time() - t1)
    return {"value": done / el, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"{done} frames ({nf} distinct frames x {done // nf} passes) of the "
                      f"same workload, {el:.1f} s wall",
            "nproc": nproc, "cpu_model": model,
            "strict_oracle_1thread_fps": strict,
            "lut_build_ms": build_s * 1e3}


SERVE_GEOMETRIES = ("pilatus1m_ps1d", "eiger4m_ss2d", "eiger9m_batch1d")


def serve_bench(rates, n_requests, seed=7):
    """Continuous-batching server over the serving mix: Poisson arrivals,
    the three detector shapes, 1-64 device-resident uint32 frames per
    request, 60 % 1D (1000-3000 bins) / 40 % 2D (500 x 360), pixel
    splitting. Tables are built once (warm cache) before the timed loads; an
    infinite rate submits every request at once (the saturating load); a
    cold-start load on a fresh server builds its plans during the load. The
    first rate's requests are also run through the plain per-frame path."""
    from dioptrin_amd import GpuPlan, synth
    from dioptrin_amd.device import synchronize
    from dioptrin_amd.serve import LoadMix, Server, make_requests, run_load

    geos = {}
    for name in SERVE_GEOMETRIES:
        w = synth.WORKLOADS[name]
        geos[name] = (synth.workload_poni(name), w["shape"],
                      synth.frames_device(name, 64, seed=11 + len(geos)))
    synchronize()
    mix = LoadMix(geometries=geos, frames_min=1, frames_max=64, p1d=0.6, bins_1d=(1000, 3000),
                  grid_2d=(500, 360), method="pixel_split", precision="f32_wide")
    out = {"requests_per_rate": n_requests, "geometries": list(SERVE_GEOMETRIES),
           "frames_per_request": [1, 64], "mix": "60% 1D 1000-3000 bins / 40% 2D 500x360",
           "frames": "device-resident uint32", "rates": []}
    with Server(n_slots=4, plan_cache_entries=64, hbm_budget_bytes=32 << 30) as srv:
        # Warm the plan cache: one 1-frame request per distinct plan.
        all_reqs = [make_requests(mix, n_requests, seed + k) for k in range(len(rates))]
        seen = {}
        for reqs in all_reqs:
            for r in reqs:
                key = (id(r.poni), r.n_radial, r.n_azimuthal)
                if key not in seen:
                    seen[key] = r
        t0 = time.perf_counter()
        for r in seen.values():
            srv.result(srv.submit(type(r)(**{**r.__dict__, "frames": r.frames[:1]})))
        warm_s = time.perf_counter() - t0
        # Then, per plan, its largest request of the load on every slot at
        # once: each slot's scratch is sized and every kernel the load uses has
        # run once before the timed load.
        big = {}
        for reqs in all_reqs:
            for r in reqs:
                key = (id(r.poni), r.n_radial, r.n_azimuthal)
                n = 1 if len(r.frames.shape) == 2 else int(r.frames.shape[0])
                if key not in big or n > big[key][0]:
                    big[key] = (n, r)
        for _, r in big.values():
            for t in [srv.submit(r) for _ in range(4)]:
                srv.result(t)
        out["plans"] = len(seen)
        out["plan_build_s"] = round(warm_s, 2)
        for rate, reqs in zip(rates, all_reqs):
            base = srv.stats()
            res = run_load(srv, reqs, rate=rate, seed=seed)
            frames = res["frames_total"]
            out["rates"].append({
                "offered_req_s": rate if np.isfinite(rate) else "saturating: all at once",
                "achieved_req_s": round(len(reqs) / res["wall_s"], 1),
                "frames": frames,
                "frames_per_s": round(res["frames_per_s_wall"], 1),
                "p50_ms": round(res["p50_ms"], 3), "p99_ms": round(res["p99_ms"], 3),
                "max_ms": round(res["max_ms"], 3),
                "peak_inflight": res["peak_inflight"],
                "plan_builds_during_load": res["plan_builds"] - base["plan_builds"]})
            log(f"[bench] serve {rate} req/s: {out['rates'][-1]}")
        st = srv.stats()
        out["device_bytes"] = {"plans": st["plan_bytes"], "slot_scratch": st["scratch_bytes"],
                               "peak_total": st["peak_bytes"],
                               "plans_resident": st["plans_resident"]}
    # Cold start: a fresh server (empty plan cache) under the first rate's
    # load; plans are built on the server's builder thread while requests of
    # already-built plans keep running.
    with Server(n_slots=4, plan_cache_entries=64, hbm_budget_bytes=32 << 30) as srv:
        res = run_load(srv, all_reqs[0], rate=rates[0], seed=seed)
        out["cold_start"] = {
            "offered_req_s": rates[0], "plan_builds_during_load": res["plan_builds"],
            "frames_per_s": round(res["frames_per_s_wall"], 1),
            "p50_ms": round(res["p50_ms"], 3), "p99_ms": round(res["p99_ms"], 3),
            "max_ms": round(res["max_ms"], 3)}
        log(f"[bench] serve cold start: {out['cold_start']}")
    # The same requests of the first rate through the plain per-frame path:
    # one plan per request shape (built beforehand), one call and one
    # synchronisation per frame.
    plans = {}
    for r in seen.values():
        plans[(id(r.poni), r.n_radial, r.n_azimuthal)] = GpuPlan(
            r.poni, r.frames.shape[1:], n_radial=r.n_radial, n_azimuthal=r.n_azimuthal,
            precision="f32_wide")
    from dioptrin_amd.device import DeviceArray
    outs = {}
    for key, pl in plans.items():
        pl.set_tile(1, 1)
        outs[key] = DeviceArray((1, pl.nbins), np.float64)
    synchronize()
    nfr, t0 = 0, time.perf_counter()
    for r in all_reqs[0]:
        key = (id(r.poni), r.n_radial, r.n_azimuthal)
        for k in range(r.frames.shape[0]):
            plans[key].integrate_device(r.frames[k:k + 1], outs[key])
            synchronize()
            nfr += 1
    out["per_frame_path_frames_per_s"] = round(nfr / (time.perf_counter() - t0), 1)
    return out


def bslz4_bench(plan, frames, errors, n_chunks=64, distinct=16, reps=10):
    """Bitshuffle + LZ4 leg: `distinct` of the step's frames written as raw
    HDF5 chunks (8 KiB blocks, the library's LZ4 encoder, 16 host threads),
    repeated to n_chunks chunks resident in device memory with their block
    descriptors. Timed: the device decode alone and decode + integrate of the
    n_chunks frames; the decoded frames are checked against the originals."""
    from concurrent.futures import ThreadPoolExecutor

    from dioptrin_amd import bslz4
    from dioptrin_amd.device import DeviceArray, synchronize
    host = frames[:distinct].numpy()
    t0 = time.time()
    with ThreadPoolExecutor(16) as ex:
        distinct_chunks = list(ex.map(bslz4.compress, host))
    log(f"[bench] bslz4: {distinct} chunks written in {time.time() - t0:.1f} s")
    chunks = [distinct_chunks[k % distinct] for k in range(n_chunks)]
    comp = sum(len(c) for c in chunks)
    shape = tuple(frames.shape[1:])
    npix = shape[0] * shape[1]
    dc = bslz4.DeviceChunks(chunks, shape, "uint32")
    out = DeviceArray((n_chunks, *shape), np.uint32)
    dc.decode(out)
    synchronize()
    dc.check()
    if not np.array_equal(out[:distinct].numpy(), host):
        raise RuntimeError("bslz4: decoded frames differ from the originals")
    out_i = DeviceArray((n_chunks, plan.nbins), np.float64)
    out_s = DeviceArray((n_chunks, plan.nbins), np.float64) if errors else None
    t0 = time.perf_counter()
    for _ in range(reps):
        dc.decode(out)
    synchronize()
    dec_s = (time.perf_counter() - t0) / reps
    for _ in range(2):
        dc.decode(out)
        plan.integrate_device(out, out_i, out_s, errors=errors)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dc.decode(out)
        plan.integrate_device(out, out_i, out_s, errors=errors)
    synchronize()
    both_s = (time.perf_counter() - t0) / reps
    dc.check()
    # Algorithmic bytes of a decode: compressed bytes and descriptors read,
    # decoded frames written.
    alg = comp + dc.n_blocks * 16 + n_chunks * npix * 4
    return {"chunks": n_chunks, "distinct_frames": distinct, "block_bytes": 8192,
            "compression_ratio": round(n_chunks * npix * 4 / comp, 3),
            "decode_ms": round(dec_s * 1e3, 3), "decode_fps": round(n_chunks / dec_s, 1),
            "decode_gbs_algorithmic": round(alg / dec_s / 1e9, 1),
            "decode_frac_hbm_peak": round(alg / dec_s / 1e9 / HBM_PEAK_GBS, 4),
            "decode_integrate_fps": round(n_chunks / both_s, 1),
            "frames": "device-resident compressed chunks (uint32 Eiger 9M frames of the step)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="f32_wide")
    ap.add_argument("--tile", default="", help="frames per pass[,frames per lane] "
                                               "(default: autotune)")
    ap.add_argument("--frames", type=int, default=256)
    ap.add_argument("--errors", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop section")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--no-bslz4", action="store_true",
                    help="skip the Bitshuffle+LZ4 decode section")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy kernel of MIB MiB 5 times "
                         "(reference for the memory-traffic counters) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        from dioptrin_amd.device import DeviceArray, synchronize
        n = args.calibrate_copy << 20
        a, b = DeviceArray((n,), np.uint8), DeviceArray((n,), np.uint8)
        from dioptrin_amd import _native as N
        for _ in range(5):
            N.check(N.lib().dio_memcpy(a._ctx, b.ptr, a.ptr, n, 2))
        synchronize()
        print(json.dumps({"calibrate_copy_bytes": n, "launches": 5}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from dioptrin_amd import GpuPlan, device_info, synth
    from dioptrin_amd.device import DeviceArray, synchronize

    w = synth.WORKLOADS[WORKLOAD]
    poni = synth.workload_poni(WORKLOAD)
    npix = w["shape"][0] * w["shape"][1]
    nfr = args.frames
    log(f"[bench] device {device_info(0)['name']} {device_info(0)['arch']}")
    t0 = time.time()
    frames = synth.frames_device(WORKLOAD, nfr, seed=2026)
    synchronize()
    log(f"[bench] generated {nfr} frames in {time.time() - t0:.1f} s")

    t0 = time.time()
    plan = GpuPlan(poni, w["shape"], n_radial=w["n_radial"], precision=args.precision)
    synchronize()
    build_wall = (time.time() - t0) * 1e3
    info = plan.info
    log(f"[bench] table: nnz {info['nnz']} segments {info['n_segments']} build "
        f"{info['build_ms']:.1f} ms uncertain {info['uncertain_angles']} singular "
        f"{info['singular_pixels']}")
    if args.tile:
        tile, lanes = (list(map(int, args.tile.split(","))) + [0])[:2]
        plan.set_tile(tile, lanes)
        lanes = lanes or min(tile, 4)
    else:
        tile, lanes = plan.autotune("uint32", 64, errors=args.errors)
    log(f"[bench] frames per pass {tile}, frames per lane {lanes}")

    bins = plan.nbins
    out_i = DeviceArray((nfr, bins), np.float64)
    out_s = DeviceArray((nfr, bins), np.float64) if args.errors else None

    if args.no_graph:
        def step():
            plan.integrate_device(frames, out_i, out_s, errors=args.errors)
    else:
        graph = plan.capture_graph(frames, out_i, out_s, errors=args.errors)

        def step():
            graph.launch()

    for _ in range(args.warmup):
        step()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    synchronize()
    el = time.perf_counter() - t0
    ms_per_step = el * 1e3 / args.steps
    fps = nfr * args.steps / el
    log(f"[bench] batched: {ms_per_step:.3f} ms/step, {fps:.0f} frames/s")

    # Dominant kernel: the gather kernel of every launch of the same step
    # (same frames, same order), timed with HIP events recorded on the plan's
    # stream around each launch, in three eager runs of the step right after
    # the timed region. A launch covers one pass, or two passes that share one
    # read of the table (the double-buffered kernel).
    launches = []
    for _ in range(3):
        launches += plan.time_passes(frames, out_i, out_s, errors=args.errors)

    # Plain per-frame path: one call and one synchronisation per frame.
    plan1 = plan
    plan1.set_tile(1, 1)
    n1 = min(nfr, 64)
    for k in range(4):
        plan1.integrate_device(frames[k:k + 1], out_i[k:k + 1], None if out_s is None else
                               out_s[k:k + 1], errors=args.errors)
    synchronize()
    t0 = time.perf_counter()
    for k in range(n1):
        plan1.integrate_device(frames[k:k + 1], out_i[k:k + 1], None if out_s is None else
                               out_s[k:k + 1], errors=args.errors)
        synchronize()
    per_frame_fps = n1 / (time.perf_counter() - t0)
    plan.set_tile(tile, lanes)
    log(f"[bench] per-frame: {per_frame_fps:.0f} frames/s")

    # Frames handed over in host memory (dio_integrate): the PCIe-inclusive
    # rate of the batched path, from pageable and from page-locked buffers.
    host_fps = host_frame_rates(plan, frames, min(nfr, 64), args.errors)
    log(f"[bench] host frames: {host_fps}")
    dyn_fps = dynamic_mask_rate(plan, frames, min(nfr, 64), args.errors)
    log(f"[bench] per-frame masks: {dyn_fps} frames/s")
    try:
        u16_fps = uint16_rate(plan, w["shape"], min(nfr, 64))
    except Exception as e:  # the 16-bit leg is a side measurement
        log(f"[bench] uint16 frames: not measured ({e})")
        u16_fps = None
    log(f"[bench] uint16 frames: {u16_fps} frames/s")
    std_fps = azimuthal_std_rate(plan, frames, min(nfr, 64))
    std_bpf = std_bytes_per_frame(npix, info["nnz_masked"], bins, 8)
    log(f"[bench] azimuthal std: {std_fps} frames/s")
    bslz4_res = None
    if not args.no_bslz4:
        bslz4_res = bslz4_bench(plan, frames, args.errors)
        log(f"[bench] bslz4: {bslz4_res}")

    fl = max(f for _, f in launches)  # frames of a full launch = frames per table read
    full = [ms for ms, f in launches if f == fl]
    seg_ms = sum(full) / len(full)
    per_launch = algorithmic_bytes_per_frame(npix, info["nnz_masked"], bins, fl,
                                             outputs=1 + int(args.errors)) * fl
    achieved = per_launch / (seg_ms * 1e-3) / 1e9
    log(f"[bench] gather kernel {seg_ms:.4f} ms per {fl}-frame launch ({len(full)} launches), "
        f"{achieved:.0f} GB/s")
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_eiger9m_batch1d.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        # Counters were taken for one configuration (precision, errors, frames
        # per launch and tile); any other reports no traffic.
        if (tr.get("precision") == args.precision and tr.get("frames_per_launch") == fl
                and tr.get("errors") == bool(args.errors)):
            for e in tr.get("entries", []):
                if e.get("tile") == [tile, lanes]:
                    traffic = e.get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline ...")
        cpu = cpu_baseline(frames, poni, w)
        log(f"[bench] CPU baseline {cpu['value']:.1f} frames/s on {cpu['cores']} threads")

    serve = None
    if not args.no_serve:
        serve = serve_bench([50.0, 200.0, 400.0, float("inf")], args.serve_requests)

    bytes_frame = algorithmic_bytes_per_frame(npix, info["nnz_masked"], bins, fl,
                                              outputs=1 + int(args.errors))
    line = {
        "metric": "frames/s + HBM GB/s (% peak), batch1d 9M-px frames; p50/p99 serve latency",
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32_compensated": "f32 compensated (f32 weights and signal, f32 partials)",
                  "f32_wide": "f32 (f32 weights and signal, f64 sums of segments of <= 48 "
                              "entries, f32 partials, f64 bin reduction)",
                  "f32": "f32", "f64": "f64"}[args.precision],
        "data": "synthetic",
        "config": {"workload": f"{WORKLOAD}: {nfr} x {w['shape'][0]}x{w['shape'][1]} uint32, "
                               f"{bins} bins pixel_split, device-resident frames",
                   "frames_per_step": nfr, "frames_per_tile": tile,
                   "frames_per_launch": fl,
                   "frames_per_lane": lanes,
                   "precision": args.precision, "errors": bool(args.errors),
                   "graph": not args.no_graph, "table_entries": info["nnz_masked"]},
        "hbm_gbs_algorithmic": round(bytes_frame * fps / 1e9, 1),
        "hbm_pct_peak": round(100 * bytes_frame * fps / 1e9 / HBM_PEAK_GBS, 1),
        "per_frame_fps": round(per_frame_fps, 1),
        "host_frames_fps_pcie": host_fps,
        "dynamic_masks_fps": dyn_fps,
        "uint16_frames_fps": u16_fps,
        "azimuthal_std": {"fps": std_fps, "outputs": "intensity, sigma, std",
                          "frames_per_pass": 8,
                          "bytes_per_frame_algorithmic": int(std_bpf),
                          "gbs_algorithmic": round(std_bpf * std_fps / 1e9, 1),
                          "frac_hbm_peak": round(std_bpf * std_fps / 1e9 / HBM_PEAK_GBS, 4)},
        "batched_vs_per_frame": round(fps / per_frame_fps, 2),
        "lut_build_ms_gpu": round(info["build_ms"], 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": "k_tile (detector-tiled gather, one launch)",
                     "kernel_ms": round(seg_ms, 5), "bytes_per_launch": int(per_launch)},
        "cpu_baseline": None if cpu is None else {k: (round(v, 2) if isinstance(v, float) else v)
                                                  for k, v in cpu.items()},
        "serve": serve,
        "bslz4": bslz4_res,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
