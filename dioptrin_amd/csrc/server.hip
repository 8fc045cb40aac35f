// Continuous-batching integration server (no reference counterpart beyond
// the reference's single-entry table caches, src/integration/cache.rs:42-96,
// and its per-plan streams, src/integration/cuda/device.rs:1354-1584).
//
// Requests (geometry + plan descriptor + host frames) arrive at any time. A
// scheduler thread admits them in arrival order whenever a stream slot is free
// and their device footprint (frames + outputs) fits the HBM budget, builds
// or reuses the device-resident table from an LRU plan cache, and enqueues
// copy-in, gather and copy-out on the slot's stream, so requests of
// different shapes, bin counts and 1D/2D modes overlap on the device.
// Completion is detected by events; latency is submit-to-results.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "plan.hpp"

using Clock = std::chrono::steady_clock;

namespace {

struct Req {
  uint64_t ticket = 0;
  dio_poni poni{};
  dio_plan_desc desc{};
  std::vector<uint8_t> mask;
  const void* frames = nullptr;
  int dtype = 0;
  uint64_t n_frames = 0;
  uint32_t flags = 0;
  double* out[3] = {nullptr, nullptr, nullptr};
  Clock::time_point t_submit, t_done;
  bool done = false;
  int status = DIO_OK;
  std::string error;
  // Execution state.
  int slot = -1;
  uint64_t footprint = 0;
  uint32_t nbins = 0;
  void* dmem = nullptr;
};

struct Entry {
  std::string key;
  dio_plan* plan;
  uint32_t users;
};

struct Slot {
  Entry* entry = nullptr;  // plan of the running request
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  // Gather partials in stream-ordered memory of this slot (hipMallocAsync /
  // hipFreeAsync on the slot's stream: growing it never synchronises the
  // device, so the other slots keep running).
  dio::TileScratch scratch;
  double* part = nullptr;
  size_t part_n = 0;
  double* pinned = nullptr;  // outputs staged for async device-to-host copies
  size_t pinned_n = 0;
  std::shared_ptr<Req> req;
};

// Output staging per slot, allocated with the server: 128 MiB holds 64 frames
// of a 500 x 360 grid; larger requests grow it (page-locking is slow, so it
// grows by doubling and is kept).
constexpr size_t kMinPinned = (128u << 20) / sizeof(double);

std::string plan_key(const Req& r) {
  std::string k(reinterpret_cast<const char*>(&r.poni), sizeof(r.poni));
  dio_plan_desc d = r.desc;
  d.mask = nullptr;
  k.append(reinterpret_cast<const char*>(&d), sizeof(d));
  k.append(std::to_string(std::hash<std::string>()(
      std::string(reinterpret_cast<const char*>(r.mask.data()), r.mask.size()))));
  return k;
}

// Host copy of request outputs out of the pinned staging buffer: large copies
// (2D grids of many frames reach tens of MB) are split over four threads.
void parallel_copy(void* dst, const void* src, size_t bytes) {
  constexpr size_t kSplit = 4u << 20;
  constexpr int kParts = 4;
  if (bytes < kSplit) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const size_t part = (bytes / kParts + 63) / 64 * 64;
  std::thread th[kParts - 1];
  for (int i = 1; i < kParts; i++) {
    const size_t off = std::min(bytes, part * i), n = std::min(bytes - off, part);
    th[i - 1] = std::thread([=] {
      std::memcpy(static_cast<char*>(dst) + off, static_cast<const char*>(src) + off, n);
    });
  }
  std::memcpy(dst, src, std::min(bytes, part));
  for (auto& t : th) t.join();
}

size_t dtype_size(int dtype) {
  switch (dtype) {
    case DIO_U16: return 2;
    case DIO_U32: case DIO_I32: case DIO_F32T: return 4;
    case DIO_F64T: return 8;
    default: return 0;
  }
}

}  // namespace

struct dio_server {
  dio_ctx* ctx = nullptr;
  dio_server_config cfg{};
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::shared_ptr<Req>> queue;
  std::unordered_map<uint64_t, std::shared_ptr<Req>> tickets;
  uint64_t next_ticket = 1;
  uint64_t bytes_in_use = 0, peak_bytes = 0;
  uint64_t completed = 0, failed = 0, frames_done = 0, plan_builds = 0, cache_hits = 0;
  uint32_t inflight = 0, peak_inflight = 0;
  std::vector<double> latency_ms;
  Clock::time_point t_first, t_last;
  bool have_first = false;
  bool stop = false;
  std::thread worker;

  // Device-resident plans, most recently used first; a plan in use by a slot
  // is never evicted.
  std::list<Entry> lru;
  std::unordered_map<std::string, std::list<Entry>::iterator> cache;

  Entry* plan_for(const Req& r) {
    std::string k = plan_key(r);
    auto it = cache.find(k);
    if (it != cache.end()) {
      lru.splice(lru.begin(), lru, it->second);
      cache_hits++;
      return &*lru.begin();
    }
    dio_plan_desc d = r.desc;
    d.mask = r.mask.empty() ? nullptr : r.mask.data();
    dio_plan* p = dio::make_plan(ctx, r.poni, d);
    plan_builds++;
    lru.push_front(Entry{k, p, 0});
    cache[k] = lru.begin();
    evict();
    return &*lru.begin();
  }

  void evict() {
    size_t cap = std::max<uint32_t>(1, cfg.plan_cache_entries);
    auto it = lru.end();
    while (lru.size() > cap && it != lru.begin()) {
      --it;
      if (it->users) continue;
      cache.erase(it->key);
      delete it->plan;
      it = lru.erase(it);
    }
  }

  void finish(Slot& s) {
    auto r = s.req;
    size_t nb = (size_t)r->n_frames * r->nbins;
    int k = 0;
    for (int o = 0; o < 3; o++) {
      if (!r->out[o]) continue;
      parallel_copy(r->out[o], s.pinned + k * nb, nb * sizeof(double));
      k++;
    }
    r->t_done = Clock::now();
    std::lock_guard<std::mutex> lk(mu);
    r->done = true;
    completed++;
    frames_done += r->n_frames;
    latency_ms.push_back(std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count());
    t_last = r->t_done;
    bytes_in_use -= r->footprint;
    inflight--;
    s.req.reset();
    s.entry->users--;
    s.entry = nullptr;
    cv.notify_all();
  }

  void fail_req(const std::shared_ptr<Req>& r, int code, const std::string& msg) {
    std::lock_guard<std::mutex> lk(mu);
    r->status = code;
    r->error = msg;
    r->done = true;
    r->t_done = Clock::now();
    failed++;
    cv.notify_all();
  }

  // After a failed enqueue: let whatever reached the stream finish, then
  // release the request's device memory.
  void settle(Slot& s, const std::shared_ptr<Req>& r) {
    (void)hipStreamSynchronize(s.stream);
    if (r->dmem) {
      (void)hipFree(r->dmem);
      r->dmem = nullptr;
    }
  }

  // Enqueue one admitted request on slot s. Returns false if it failed.
  bool launch(Slot& s, const std::shared_ptr<Req>& r, dio_plan* p) {
    try {
      DIO_HIP(hipSetDevice(ctx->device));
      size_t fbytes = (size_t)r->n_frames * p->gt.npix * dtype_size(r->dtype);
      size_t nb = (size_t)r->n_frames * p->gt.nbins;
      int nout = 1 + ((r->flags & DIO_WITH_ERRORS) ? 1 : 0) +
                 ((r->flags & DIO_WITH_AZIMUTHAL_STD) ? 1 : 0);
      uint64_t expect = (uint64_t)r->desc.n_radial * (r->desc.ndim == 2 ? r->desc.n_azimuthal : 1);
      if (p->gt.nbins != expect) dio::fail(DIO_ERR_INVALID, "plan bin count differs from the request's");
      r->nbins = p->gt.nbins;
      if (s.pinned_n < nb * nout) {  // the slot is idle here: its stream has drained
        const size_t n = std::max({nb * nout, 2 * s.pinned_n, kMinPinned});
        if (s.pinned) (void)hipHostFree(s.pinned);
        s.pinned = nullptr;
        s.pinned_n = 0;
        DIO_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.pinned), n * sizeof(double),
                              hipHostMallocDefault));
        s.pinned_n = n;
      }
      const size_t need = dio::TileScratch::need(p->tt, p->tile, true);
      if (s.part_n < need) {
        const size_t n = std::max(need, 2 * s.part_n);
        if (s.part) DIO_HIP(hipFreeAsync(s.part, s.stream));
        s.part = nullptr;
        s.part_n = 0;
        DIO_HIP(hipMallocAsync(reinterpret_cast<void**>(&s.part), n * sizeof(double), s.stream));
        s.part_n = n;
      }
      s.scratch.ext = s.part;
      s.scratch.ext_n = s.part_n;
      const bool on_device = r->flags & DIO_DEVICE_FRAMES;
      const size_t fpad = on_device ? 0 : ((fbytes + 255) / 256) * 256;
      DIO_HIP(hipMallocAsync(&r->dmem, fpad + nb * nout * sizeof(double), s.stream));
      char* base = static_cast<char*>(r->dmem);
      double* dout = reinterpret_cast<double*>(base + fpad);
      const void* frames = r->frames;
      if (!on_device) {
        DIO_HIP(hipMemcpyAsync(base, r->frames, fbytes, hipMemcpyHostToDevice, s.stream));
        frames = base;
      }
      double* I = dout;
      double* S = (r->flags & DIO_WITH_ERRORS) ? dout + nb : nullptr;
      double* D = (r->flags & DIO_WITH_AZIMUTHAL_STD) ? dout + nb * (S ? 2 : 1) : nullptr;
      dio::run_device(p, s.scratch, frames, r->dtype, r->n_frames,
                      r->flags & (DIO_WITH_ERRORS | DIO_WITH_AZIMUTHAL_STD), I, S, D, s.stream);
      DIO_HIP(hipMemcpyAsync(s.pinned, dout, nb * nout * sizeof(double), hipMemcpyDeviceToHost,
                             s.stream));
      DIO_HIP(hipFreeAsync(r->dmem, s.stream));
      DIO_HIP(hipEventRecord(s.done, s.stream));
      return true;
    } catch (const dio::Error& e) {
      settle(s, r);
      fail_req(r, e.code, e.what());
    } catch (const std::exception& e) {
      settle(s, r);
      fail_req(r, DIO_ERR_DEVICE, e.what());
    }
    return false;
  }

  void run() {
    (void)hipSetDevice(ctx->device);
    while (true) {
      bool progressed = false;
      // Completions.
      for (auto& s : slots) {
        if (s.req && hipEventQuery(s.done) == hipSuccess) {
          finish(s);
          progressed = true;
        }
      }
      // Admissions, in arrival order.
      while (true) {
        std::shared_ptr<Req> r;
        Slot* free_slot = nullptr;
        {
          std::lock_guard<std::mutex> lk(mu);
          if (stop && queue.empty()) break;
          if (queue.empty()) break;
          for (auto& s : slots)
            if (!s.req) {
              free_slot = &s;
              break;
            }
          if (!free_slot) break;
          r = queue.front();
        }
        Entry* ent = nullptr;
        try {
          ent = plan_for(*r);
        } catch (const dio::Error& e) {
          {
            std::lock_guard<std::mutex> lk(mu);
            queue.pop_front();
          }
          fail_req(r, e.code, e.what());
          progressed = true;
          continue;
        } catch (const std::exception& e) {
          {
            std::lock_guard<std::mutex> lk(mu);
            queue.pop_front();
          }
          fail_req(r, DIO_ERR_DEVICE, e.what());
          progressed = true;
          continue;
        }
        dio_plan* p = ent->plan;
        int nout = 1 + ((r->flags & DIO_WITH_ERRORS) ? 1 : 0) +
                   ((r->flags & DIO_WITH_AZIMUTHAL_STD) ? 1 : 0);
        const bool on_device = r->flags & DIO_DEVICE_FRAMES;
        uint64_t foot = (on_device ? 0 : r->n_frames * p->gt.npix * dtype_size(r->dtype)) +
                        r->n_frames * (uint64_t)p->gt.nbins * 8 * nout;
        {
          std::lock_guard<std::mutex> lk(mu);
          // HBM-footprint admission: wait for running requests to free memory;
          // a request larger than the whole budget runs alone.
          if (inflight > 0 && bytes_in_use + foot > cfg.hbm_budget_bytes) break;
          queue.pop_front();
          r->footprint = foot;
          r->slot = (int)(free_slot - slots.data());
          bytes_in_use += foot;
          peak_bytes = std::max(peak_bytes, bytes_in_use);
          inflight++;
          peak_inflight = std::max(peak_inflight, inflight);
          free_slot->req = r;
          free_slot->entry = ent;
          ent->users++;
        }
        if (!launch(*free_slot, r, p)) {
          std::lock_guard<std::mutex> lk(mu);
          bytes_in_use -= foot;
          inflight--;
          free_slot->req.reset();
          free_slot->entry = nullptr;
          ent->users--;
        }
        progressed = true;
      }
      {
        std::unique_lock<std::mutex> lk(mu);
        if (stop && queue.empty() && inflight == 0) break;
        if (!progressed) cv.wait_for(lk, std::chrono::microseconds(50));
      }
    }
  }
};

extern "C" {

namespace {
// Frees the streams, events and staging buffers of every slot (the slots must
// be idle).
void free_slots(dio_server* s) {
  for (auto& sl : s->slots) {
    if (sl.part) (void)hipFreeAsync(sl.part, sl.stream);
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.pinned) (void)hipHostFree(sl.pinned);
    sl = Slot{};
  }
}
}  // namespace

int dio_server_create(dio_ctx* ctx, const dio_server_config* cfg, dio_server** out) {
  std::unique_ptr<dio_server> s;
  try {
    if (!ctx || !out) throw dio::Error(DIO_ERR_INVALID, "null argument");
    s = std::make_unique<dio_server>();
    s->ctx = ctx;
    dio_server_config c{};
    c.n_slots = 4;
    c.hbm_budget_bytes = 32ull << 30;
    c.plan_cache_entries = 32;
    if (cfg) c = *cfg;
    if (c.n_slots == 0 || c.n_slots > 64) throw dio::Error(DIO_ERR_INVALID, "n_slots must be in 1..64");
    if (c.hbm_budget_bytes == 0) throw dio::Error(DIO_ERR_INVALID, "hbm budget must be positive");
    s->cfg = c;
    DIO_HIP(hipSetDevice(ctx->device));
    s->slots.resize(c.n_slots);
    for (auto& sl : s->slots) {
      DIO_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
      DIO_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
      DIO_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.pinned), kMinPinned * sizeof(double),
                            hipHostMallocDefault));
      sl.pinned_n = kMinPinned;
    }
    dio_server* raw = s.release();
    raw->worker = std::thread([raw] { raw->run(); });
    *out = raw;
    return DIO_OK;
  } catch (const dio::Error& e) {
    if (s) free_slots(s.get());
    dio_set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    if (s) free_slots(s.get());
    dio_set_last_error(e.what());
    return DIO_ERR_DEVICE;
  }
}

int dio_server_submit(dio_server* s, const dio_poni* poni, const dio_plan_desc* desc,
                      const void* frames, int dtype, uint64_t n_frames, uint32_t flags,
                      double* intensity, double* sigma, double* std_out, uint64_t* ticket) {
  if (!s || !poni || !desc || !frames || !intensity || !ticket || n_frames == 0 ||
      dtype_size(dtype) == 0 || ((flags & DIO_WITH_ERRORS) && !sigma) ||
      ((flags & DIO_WITH_AZIMUTHAL_STD) && !std_out)) {
    dio_set_last_error("invalid request");
    return DIO_ERR_INVALID;
  }
  auto r = std::make_shared<Req>();
  r->poni = *poni;
  r->desc = *desc;
  if (desc->mask) r->mask.assign(desc->mask, desc->mask + (size_t)desc->dim1 * desc->dim2);
  r->desc.mask = nullptr;
  r->frames = frames;
  r->dtype = dtype;
  r->n_frames = n_frames;
  r->flags = flags;
  r->out[0] = intensity;
  r->out[1] = (flags & DIO_WITH_ERRORS) ? sigma : nullptr;
  r->out[2] = (flags & DIO_WITH_AZIMUTHAL_STD) ? std_out : nullptr;
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
      dio_set_last_error("server is shutting down");
      return DIO_ERR_INVALID;
    }
    r->ticket = s->next_ticket++;
    if (!s->have_first) {
      s->t_first = r->t_submit;
      s->have_first = true;
    }
    s->tickets[r->ticket] = r;
    s->queue.push_back(r);
    *ticket = r->ticket;
  }
  s->cv.notify_all();
  return DIO_OK;
}

int dio_server_poll(dio_server* s, uint64_t ticket, double timeout_s, int* done,
                    double* latency_ms) {
  if (!s || !done) {
    dio_set_last_error("null argument");
    return DIO_ERR_INVALID;
  }
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) {
    dio_set_last_error("unknown ticket");
    return DIO_ERR_INVALID;
  }
  auto r = it->second;
  auto deadline = Clock::now() + std::chrono::duration<double>(timeout_s);
  while (!r->done && timeout_s > 0 && Clock::now() < deadline)
    s->cv.wait_until(lk, deadline);
  *done = r->done ? 1 : 0;
  if (!r->done) return DIO_OK;
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  s->tickets.erase(it);
  if (r->status != DIO_OK) {
    dio_set_last_error(r->error.c_str());
    return r->status;
  }
  return DIO_OK;
}

int dio_server_stats_get(dio_server* s, dio_server_stats* out) {
  if (!s || !out) {
    dio_set_last_error("null argument");
    return DIO_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof(*out));
  out->completed = s->completed;
  out->failed = s->failed;
  out->frames = s->frames_done;
  out->queued = s->queue.size();
  out->inflight = s->inflight;
  out->peak_inflight = s->peak_inflight;
  out->bytes_in_use = s->bytes_in_use;
  out->peak_bytes = s->peak_bytes;
  out->plan_builds = s->plan_builds;
  out->cache_hits = s->cache_hits;
  std::vector<double> l = s->latency_ms;
  if (!l.empty()) {
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
      size_t i = (size_t)std::min<double>(l.size() - 1, std::ceil(q * l.size()) - 1);
      return l[i];
    };
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    out->max_ms = l.back();
    double sum = 0;
    for (double v : l) sum += v;
    out->mean_ms = sum / l.size();
  }
  if (s->have_first && s->completed) {
    double span = std::chrono::duration<double>(s->t_last - s->t_first).count();
    out->span_s = span;
    out->frames_per_s = span > 0 ? s->frames_done / span : 0.0;
  }
  return DIO_OK;
}

int dio_server_destroy(dio_server* s) {
  if (!s) return DIO_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)hipSetDevice(s->ctx->device);
  free_slots(s);
  for (auto& e : s->lru) delete e.plan;
  delete s;
  return DIO_OK;
}

}  // extern "C"
