// Continuous-batching integration server (no reference counterpart beyond
// the reference's single-entry table caches, src/integration/cache.rs:42-96,
// and its per-plan streams, src/integration/cuda/device.rs:1354-1584).
//
// Requests (geometry + plan descriptor + frames) arrive at any time. A
// scheduler thread attaches each to a device-resident plan from an LRU plan
// cache; plans not yet resident are built by a separate builder thread, so a
// cold plan delays only the requests that need it. Requests whose plan is
// ready are admitted in arrival order whenever a stream slot is free and the
// server's device bytes (resident plans + slot scratch + frames and outputs
// of requests in flight) stay within the HBM budget, evicting idle plans when
// that makes room. Copy-in, gather and copy-out go on the slot's stream, so
// requests of different shapes, bin counts and 1D/2D modes overlap on the
// device. Completion is detected by events; latency is submit-to-results.
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

struct Entry;

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
  Entry* entry = nullptr;  // plan cache entry, held (users) until the request ends
  int slot = -1;
  uint64_t footprint = 0;
  uint32_t nbins = 0;
  void* dmem = nullptr;
};

// One plan of the cache. The key holds every descriptor field that shapes the
// plan and the full mask bytes, so two requests share an entry only if their
// plans are identical.
struct Entry {
  enum State { kBuilding, kReady, kFailed };
  std::string key;
  size_t hash = 0;
  State state = kBuilding;
  dio_plan* plan = nullptr;
  uint64_t bytes = 0;  // device bytes of the plan once built
  uint32_t users = 0;  // queued and running requests attached to this entry
  int code = DIO_OK;   // build failure
  std::string error;
  // Build inputs, released once the build ends.
  dio_poni poni{};
  dio_plan_desc desc{};
  std::vector<uint8_t> mask;
};

struct Slot {
  Entry* entry = nullptr;  // plan of the running request
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  // Gather partials in stream-ordered memory of this slot (hipMallocAsync /
  // hipFreeAsync on the slot's stream: growing it never synchronises the
  // device, so the other slots keep running). part_n is what is allocated,
  // part_want what admission reserved for the request being launched.
  dio::TileScratch scratch;
  double* part = nullptr;
  size_t part_n = 0, part_want = 0;
  double* pinned = nullptr;  // outputs staged for async device-to-host copies
  size_t pinned_n = 0;
  std::shared_ptr<Req> req;
};

// Output staging per slot, allocated with the server: 128 MiB holds 64 frames
// of a 500 x 360 grid; larger requests grow it (page-locking is slow, so it
// grows by doubling and is kept).
constexpr size_t kMinPinned = (128u << 20) / sizeof(double);

template <class T>
void put(std::string& k, T v) {
  k.append(reinterpret_cast<const char*>(&v), sizeof(v));
}

// Plan identity, serialised field by field (no struct padding, no mask
// pointer; fields a setting switches off are left out, so leftover values in
// them do not split the cache), followed by the full mask bytes.
std::string plan_key(const dio_poni& p, const dio_plan_desc& d, const std::vector<uint8_t>& mask) {
  std::string k;
  k.reserve(160 + mask.size());
  for (double v : {p.pixel1, p.pixel2, p.distance, p.poni1, p.poni2, p.rot1, p.rot2, p.rot3,
                   p.wavelength})
    put(k, v);
  put(k, d.dim1);
  put(k, d.dim2);
  put(k, d.method);
  put(k, d.method == DIO_SUPERSAMPLED ? d.n_ss : 0u);
  put(k, d.unit);
  put(k, d.ndim);
  put(k, d.n_radial);
  put(k, d.ndim == 2 ? d.n_azimuthal : 0u);
  put(k, d.has_radial_range);
  if (d.has_radial_range) {
    put(k, d.radial_min);
    put(k, d.radial_max);
  }
  const bool az = d.ndim == 2 && d.has_azimuthal_range;
  put(k, (int32_t)az);
  if (az) {
    put(k, d.azimuthal_min);
    put(k, d.azimuthal_max);
  }
  put(k, d.has_polarization);
  if (d.has_polarization) put(k, d.polarization_factor);
  put(k, d.precision);
  put(k, d.input_order);
  put(k, (uint64_t)mask.size());
  k.append(reinterpret_cast<const char*>(mask.data()), mask.size());
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

int n_outputs(uint32_t flags) {
  return 1 + ((flags & DIO_WITH_ERRORS) ? 1 : 0) + ((flags & DIO_WITH_AZIMUTHAL_STD) ? 1 : 0);
}

}  // namespace

struct dio_server {
  dio_ctx* ctx = nullptr;
  dio_server_config cfg{};
  std::vector<Slot> slots;
  // Everything below is guarded by mu, except slot contents, which only the
  // scheduler thread touches.
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::shared_ptr<Req>> queue;
  std::unordered_map<uint64_t, std::shared_ptr<Req>> tickets;
  uint64_t next_ticket = 1;
  uint64_t request_bytes = 0, plan_bytes = 0, scratch_bytes = 0, peak_bytes = 0;
  uint64_t completed = 0, failed = 0, frames_done = 0, plan_builds = 0, cache_hits = 0;
  uint32_t inflight = 0, peak_inflight = 0;
  std::vector<double> latency_ms;
  Clock::time_point t_first, t_last;
  bool have_first = false;
  bool stop = false;
  std::thread worker, builder;

  // Plan cache, most recently used first; entries with users (queued or
  // running requests) and entries being built are never evicted. index maps
  // the key hash to its entries; a hit also compares the whole key.
  std::list<Entry> lru;
  std::unordered_multimap<size_t, std::list<Entry>::iterator> index;
  std::deque<Entry*> build_queue;
  uint32_t builds_pending = 0;
  std::vector<dio_plan*> dead;  // evicted plans, freed by the builder thread
  bool stop_builder = false;

  uint64_t total_bytes() const { return request_bytes + plan_bytes + scratch_bytes; }
  void note_peak() { peak_bytes = std::max(peak_bytes, total_bytes()); }

  // Attaches r to the cache entry of its plan, queueing a build on a miss.
  // Called with mu held.
  void attach(Req& r) {
    std::string k = plan_key(r.poni, r.desc, r.mask);
    const size_t h = std::hash<std::string>()(k);
    auto range = index.equal_range(h);
    for (auto it = range.first; it != range.second; ++it) {
      if (it->second->key != k) continue;
      lru.splice(lru.begin(), lru, it->second);
      Entry& e = *lru.begin();
      e.users++;
      r.entry = &e;
      cache_hits++;
      return;
    }
    lru.emplace_front();
    Entry& e = lru.front();
    e.key = std::move(k);
    e.hash = h;
    e.poni = r.poni;
    e.desc = r.desc;
    e.mask = r.mask;
    e.users = 1;
    index.emplace(h, lru.begin());
    r.entry = &e;
    build_queue.push_back(&e);
    builds_pending++;
    cv.notify_all();
  }

  void drop_entry(std::list<Entry>::iterator it) {
    auto range = index.equal_range(it->hash);
    for (auto x = range.first; x != range.second; ++x)
      if (x->second == it) {
        index.erase(x);
        break;
      }
    if (it->state == Entry::kReady) plan_bytes -= it->bytes;
    // Freeing device memory waits for the device; the builder thread does it.
    if (it->plan) dead.push_back(it->plan);
    lru.erase(it);
    cv.notify_all();
  }

  // Releases one user of e; failed entries leave the cache with their last
  // user. Called with mu held.
  void release(Entry* e) {
    e->users--;
    if (e->users == 0 && e->state == Entry::kFailed) {
      for (auto it = lru.begin(); it != lru.end(); ++it)
        if (&*it == e) {
          drop_entry(it);
          break;
        }
    }
  }

  // Evicts idle plans, least recently used first, while more than the entry
  // cap are cached. Called with mu held.
  void trim_locked() {
    const size_t cap = std::max<uint32_t>(1, cfg.plan_cache_entries);
    for (auto it = lru.end(); it != lru.begin() && lru.size() > cap;) {
      --it;
      if (it->users || it->state != Entry::kReady) continue;
      auto next = std::next(it);
      drop_entry(it);
      it = next;
    }
  }

  // Whether need more bytes fit the budget, evicting idle plans (least
  // recently used first) when that makes them fit; nothing is evicted when
  // even dropping every idle plan would not. Called with mu held.
  bool make_room_locked(uint64_t need) {
    const uint64_t budget = cfg.hbm_budget_bytes;
    if (total_bytes() + need <= budget) return true;
    uint64_t idle = 0;
    for (auto& e : lru)
      if (!e.users && e.state == Entry::kReady) idle += e.bytes;
    if (total_bytes() - idle + need > budget) return false;
    for (auto it = lru.end(); it != lru.begin() && total_bytes() + need > budget;) {
      --it;
      if (it->users || it->state != Entry::kReady) continue;
      auto next = std::next(it);
      drop_entry(it);
      it = next;
    }
    return true;
  }

  void build_loop() {
    (void)hipSetDevice(ctx->device);
    while (true) {
      Entry* e = nullptr;
      std::vector<dio_plan*> reap;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop_builder || !build_queue.empty() || !dead.empty(); });
        reap.swap(dead);
        if (!build_queue.empty()) {
          e = build_queue.front();
          build_queue.pop_front();
        } else if (reap.empty()) {
          return;  // stopping, nothing left to build or free
        }
      }
      for (dio_plan* p : reap) delete p;
      if (!e) continue;
      // The entry has users until its build ends, so it stays in the cache.
      dio_plan_desc d = e->desc;
      d.mask = e->mask.empty() ? nullptr : e->mask.data();
      dio_plan* p = nullptr;
      int code = DIO_OK;
      std::string msg;
      try {
        p = dio::make_plan(ctx, e->poni, d);
      } catch (const dio::Error& x) {
        code = x.code;
        msg = x.what();
      } catch (const std::exception& x) {
        code = DIO_ERR_DEVICE;
        msg = x.what();
      }
      std::lock_guard<std::mutex> lk(mu);
      e->mask = std::vector<uint8_t>();
      builds_pending--;
      if (p) {
        e->plan = p;
        e->bytes = dio::plan_device_bytes(p);
        e->state = Entry::kReady;
        plan_builds++;
        plan_bytes += e->bytes;
        note_peak();
        make_room_locked(0);
        trim_locked();
      } else {
        e->state = Entry::kFailed;
        e->code = code;
        e->error = msg;
      }
      cv.notify_all();
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
    request_bytes -= r->footprint;
    inflight--;
    s.req.reset();
    release(s.entry);
    s.entry = nullptr;
    cv.notify_all();
  }

  // Marks r failed. Called with mu held.
  void fail_locked(Req& r, int code, const std::string& msg) {
    r.status = code;
    r.error = msg;
    r.done = true;
    r.t_done = Clock::now();
    failed++;
    if (r.entry) release(r.entry);
    r.entry = nullptr;
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

  // Enqueue one admitted request on slot s. Returns the failure (code 0 on
  // success).
  std::pair<int, std::string> launch(Slot& s, const std::shared_ptr<Req>& r, dio_plan* p) {
    try {
      DIO_HIP(hipSetDevice(ctx->device));
      size_t fbytes = (size_t)r->n_frames * p->gt.npix * dtype_size(r->dtype);
      size_t nb = (size_t)r->n_frames * p->gt.nbins;
      const int nout = n_outputs(r->flags);
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
      if (s.part_n < s.part_want) {
        if (s.part) DIO_HIP(hipFreeAsync(s.part, s.stream));
        s.part = nullptr;
        s.part_n = 0;
        DIO_HIP(hipMallocAsync(reinterpret_cast<void**>(&s.part), s.part_want * sizeof(double),
                               s.stream));
        s.part_n = s.part_want;
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
      return {DIO_OK, {}};
    } catch (const dio::Error& e) {
      settle(s, r);
      return {e.code, e.what()};
    } catch (const std::exception& e) {
      settle(s, r);
      return {DIO_ERR_DEVICE, e.what()};
    }
  }

  // One admission pass over the queue, in arrival order; requests whose plan
  // is still being built are passed over (they keep their place). Returns the
  // requests admitted, each with its slot reserved. Called with mu held.
  std::vector<std::pair<Slot*, std::shared_ptr<Req>>> admit_locked(bool& progressed) {
    std::vector<std::pair<Slot*, std::shared_ptr<Req>>> out;
    std::vector<Slot*> idle;
    for (auto& s : slots)
      if (!s.req) idle.push_back(&s);
    std::reverse(idle.begin(), idle.end());
    for (auto it = queue.begin(); it != queue.end();) {
      Req& r = **it;
      if (!r.entry) attach(r);
      Entry* e = r.entry;
      if (e->state == Entry::kBuilding) {
        ++it;
        continue;
      }
      if (e->state == Entry::kFailed) {
        fail_locked(r, e->code, e->error);
        it = queue.erase(it);
        progressed = true;
        continue;
      }
      const uint64_t expect =
          (uint64_t)r.desc.n_radial * (r.desc.ndim == 2 ? r.desc.n_azimuthal : 1);
      if (e->plan->gt.nbins != expect) {
        fail_locked(r, DIO_ERR_INVALID, "plan bin count differs from the request's");
        it = queue.erase(it);
        progressed = true;
        continue;
      }
      if (idle.empty()) break;
      Slot* s = idle.back();
      dio_plan* p = e->plan;
      const bool on_device = r.flags & DIO_DEVICE_FRAMES;
      const uint64_t foot = (on_device ? 0 : r.n_frames * p->gt.npix * dtype_size(r.dtype)) +
                            r.n_frames * (uint64_t)p->gt.nbins * 8 * n_outputs(r.flags);
      // The slot's partial scratch grows (by doubling, and is kept) when this
      // plan's tiles need more than it holds.
      const size_t need = dio::TileScratch::need(p->tt, p->tile, true);
      const size_t want = s->part_n < need ? std::max(need, 2 * s->part_n) : s->part_n;
      const uint64_t grow = (uint64_t)(want - s->part_n) * sizeof(double);
      // HBM admission: evict idle plans or wait for running requests to
      // finish; with nothing in flight the request runs whatever its size.
      if (!make_room_locked(foot + grow) && inflight > 0) break;
      std::shared_ptr<Req> rp = *it;
      it = queue.erase(it);
      idle.pop_back();
      r.footprint = foot;
      r.slot = (int)(s - slots.data());
      s->part_want = want;
      request_bytes += foot;
      scratch_bytes += grow;
      note_peak();
      inflight++;
      peak_inflight = std::max(peak_inflight, inflight);
      // Move the entry to the front: it is in use.
      for (auto li = lru.begin(); li != lru.end(); ++li)
        if (&*li == e) {
          lru.splice(lru.begin(), lru, li);
          break;
        }
      s->entry = e;
      s->req = rp;
      out.emplace_back(s, rp);
      progressed = true;
    }
    return out;
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
      std::vector<std::pair<Slot*, std::shared_ptr<Req>>> admitted;
      {
        std::lock_guard<std::mutex> lk(mu);
        admitted = admit_locked(progressed);
        trim_locked();
      }
      for (auto& [s, r] : admitted) {
        auto [code, msg] = launch(*s, r, r->entry->plan);
        if (code == DIO_OK) continue;
        std::lock_guard<std::mutex> lk(mu);
        request_bytes -= r->footprint;
        scratch_bytes -= (uint64_t)(s->part_want - s->part_n) * sizeof(double);
        s->part_want = s->part_n;
        inflight--;
        s->req.reset();
        s->entry = nullptr;
        fail_locked(*r, code, msg);
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
    raw->builder = std::thread([raw] { raw->build_loop(); });
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
  out->bytes_in_use = s->total_bytes();
  out->peak_bytes = s->peak_bytes;
  out->plan_builds = s->plan_builds;
  out->cache_hits = s->cache_hits;
  out->plan_bytes = s->plan_bytes;
  out->scratch_bytes = s->scratch_bytes;
  out->request_bytes = s->request_bytes;
  out->plans_resident = 0;
  for (auto& e : s->lru)
    if (e.state == Entry::kReady) out->plans_resident++;
  out->builds_pending = s->builds_pending;
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
  // The scheduler drains the queue (its requests may still wait for plan
  // builds), then the builder stops.
  if (s->worker.joinable()) s->worker.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop_builder = true;
  }
  s->cv.notify_all();
  if (s->builder.joinable()) s->builder.join();
  (void)hipSetDevice(s->ctx->device);
  free_slots(s);
  for (auto& e : s->lru) delete e.plan;
  for (dio_plan* p : s->dead) delete p;
  delete s;
  return DIO_OK;
}

}  // extern "C"
