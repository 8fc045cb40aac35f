// Internal definitions of the C-ABI handles, shared by the API and the
// serving loop.
#pragma once
#include <mutex>

#include "builder.hpp"
#include "gather.hpp"
#include "tile.hpp"

struct dio_ctx {
  int device = 0;
  hipDeviceProp_t prop{};
  int refs = 0;
};

struct dio_plan {
  dio_ctx* ctx = nullptr;
  int device = 0;
  hipStream_t stream = nullptr;
  dio_plan_desc desc{};
  dio::Geo geo{};
  dio::Built b;
  dio::GatherTable gt;
  dio::StdTable st;
  dio::TiledTable tt;
  dio::TileScratch sc;
  uint32_t tile = 16;   // frames per pass
  uint32_t lanes = 4;   // frames per lane
  double build_ms = 0;
  std::mutex mu;
  dio::DevBuf<uint8_t> stage;
  dio::DevBuf<uint8_t> mstage;  // per-frame masks of dio_integrate_masked
  // Dynamic masks on the tile kernel: zeroed signal and normalization frames,
  // per-frame denominators of one chunk.
  dio::DevBuf<uint8_t> dyn_x, dyn_n;
  dio::DevBuf<double> dyn_d;
  // Dynamic masks folded into the tile kernel: the normalization as float32,
  // one chunk's mask bits and the passes' row sums.
  dio::DevBuf<float> norm32;
  dio::DevBuf<uint8_t> dyn_m;
  dio::DevBuf<double> dyn_r;
  dio::DevBuf<double> out_i, out_s, out_d;
  ~dio_plan() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct dio_graph {
  dio_plan* plan = nullptr;
  dio::TileScratch sc;  // owned: the plan's scratch may be reallocated later
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~dio_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

namespace dio {
// Enqueues the integration of device frames on stream s (no synchronisation).
void run_device(dio_plan* p, TileScratch& sc, const void* frames, int dtype, uint64_t n_frames,
                uint32_t flags, double* I, double* S, double* SD, hipStream_t s,
                dio::LaunchEvents* launches = nullptr);
// Fixed scratch of dio_integrate_device_noalloc (noalloc.hip): its device
// address is taken once per device, with that device current (dio_open);
// noalloc_scratch returns it and the usable bytes.
void noalloc_init(int device);
void* noalloc_scratch(int device, size_t* bytes);
size_t noalloc_capacity();
// Device bytes a plan holds (tables, corrections, tile layout, scratch and
// staging buffers it keeps).
size_t plan_device_bytes(const dio_plan* p);
// Plan construction behind dio_plan_create (validates, builds on the device).
dio_plan* make_plan(dio_ctx* ctx, const dio_poni& poni, const dio_plan_desc& desc);
}  // namespace dio
