// Fixed device scratch of dio_integrate_device_noalloc: the tile kernel's
// segment partials of that entry point live in this array instead of memory
// the library allocates, so the call itself allocates nothing (for callers
// under HIP graph capture or in loops that cannot allocate).
//
// The array is a file-scope device variable: each device the library runs
// on holds its own copy, at its own address, until the process ends. In host
// code the array's name is only a stand-in, so its device address is taken
// once per device (dio_open) with hipGetSymbolAddress, with that device
// current, and only that address is handed to kernels. Every call on a
// device shares that device's array and leaves the last call's partials in
// it; calls that overlap on one device race (the caller serialises them).
#include <cstdlib>
#include <mutex>

#include "plan.hpp"

namespace {

constexpr size_t kScratchBytes = 64ull << 20;  // 64 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

void check_ordinal(int device) {
  if (device < 0 || device >= kMaxDevices)
    dio::fail(DIO_ERR_UNAVAILABLE, "device ordinal out of range");
}

}  // namespace

namespace dio {

void noalloc_init(int device) {
  check_ordinal(device);
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return;
  void* ptr = nullptr;
  DIO_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) fail(DIO_ERR_DEVICE, "no device address for the fixed scratch");
  g_ptr[device] = ptr;
}

void* noalloc_scratch(int device, size_t* bytes) {
  check_ordinal(device);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) fail(DIO_ERR_INVALID, "device not opened with dio_open");
  size_t cap = kScratchBytes;
  // DIOPTRIN_NOALLOC_LIMIT (bytes) lowers the usable part, to test the
  // refusal of plans whose partials do not fit.
  if (const char* e = std::getenv("DIOPTRIN_NOALLOC_LIMIT")) {
    const long long v = std::atoll(e);
    if (v > 0 && (size_t)v < cap) cap = (size_t)v;
  }
  *bytes = cap;
  return g_ptr[device];
}

size_t noalloc_capacity() { return kScratchBytes; }

}  // namespace dio
