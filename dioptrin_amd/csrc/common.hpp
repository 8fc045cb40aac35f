// Shared host-side plumbing: errors, device buffers, launch helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>

#include "../../include/dioptrin_hip.h"

namespace dio {

// Typed error carried inside the library and turned into a status code and
// thread-local message at the C boundary (mirrors CudaError, cuda.rs:218-237).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

}  // namespace dio

// Sets the thread-local text returned by dio_last_error().
void dio_set_last_error(const char* msg);

namespace dio {

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(DIO_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define DIO_HIP(call) ::dio::hip_check((call), #call)

// Owning device allocation.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) DIO_HIP(hipMalloc(&p, count * sizeof(T)));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

inline unsigned blocks_for(size_t n, unsigned threads) {
  return static_cast<unsigned>((n + threads - 1) / threads);
}

inline void require(bool ok, int code, const char* msg) {
  if (!ok) fail(code, msg);
}

// Bytes per element of a frame dtype (DIO_U16 .. DIO_F64T).
inline size_t dtype_bytes(int dtype) {
  switch (dtype) {
    case DIO_U16: return 2;
    case DIO_U32: case DIO_I32: case DIO_F32T: return 4;
    case DIO_F64T: return 8;
    default: fail(DIO_ERR_INVALID, "unknown frame dtype");
  }
}

// Runs fn at the C boundary: exceptions become a status code and the
// thread-local text of dio_last_error(), which success clears.
template <class Fn>
int guarded(Fn&& fn) {
  try {
    fn();
    dio_set_last_error("");
    return DIO_OK;
  } catch (const Error& e) {
    dio_set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    dio_set_last_error("host out of memory");
    return DIO_ERR_DEVICE;
  } catch (const std::exception& e) {
    dio_set_last_error(e.what());
    return DIO_ERR_DEVICE;
  }
}

}  // namespace dio
