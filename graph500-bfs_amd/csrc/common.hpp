// common.hpp -- error plumbing and wave64 helpers shared by every kernel file.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/g500bfs.h"

namespace g500 {

// Internal errors travel as exceptions and are turned into status codes plus
// a thread-local message at the C boundary (api.hip).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const char* fmt, ...);

#define G500_CHECK_HIP(expr)                                                  \
  do {                                                                        \
    hipError_t e__ = (expr);                                                  \
    if (e__ != hipSuccess)                                                    \
      ::g500::fail(G500_ERR_HIP, "%s: %s (%s:%d)", #expr,                     \
                   hipGetErrorString(e__), __FILE__, __LINE__);               \
  } while (0)

#define G500_CHECK_LAUNCH() G500_CHECK_HIP(hipGetLastError())

constexpr int kWave = 64;

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) {
  return (a + b - 1) / b;
}

__device__ inline int lane_id() { return (int)__lane_id(); }
// mask of lanes strictly below this one
__device__ inline uint64_t lanemask_lt() {
  const int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Device allocation owned by a host object.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess)
      fail(G500_ERR_NOMEM, "hipMalloc(%zu bytes): %s", count * sizeof(T),
           hipGetErrorString(e));
    p = static_cast<T*>(q);
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  T* get() const { return p; }
};

}  // namespace g500
