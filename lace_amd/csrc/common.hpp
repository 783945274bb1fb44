// common.hpp -- runtime plumbing shared by every translation unit of
// liblace_mi355x.so: error reporting, the library stream, device buffers,
// host allocation in the reference's convention, and kernel timing.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lace.h"

namespace lace {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
const char* get_error();

#define LACE_HIP(call)                                                   \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) {                                              \
      ::lace::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call,     \
                        hipGetErrorString(e_));                          \
      return DEV_ERR_HIP;                                                \
    }                                                                    \
  } while (0)

#define LACE_TRY(call)            \
  do {                            \
    int r_ = (call);              \
    if (r_ != DEV_SUCCESS) return r_; \
  } while (0)

#define LACE_REQUIRE(cond, code, ...)   \
  do {                                  \
    if (!(cond)) {                      \
      ::lace::set_error(__VA_ARGS__);   \
      return (code);                    \
    }                                   \
  } while (0)

// ---------------------------------------------------------------- runtime
// The library runs everything on one non-blocking stream of default
// priority, created on first use on the current device.
int ensure_init();
hipStream_t stream();
int device_id();
int num_cus();
// Device address of the allocation-free entry point's scratch array on
// this GPU, looked up once per device (noalloc.hip).
int noalloc_setup(int device);
// Kernel-shape option set through lace_set_option (dflt when unset).
int get_option(const char* name, int dflt);

// ---------------------------------------------------------------- memory
// Host arrays in the reference convention (include/dense_types.h:430-450):
// 64-byte aligned, zero-filled, padded to a whole 64-byte block, released
// with free().
template <typename T>
int host_calloc(T** p, size_t n) {
  size_t bytes = (n * sizeof(T) / DEV_ALIGN + 1) * DEV_ALIGN;
  void* q = nullptr;
  if (posix_memalign(&q, DEV_ALIGN, bytes) != 0) {
    set_error("host allocation of %zu bytes failed", bytes);
    return DEV_ERR_HOST_ALLOC;
  }
  memset(q, 0, bytes);
  *p = static_cast<T*>(q);
  return DEV_SUCCESS;
}

// Device buffer owned by a handle (freed in the destructor).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int alloc(size_t nbytes) {
    release();
    if (nbytes == 0) nbytes = 256;
    hipError_t e = hipMalloc(&p, nbytes);
    if (e != hipSuccess) {
      set_error("hipMalloc(%zu) failed: %s", nbytes, hipGetErrorString(e));
      p = nullptr;
      return DEV_ERR_DEVICE_ALLOC;
    }
    bytes = nbytes;
    return DEV_SUCCESS;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

int upload(DevBuf& b, const void* host, size_t bytes);
int download(void* host, const DevBuf& b, size_t bytes);

// ---------------------------------------------------------------- timing
// Named kernel classes timed with HIP events on the library stream when
// timing is enabled; the averages feed the roofline report.
// A region may carry its algorithmic work (flops and HBM bytes of that
// launch), summed per region for launches whose shape varies.
// A region made with kernel_events = true times one kernel: launched
// through launch(), that kernel's own start and end set the two events
// (hipExtLaunchKernel), so the duration excludes the dispatch latency that
// marker events around a short kernel would add.
struct TimedRegion {
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  bool on = false, ext = false, used = false;
  double flops = 0.0, bytes = 0.0;
  explicit TimedRegion(const char* n);
  TimedRegion(const char* n, double flops_, double bytes_, bool kernel_events = false);
  ~TimedRegion();
  template <typename F, typename... Args>
  void launch(F kernel, dim3 grid, dim3 block, Args... args);
};
bool timing_on();

template <typename F, typename... Args>
void TimedRegion::launch(F kernel, dim3 grid, dim3 block, Args... args) {
  const bool timed = on && ext && !used;
  hipExtLaunchKernelGGL(kernel, grid, block, 0, stream(), timed ? a : nullptr, timed ? b : nullptr, 0,
                        args...);
  used = used || timed;
}

// ---------------------------------------------------------------- launch
constexpr int kWave = 64;

inline int ceil_div(long a, long b) { return static_cast<int>((a + b - 1) / b); }

// Reorders block indices so that consecutive logical blocks share an XCD:
// the hardware deals workgroups round-robin over the 8 XCDs, so logical
// block L runs on XCD (L % 8); remapping gives each XCD one contiguous
// range of the row space and keeps neighbouring rows in one L2.
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
  const int nx = 8;
  int per = nblocks / nx;
  int rem = nblocks - per * nx;
  int xcd = b % nx;
  int idx = b / nx;
  // XCDs [0, rem) own per+1 blocks, the rest own per.
  int start = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
  return start + idx;
}

}  // namespace lace
