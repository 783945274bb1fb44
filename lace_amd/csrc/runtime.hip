// runtime.hip -- library initialisation, error text, device memory entry
// points and the kernel-timing registry.
#include <cstdarg>
#include <algorithm>
#include <map>
#include <mutex>

#include "common.hpp"

namespace lace {

static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
const char* get_error() { return g_err; }

static std::mutex g_init_mu;
static hipStream_t g_stream = nullptr;
static int g_device = -1;
static int g_cus = 0;

static int init_device(int dev) {
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_stream && g_device == dev) return DEV_SUCCESS;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    set_error("no HIP device available (%s)", hipGetErrorString(e));
    return DEV_ERR_NOT_INITIALIZED;
  }
  if (dev < 0 || dev >= count) {
    set_error("device %d out of range (%d devices)", dev, count);
    return DEV_ERR_ILLEGAL_VALUE;
  }
  LACE_HIP(hipSetDevice(dev));
  hipDeviceProp_t prop;
  LACE_HIP(hipGetDeviceProperties(&prop, dev));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error("device %d is %s; this library is built for gfx950 only", dev,
              prop.gcnArchName);
    return DEV_ERR_NOT_SUPPORTED;
  }
  if (g_stream) {
    (void)hipStreamDestroy(g_stream);
    g_stream = nullptr;
  }
  LACE_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  g_device = dev;
  g_cus = prop.multiProcessorCount;
  // this GPU's allocation-free scratch address (noalloc.hip)
  LACE_TRY(noalloc_setup(dev));
  return DEV_SUCCESS;
}

int ensure_init() {
  if (g_stream) {
    LACE_HIP(hipSetDevice(g_device));
    return DEV_SUCCESS;
  }
  return init_device(0);
}
hipStream_t stream() { return g_stream; }
int device_id() { return g_device; }
int num_cus() { return g_cus; }

int upload(DevBuf& b, const void* host, size_t bytes) {
  LACE_TRY(b.alloc(bytes));
  if (bytes) {
    // an empty host array (a matrix without off-diagonal entries, say)
    // still gets its placeholder allocation, zeroed
    if (host) LACE_HIP(hipMemcpyAsync(b.p, host, bytes, hipMemcpyHostToDevice, stream()));
    else LACE_HIP(hipMemsetAsync(b.p, 0, bytes, stream()));
    LACE_HIP(hipStreamSynchronize(stream()));
  }
  return DEV_SUCCESS;
}

int download(void* host, const DevBuf& b, size_t bytes) {
  if (bytes) {
    LACE_HIP(hipMemcpyAsync(host, b.p, bytes, hipMemcpyDeviceToHost, stream()));
    LACE_HIP(hipStreamSynchronize(stream()));
  }
  return DEV_SUCCESS;
}

// ---------------------------------------------------------------- timing
struct TimingEntry {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  double total_ms = 0.0;
  long count = 0;  // timed entries
  long calls = 0;  // all entries while timing was on (timed or skipped by sampling)
  double flops = 0.0, bytes = 0.0;
};
static std::mutex g_tmu;
static std::map<std::string, TimingEntry> g_timing;
static bool g_timing_on = false;
static std::string g_timing_only;  // empty: time every region
static int g_timing_sample = 1;    // time every k-th entry of each region
// The first entries of each region since the last reset are always timed,
// so a region entered only a few times still gets several samples.
constexpr long kTimingFirst = 4;
static bool sampled(long call) { return call < kTimingFirst || call % g_timing_sample == 0; }
// Events are reused: creating two per timed region costs several
// microseconds of host time per launch, which a launch-bound solve loop
// would otherwise carry into its wall time.
static std::vector<hipEvent_t> g_event_pool;

static hipEvent_t pooled_event() {  // caller holds g_tmu
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

bool timing_on() { return g_timing_on; }

// g_timing_only: a comma-separated list of region names (empty: all).
static bool timing_selected(const char* n) {
  if (g_timing_only.empty()) return true;
  const size_t len = strlen(n);
  size_t pos = 0;
  while (pos <= g_timing_only.size()) {
    size_t end = g_timing_only.find(',', pos);
    if (end == std::string::npos) end = g_timing_only.size();
    if (end - pos == len && g_timing_only.compare(pos, len, n) == 0) return true;
    pos = end + 1;
  }
  return false;
}

TimedRegion::TimedRegion(const char* n) : name(n) {
  if (!g_timing_on) return;
  std::lock_guard<std::mutex> lk(g_tmu);
  if (!timing_selected(n)) return;
  TimingEntry& e = g_timing[name];
  if (!sampled(e.calls++)) return;
  a = pooled_event();
  b = pooled_event();
  if (!a || !b) {
    if (a) g_event_pool.push_back(a);
    if (b) g_event_pool.push_back(b);
    a = b = nullptr;
    return;
  }
  on = hipEventRecord(a, stream()) == hipSuccess;
  if (!on) {
    g_event_pool.push_back(a);
    g_event_pool.push_back(b);
  }
}

TimedRegion::TimedRegion(const char* n, double flops_, double bytes_, bool kernel_events) : name(n) {
  flops = flops_;
  bytes = bytes_;
  if (!g_timing_on) return;
  std::lock_guard<std::mutex> lk(g_tmu);
  if (!timing_selected(n)) return;
  TimingEntry& e = g_timing[name];
  if (!sampled(e.calls++)) return;
  a = pooled_event();
  b = pooled_event();
  if (!a || !b) {
    if (a) g_event_pool.push_back(a);
    if (b) g_event_pool.push_back(b);
    a = b = nullptr;
    return;
  }
  if (kernel_events) {
    // recorded by the kernel launched through launch()
    ext = true;
    on = true;
    return;
  }
  on = hipEventRecord(a, stream()) == hipSuccess;
  if (!on) {
    g_event_pool.push_back(a);
    g_event_pool.push_back(b);
  }
}

TimedRegion::~TimedRegion() {
  if (!on) return;
  if (ext && !used) {
    // no kernel took the events: nothing was timed
    std::lock_guard<std::mutex> lk(g_tmu);
    g_event_pool.push_back(a);
    g_event_pool.push_back(b);
    return;
  }
  if (!ext) (void)hipEventRecord(b, stream());
  std::lock_guard<std::mutex> lk(g_tmu);
  TimingEntry& e = g_timing[name];
  e.pending.emplace_back(a, b);
  e.flops += flops;
  e.bytes += bytes;
}

static void drain_timing() {
  std::lock_guard<std::mutex> lk(g_tmu);
  for (auto& kv : g_timing) {
    for (auto& ev : kv.second.pending) {
      (void)hipEventSynchronize(ev.second);
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev.first, ev.second) == hipSuccess) {
        kv.second.total_ms += ms;
        kv.second.count++;
      }
      g_event_pool.push_back(ev.first);
      g_event_pool.push_back(ev.second);
    }
    kv.second.pending.clear();
  }
}

// Plain streaming copy, 16 bytes per lane per iteration; the reference
// kernel for calibrating HBM traffic counters (bytes known exactly).
__global__ __launch_bounds__(256) void stream_copy_kernel(size_t n16, const uint4* __restrict__ src,
                                                         uint4* __restrict__ dst) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (size_t)gridDim.x * 256)
    dst[i] = src[i];
}

// ---------------------------------------------------------------- options
// Kernel-shape options (benchmarking and tuning); read by the launchers.
static std::mutex g_omu;
static std::map<std::string, int> g_options;

int get_option(const char* name, int dflt) {
  std::lock_guard<std::mutex> lk(g_omu);
  auto it = g_options.find(name);
  return it == g_options.end() ? dflt : it->second;
}

}  // namespace lace

using namespace lace;

extern "C" {

int lace_init(int device) { return init_device(device); }

const char* lace_last_error(void) { return get_error(); }

int lace_device_synchronize(void) {
  LACE_TRY(ensure_init());
  LACE_HIP(hipStreamSynchronize(stream()));
  return DEV_SUCCESS;
}

int lace_dmalloc(void** ptr, size_t bytes) {
  LACE_TRY(ensure_init());
  LACE_REQUIRE(ptr != nullptr, DEV_ERR_INVALID_PTR, "lace_dmalloc: null out pointer");
  LACE_HIP(hipMalloc(ptr, bytes ? bytes : 256));
  return DEV_SUCCESS;
}

int lace_dfree(void* ptr) {
  LACE_TRY(ensure_init());
  if (ptr) LACE_HIP(hipFree(ptr));
  return DEV_SUCCESS;
}

int lace_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  LACE_TRY(ensure_init());
  LACE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
  LACE_HIP(hipStreamSynchronize(stream()));
  return DEV_SUCCESS;
}

int lace_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  LACE_TRY(ensure_init());
  LACE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
  LACE_HIP(hipStreamSynchronize(stream()));
  return DEV_SUCCESS;
}

int lace_stream_copy(void* dst, const void* src, size_t bytes) {
  LACE_TRY(ensure_init());
  LACE_REQUIRE(bytes % 16 == 0, DEV_ERR_ILLEGAL_VALUE, "lace_stream_copy: bytes must be a multiple of 16");
  const size_t n16 = bytes / 16;
  const int grid = (int)std::min<size_t>((n16 + 255) / 256, (size_t)num_cus() * 16);
  // timed by events its own launch sets, as the hot kernels are
  TimedRegion tr("stream_copy", 0.0, 2.0 * (double)bytes, true);
  tr.launch(stream_copy_kernel, dim3(grid > 0 ? grid : 1), dim3(256), n16, static_cast<const uint4*>(src),
            static_cast<uint4*>(dst));
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

int lace_timing_enable(int on) {
  LACE_TRY(ensure_init());
  g_timing_on = on != 0;
  return DEV_SUCCESS;
}

int lace_timing_only(const char* name) {
  std::lock_guard<std::mutex> lk(g_tmu);
  g_timing_only = name ? name : "";
  return DEV_SUCCESS;
}

int lace_timing_reset(void) {
  drain_timing();
  std::lock_guard<std::mutex> lk(g_tmu);
  g_timing.clear();
  return DEV_SUCCESS;
}

int lace_timing_get(const char* name, double* total_ms, long* count) {
  drain_timing();
  std::lock_guard<std::mutex> lk(g_tmu);
  auto it = g_timing.find(name);
  if (it == g_timing.end()) {
    *total_ms = 0.0;
    *count = 0;
    return DEV_ERR_NOT_FOUND;
  }
  *total_ms = it->second.total_ms;
  *count = it->second.count;
  return DEV_SUCCESS;
}

int lace_timing_get_calls(const char* name, long* calls) {
  std::lock_guard<std::mutex> lk(g_tmu);
  auto it = g_timing.find(name);
  *calls = it == g_timing.end() ? 0 : it->second.calls;
  return it == g_timing.end() ? DEV_ERR_NOT_FOUND : DEV_SUCCESS;
}

int lace_timing_get_work(const char* name, double* flops, double* bytes) {
  drain_timing();
  std::lock_guard<std::mutex> lk(g_tmu);
  auto it = g_timing.find(name);
  *flops = it == g_timing.end() ? 0.0 : it->second.flops;
  *bytes = it == g_timing.end() ? 0.0 : it->second.bytes;
  return it == g_timing.end() ? DEV_ERR_NOT_FOUND : DEV_SUCCESS;
}

}  // extern "C"

extern "C" int lace_set_option(const char* name, int value) {
  LACE_REQUIRE(name != nullptr, DEV_ERR_INVALID_PTR, "lace_set_option: null name");
  const std::string key(name);
  if (key == "stream_rows_per_thread") {
    LACE_REQUIRE(value == 1 || value == 2 || value == 4, DEV_ERR_ILLEGAL_VALUE,
                 "stream_rows_per_thread must be 1, 2 or 4 (have %d)", value);
  } else if (key == "jacobi_ell" || key == "spmv_ell" || key == "jacobi_dia" || key == "spmv_dia" ||
             key == "jacobi_fuse" || key == "jacobi_dia_vec" || key == "jacobi_pipe" ||
             key == "matlab_log" || key == "jacobi_dia16" || key == "inner_ortho" || key == "trsv_exact_tiles" ||
             key == "mrhs_plane_order" || key == "jacobi_pair") {
    LACE_REQUIRE(value == 0 || value == 1, DEV_ERR_ILLEGAL_VALUE, "%s must be 0 or 1 (have %d)", name, value);
  } else if (key == "trsv_exact_wgs_per_cu") {
    LACE_REQUIRE(value >= 0 && value <= 16, DEV_ERR_ILLEGAL_VALUE, "%s must be 0..16 (have %d)", name, value);
  } else if (key == "trsv_exact") {
    LACE_REQUIRE(value == 0 || value == 1, DEV_ERR_ILLEGAL_VALUE, "%s must be 0 or 1 (have %d)", name, value);
  } else if (key == "trsv_exact_rows") {
    LACE_REQUIRE(value == 256 || value == 512, DEV_ERR_ILLEGAL_VALUE, "%s must be 256 or 512 (have %d)", name,
                 value);
  } else if (key == "trsv_exact_idle") {
    LACE_REQUIRE(value >= 0 && value <= 16, DEV_ERR_ILLEGAL_VALUE, "%s must be 0..16 (have %d)", name, value);
  } else if (key == "trsv_exact_poll_sleep") {
    LACE_REQUIRE(value >= 0 && value <= 64, DEV_ERR_ILLEGAL_VALUE, "%s must be 0..64 (have %d)", name, value);
  } else if (key == "timing_sample") {
    LACE_REQUIRE(value >= 1, DEV_ERR_ILLEGAL_VALUE, "timing_sample must be >= 1 (have %d)", value);
    std::lock_guard<std::mutex> lk(g_tmu);
    g_timing_sample = value;
  } else {
    LACE_REQUIRE(false, DEV_ERR_ILLEGAL_VALUE, "unknown option %s", name);
  }
  std::lock_guard<std::mutex> lk(g_omu);
  g_options[key] = value;
  return DEV_SUCCESS;
}
