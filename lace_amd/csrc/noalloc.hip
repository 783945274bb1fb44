// noalloc.hip -- the allocation-free variant of data_parcsrtrsv on device
// arrays, for callers that cannot allocate device memory (code under HIP
// graph capture, serving loops, harnesses that forbid hipMalloc).
//
// Its scratch is one file-scope device array per GPU, sized for
// kNoallocMaxRows rows: the caller's initial y (for the reference's step
// rule), per-block partial sums of the step, the persistent solve's block
// ticket and error word.  The host reaches it only through the address
// hipGetSymbolAddress returns on each device (looked up once per device at
// library initialisation); the array's own name is not a device address in
// host code.  Every call on a GPU shares that GPU's array and finds the
// previous call's contents in it, so each call clears the words that must
// start at zero itself, with a kernel; calls that overlap on one GPU race.
#include <cmath>

#include "common.hpp"
#include "device.hpp"
#include "parilu.hpp"

namespace lace {

// The scratch array (32 MiB + 8 KiB + 256 B on each GPU the library runs on,
// held until the process ends).
static __device__ __align__(256) double scratch[(size_t)LACE_NOALLOC_MAX_ROWS + 1024 + 32];

namespace {

constexpr int kNoallocMaxRows = LACE_NOALLOC_MAX_ROWS;  // 32 MiB of y0
constexpr int kStepBlocks = 1024;                       // step partial sums
constexpr int kCtlWords = 32;                           // ticket and error word
constexpr size_t kScratchDoubles = (size_t)kNoallocMaxRows + kStepBlocks + kCtlWords;

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {nullptr};

struct ScratchView {
  double* y0;        // kNoallocMaxRows
  double* partial;   // kStepBlocks
  unsigned* ticket;  // control words
  int* err;
};

ScratchView view(double* base) {
  ScratchView v;
  v.y0 = base;
  v.partial = base + kNoallocMaxRows;
  unsigned* ctl = reinterpret_cast<unsigned*>(base + kNoallocMaxRows + kStepBlocks);
  v.ticket = ctl;
  v.err = reinterpret_cast<int*>(ctl + 1);
  return v;
}

// y0 <- y, y <- the unsolved-row sentinel, ticket <- 0, err <- 0.
__global__ void noalloc_prep_kernel(int n, double* __restrict__ y, double* __restrict__ y0,
                                    unsigned* __restrict__ ticket, int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    y0[i] = y[i];
    y[i] = __builtin_bit_cast(double, Sentinel<double>::bits);
  }
  if (i == 0) {
    *ticket = 0u;
    *err = 0;
  }
}

// Stage 1 of the step: per-block sums of (y0_i - y_i)^2 over a fixed range
// of rows per block, fixed reduction tree.
__global__ __launch_bounds__(256) void noalloc_step_partial_kernel(int n, int rows_per_block,
                                                                   const double* __restrict__ y0,
                                                                   const double* __restrict__ y,
                                                                   double* __restrict__ partial) {
  __shared__ double sm[4];
  const int r0 = blockIdx.x * rows_per_block;
  const int r1 = min(n, r0 + rows_per_block);
  double s = 0.0;
  for (int i = r0 + (int)threadIdx.x; i < r1; i += 256) {
    const double d = y0[i] - y[i];
    s = fma(d, d, s);
  }
  const double tot = block_sum<256>(s, sm);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Stage 2: the step in block order, the sweep count of the reference loop
// (1 when step <= tol or not finite, else 2), -1 if the solve gave up.
__global__ __launch_bounds__(256) void noalloc_step_final_kernel(int nb, const double* __restrict__ partial,
                                                                 double tol, const int* __restrict__ err,
                                                                 int* __restrict__ iter) {
  __shared__ double sm[4];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += partial[b];
  const double step = block_sum<256>(s, sm);
  if (threadIdx.x == 0) *iter = *err != 0 ? -1 : ((step > tol && isfinite(step)) ? 2 : 1);
}

// tol >= 1e8: the loop runs no sweep.
__global__ void noalloc_iter_zero_kernel(int* __restrict__ iter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *iter = 0;
}

}  // namespace

// Looks up this GPU's scratch address once; called by the library's
// device initialisation with that device current.
int noalloc_setup(int device) {
  LACE_REQUIRE(device >= 0 && device < kMaxDevices, DEV_ERR_ILLEGAL_VALUE, "device %d out of range", device);
  if (g_scratch[device]) return DEV_SUCCESS;
  void* p = nullptr;
  LACE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  static_assert(sizeof(scratch) == kScratchDoubles * sizeof(double), "scratch layout");
  LACE_REQUIRE(p != nullptr, DEV_ERR_HIP, "no device address for the allocation-free scratch");
  g_scratch[device] = static_cast<double*>(p);
  return DEV_SUCCESS;
}

}  // namespace lace

using namespace lace;

extern "C" int lace_parcsrtrsv_noalloc(const data_uplo_t uplo, const data_storage_t storage,
                                       const data_diagorder_t diag, const int num_rows, const dataType* Aval,
                                       const int* row, const int* col, const dataType* rhsval, dataType* yval,
                                       const dataType tol, int* iter) {
  LACE_TRY(ensure_init());
  const int dev = device_id();
  LACE_REQUIRE(dev >= 0 && dev < kMaxDevices && g_scratch[dev], DEV_ERR_NOT_INITIALIZED,
               "lace_parcsrtrsv_noalloc: scratch of device %d not set up", dev);
  bool upper;
  if (storage == Magma_CSRL && uplo == MagmaLower && diag != Magma_NODIAG) upper = false;
  else if (storage == Magma_CSRU && uplo == MagmaUpper && diag != Magma_NODIAG) upper = true;
  else {
    set_error("lace_parcsrtrsv_noalloc: storage %d / uplo %d must be CSRL+Lower or CSRU+Upper", (int)storage,
              (int)uplo);
    return -1;
  }
  LACE_REQUIRE(num_rows >= 0 && num_rows <= kNoallocMaxRows, DEV_ERR_ILLEGAL_VALUE,
               "lace_parcsrtrsv_noalloc: %d rows; the allocation-free solve takes at most %d", num_rows,
               kNoallocMaxRows);
  LACE_REQUIRE(iter, DEV_ERR_INVALID_PTR, "lace_parcsrtrsv_noalloc: null iter");
  if (!(1.e8 > tol)) {
    noalloc_iter_zero_kernel<<<1, 64, 0, stream()>>>(iter);
    LACE_HIP(hipGetLastError());
    return 0;
  }
  const int n = num_rows;
  LACE_REQUIRE(n == 0 || (Aval && row && col && rhsval && yval), DEV_ERR_INVALID_PTR,
               "lace_parcsrtrsv_noalloc: null array");
  LACE_REQUIRE((const void*)rhsval != (const void*)yval, DEV_ERR_ILLEGAL_VALUE,
               "lace_parcsrtrsv_noalloc: rhs and y must be distinct");
  const ScratchView sv = view(g_scratch[dev]);
  if (n == 0) {
    // an empty sweep has step 0 <= tol
    noalloc_prep_kernel<<<1, 64, 0, stream()>>>(0, yval, sv.y0, sv.ticket, sv.err);
    noalloc_step_final_kernel<<<1, 256, 0, stream()>>>(0, sv.partial, tol, sv.err, iter);
    LACE_HIP(hipGetLastError());
    return 0;
  }
  noalloc_prep_kernel<<<ceil_div(n, 256), 256, 0, stream()>>>(n, yval, sv.y0, sv.ticket, sv.err);
  LACE_HIP(hipGetLastError());
  LACE_TRY(tri_solve_persist_raw<double>(n, upper, row, col, Aval, rhsval, yval, sv.ticket, sv.err, false));
  const int rpb = ceil_div(n, kStepBlocks);
  const int nb = ceil_div(n, rpb);
  noalloc_step_partial_kernel<<<nb, 256, 0, stream()>>>(n, rpb, sv.y0, yval, sv.partial);
  noalloc_step_final_kernel<<<1, 256, 0, stream()>>>(nb, sv.partial, tol, sv.err, iter);
  LACE_HIP(hipGetLastError());
  return 0;
}
