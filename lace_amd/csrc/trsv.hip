// trsv.hip -- sparse triangular solves on the GPU.
//
// Reference: src/trisolve.cpp:20-119 (data_forward_solve/backward_solve),
// :256-351 (data_parcsrtrsv), and the exact mkl_dcsrtrsv solves of
// src/fgmres.cpp:216-228.  The reference has no level scheduling; its
// ParCSRTRSV iterates in-place OpenMP sweeps.  Three device variants here:
//   - level-scheduled, one launch per level (exact);
//   - sync-free, one launch: rows run in level order, each lane waits on
//     the values of the rows it depends on (exact);
//   - synchronous Jacobi sweeps (the update formula of trisolve.cpp:281-290
//     with all reads from the previous sweep).
// Every variant forms a row's off-diagonal sum left to right and then
// divides (r_i - sum) by the diagonal, as the reference loop does, with
// contraction disabled, so exact solves equal forward/backward
// substitution on the CPU bit for bit.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "parilu.hpp"

namespace lace {

// ---------------------------------------------------------------- plans
int TriPlan::build(int n_, const int* row_, const int* col_, bool upper_) {
  n = n_;
  nnz = row_[n_];
  upper = upper_;
  std::vector<int> perm_h(n), lptr_h(n + 1);
  nlev = lace_level_sets(n, row_, col_, upper ? 1 : 0, perm_h.data(), lptr_h.data());
  h_lptr.assign(lptr_h.begin(), lptr_h.begin() + nlev + 1);
  std::vector<int> sched_h;
  sched_h.reserve(n + 64 * (size_t)nlev);
  for (int l = 0; l < nlev; l++) {
    for (int t = h_lptr[l]; t < h_lptr[l + 1]; t++) sched_h.push_back(perm_h[t]);
    while (sched_h.size() % 64) sched_h.push_back(-1);
  }
  nslots = (int)sched_h.size();
  LACE_TRY(upload(row, row_, sizeof(int) * (n + 1)));
  LACE_TRY(upload(col, col_, sizeof(int) * std::max(nnz, 1)));
  LACE_TRY(upload(perm, perm_h.data(), sizeof(int) * std::max(n, 1)));
  LACE_TRY(upload(lptr, h_lptr.data(), sizeof(int) * (nlev + 1)));
  LACE_TRY(upload(sched, sched_h.data(), sizeof(int) * std::max(nslots, 1)));
  // off-diagonal form: every entry but the row's diagonal slot (first for
  // upper, last for lower), in row order
  std::vector<int> sr(n + 1, 0), sc, sm, dp(n);
  sc.reserve(std::max(0, nnz - n));
  sm.reserve(std::max(0, nnz - n));
  for (int i = 0; i < n; i++) {
    const int kd = upper ? row_[i] : row_[i + 1] - 1;
    dp[i] = kd;
    for (int k = row_[i]; k < row_[i + 1]; k++)
      if (k != kd) {
        sc.push_back(col_[k]);
        sm.push_back(k);
      }
    sr[i + 1] = (int)sc.size();
  }
  snnz = (int)sc.size();
  LACE_TRY(upload(srow, sr.data(), sizeof(int) * (n + 1)));
  LACE_TRY(upload(scol, sc.data(), sizeof(int) * std::max(snnz, 1)));
  LACE_TRY(upload(smap, sm.data(), sizeof(int) * std::max(snnz, 1)));
  LACE_TRY(upload(dpos, dp.data(), sizeof(int) * std::max(n, 1)));
  // ELLPACK form for short rows
  int wmax = 0;
  for (int i = 0; i < n; i++) wmax = std::max(wmax, sr[i + 1] - sr[i]);
  ew = 0;
  if (n > 0 && wmax >= 1 && wmax <= 8 && (double)wmax * n <= 1.25 * snnz) {
    ew = wmax;
    std::vector<int> ec((size_t)ew * n, -1), em((size_t)ew * n, -1);
    for (int i = 0; i < n; i++)
      for (int k = sr[i]; k < sr[i + 1]; k++) {
        ec[(size_t)(k - sr[i]) * n + i] = sc[k];
        em[(size_t)(k - sr[i]) * n + i] = sm[k];
      }
    LACE_TRY(upload(ecol, ec.data(), sizeof(int) * ec.size()));
    LACE_TRY(upload(emap, em.data(), sizeof(int) * em.size()));
  }
  // diagonal form of the same entries
  LACE_TRY(dia.build(n, row_, col_, [&](int i, int k) { return k != dp[i]; }));
  strict_src = nullptr;
  return DEV_SUCCESS;
}

template <typename T>
__global__ void ell_gather_kernel(size_t m, const int* __restrict__ emap, const T* __restrict__ val,
                                  T* __restrict__ eval) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < m) eval[e] = emap[e] >= 0 ? val[emap[e]] : T(0);
}

template <typename T>
__global__ void strict_gather_kernel(int snnz, int n, const int* __restrict__ smap, const int* __restrict__ dpos,
                                     const T* __restrict__ val, T* __restrict__ sval, T* __restrict__ dval) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < snnz) sval[e] = val[smap[e]];
  if (e < n) dval[e] = val[dpos[e]];
}

int TriPlan::refresh_strict(const void* val, size_t es) {
  if (sval.bytes < es * std::max(snnz, 1)) LACE_TRY(sval.alloc(es * std::max(snnz, 1)));
  if (dval.bytes < es * std::max(n, 1)) LACE_TRY(dval.alloc(es * std::max(n, 1)));
  const int m = std::max(snnz, n);
  if (m > 0) {
    if (es == 8)
      strict_gather_kernel<double><<<ceil_div(m, 256), 256, 0, stream()>>>(
          snnz, n, smap.as<int>(), dpos.as<int>(), static_cast<const double*>(val), sval.as<double>(),
          dval.as<double>());
    else
      strict_gather_kernel<float><<<ceil_div(m, 256), 256, 0, stream()>>>(
          snnz, n, smap.as<int>(), dpos.as<int>(), static_cast<const float*>(val), sval.as<float>(),
          dval.as<float>());
    LACE_HIP(hipGetLastError());
  }
  if (ew > 0) {
    const size_t m2 = (size_t)ew * n;
    if (eval.bytes < es * m2) LACE_TRY(eval.alloc(es * m2));
    if (es == 8)
      ell_gather_kernel<double><<<ceil_div((long)m2, 256), 256, 0, stream()>>>(
          m2, emap.as<int>(), static_cast<const double*>(val), eval.as<double>());
    else
      ell_gather_kernel<float><<<ceil_div((long)m2, 256), 256, 0, stream()>>>(
          m2, emap.as<int>(), static_cast<const float*>(val), eval.as<float>());
    LACE_HIP(hipGetLastError());
  }
  LACE_TRY(dia.gather(val, es));
  strict_src = val;
  return DEV_SUCCESS;
}

// ---------------------------------------------------------------- kernels
template <typename T, bool UPPER>
__device__ __forceinline__ void row_range(const int* __restrict__ row, int i, int& k0, int& k1,
                                          int& kd) {
  if (UPPER) {
    kd = row[i];
    k0 = kd + 1;
    k1 = row[i + 1];
  } else {
    kd = row[i + 1] - 1;
    k0 = row[i];
    k1 = kd;
  }
}

template <typename T, bool UPPER>
__global__ __launch_bounds__(256) void tri_level_kernel(int begin, int count,
                                                        const int* __restrict__ perm,
                                                        const int* __restrict__ row,
                                                        const int* __restrict__ col,
                                                        const T* __restrict__ val,
                                                        const T* __restrict__ r, T* y) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= count) return;
  const int i = perm[begin + t];
  int k0, k1, kd;
  row_range<T, UPPER>(row, i, k0, k1, kd);
  T s = T(0);
  for (int k = k0; k < k1; k++) s = s + val[k] * y[col[k]];
  y[i] = (r[i] - s) / val[kd];
}

template <typename T>
struct Sentinel;
template <>
struct Sentinel<double> {
  using U = unsigned long long;
  static constexpr U bits = 0x7FF4DEADBEEF5A5AULL;  // signalling-NaN payload
};
template <>
struct Sentinel<float> {
  using U = unsigned int;
  static constexpr U bits = 0x7FA5A5A5U;
};

template <typename T>
__global__ void fill_sentinel_kernel(int n, T* y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    typename Sentinel<T>::U b = Sentinel<T>::bits;
    y[i] = __builtin_bit_cast(T, b);
  }
}

// Sync-free exact solve.  Slot s of the level-ordered schedule is handled
// by thread s; a wavefront only ever holds rows of one level, so every row
// it waits on belongs to an earlier wavefront of this block or of a block
// with a smaller index (dispatched earlier), which guarantees progress.
// The solution vector doubles as the ready flag: y[j] holds the sentinel
// until row j is solved; each value is published with one agent-scope
// atomic store and read back with agent-scope atomic loads (coherent
// across the XCD L2s).  Arithmetic results are never the sentinel (it is a
// signalling NaN; arithmetic yields quiet NaNs).  A wait that exceeds
// spin_limit polls sets *err and stops waiting, so every wave terminates.
template <typename T, bool UPPER>
__global__ __launch_bounds__(256) void tri_syncfree_kernel(int nslots, const int* __restrict__ sched,
                                                           const int* __restrict__ row,
                                                           const int* __restrict__ col,
                                                           const T* __restrict__ val,
                                                           const T* __restrict__ r, T* y, int* err,
                                                           int spin_limit) {
#pragma clang fp contract(off)
  using UB = typename Sentinel<T>::U;
  const int slot = blockIdx.x * 256 + threadIdx.x;
  if (slot >= nslots) return;
  const int i = sched[slot];
  if (i < 0) return;
  int k0, k1, kd;
  row_range<T, UPPER>(row, i, k0, k1, kd);
  T s = T(0);
  for (int k = k0; k < k1; k++) {
    const int j = col[k];
    UB* pj = reinterpret_cast<UB*>(y + j);
    UB b = __hip_atomic_load(pj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int spins = 0;
    while (b == Sentinel<T>::bits) {
      if (++spins > spin_limit) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
      b = __hip_atomic_load(pj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s = s + val[k] * __builtin_bit_cast(T, b);
  }
  const T yi = (r[i] - s) / val[kd];
  __hip_atomic_store(reinterpret_cast<UB*>(y + i), __builtin_bit_cast(UB, yi), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// One synchronous Jacobi sweep: yout_i = (r_i - sum_{k != diag} T_ik yin_k)/T_ii.
// yin == nullptr means yin = 0.  STEP accumulates (yout - yin)^2 per block.
template <typename T, bool UPPER, bool STEP>
__global__ __launch_bounds__(256) void tri_jacobi_kernel(int n, const int* __restrict__ row,
                                                         const int* __restrict__ col,
                                                         const T* __restrict__ val,
                                                         const T* __restrict__ r,
                                                         const T* __restrict__ yin,
                                                         T* __restrict__ yout,
                                                         double* __restrict__ partial) {
#pragma clang fp contract(off)
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int i = b * 256 + threadIdx.x;
  double d2 = 0.0;
  if (i < n) {
    int k0, k1, kd;
    row_range<T, UPPER>(row, i, k0, k1, kd);
    T s = T(0);
    if (yin)
      for (int k = k0; k < k1; k++) s = s + val[k] * yin[col[k]];
    const T v = (r[i] - s) / val[kd];
    if (STEP) {
      double d = (double)v - (yin ? (double)yin[i] : 0.0);
      d2 = d * d;
    }
    yout[i] = v;
  }
  if (STEP) {
    __shared__ double sm[4];
    double tot = block_sum<256>(d2, sm);
    if (threadIdx.x == 0) partial[b] = tot;
  }
}

// Jacobi sweep in CSR-stream form (yin != null, no step): a block owns
// 256*R consecutive rows (thread t: rows r0 + t + 256 q); their nonzeros
// form one contiguous range of col/val, read by consecutive threads into
// LDS as val*yin[col]; each thread then sums its rows' off-diagonal
// products in column order and divides by the diagonal -- the same
// sequence as tri_jacobi_kernel.  Row bounds, r and the diagonal are
// loaded before the staging so their latency overlaps it.
template <typename T, int R>
struct JacCap {
  static constexpr int rows = 256 * R;
  static constexpr int cap = (rows * 8 * (int)sizeof(T) <= 32768) ? rows * 8 : 32768 / (int)sizeof(T);
};

template <typename T, bool UPPER, int R>
__global__ __launch_bounds__(256) void tri_jacobi_stream_kernel(
    int n, const int* __restrict__ row, const int* __restrict__ col, const T* __restrict__ val,
    const T* __restrict__ r, const T* __restrict__ yin, T* __restrict__ yout) {
#pragma clang fp contract(off)
  constexpr int ROWS = JacCap<T, R>::rows, CAP = JacCap<T, R>::cap;
  __shared__ T prod[CAP];
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int r0 = b * ROWS;
  const int r1 = min(r0 + ROWS, n);
  int k0[R], k1[R];
  T rr[R], dg[R], s[R];
#pragma unroll
  for (int q = 0; q < R; q++) {
    const int i = r0 + (int)threadIdx.x + 256 * q;
    k0[q] = k1[q] = 0;
    rr[q] = s[q] = T(0);
    dg[q] = T(1);
    if (i < r1) {
      int kd;
      row_range<T, UPPER>(row, i, k0[q], k1[q], kd);
      rr[q] = r[i];
      dg[q] = val[kd];
    }
  }
  const int c_begin = row[r0], c_end = row[r1];
  for (int c0 = c_begin; c0 < c_end; c0 += CAP) {
    const int cn = min(CAP, c_end - c0);
    stage_products<256, T>(col, val, yin, c0, cn, prod);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R; q++) {
      const int lo = max(k0[q], c0), hi = min(k1[q], c0 + cn);
      for (int k = lo; k < hi; k++) s[q] = s[q] + prod[k - c0];
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < R; q++) {
    const int i = r0 + (int)threadIdx.x + 256 * q;
    if (i < r1) yout[i] = (rr[q] - s[q]) / dg[q];
  }
}

// The same sweep on the plan's off-diagonal form: only off-diagonal values
// and columns are streamed and the diagonal comes from its own array, so a
// row no longer costs a diagonal value, column and gather; with a unit
// diagonal the division by 1 (exact) is dropped.  Same sums in the same
// order, so the result equals the CSR-stream kernel bit for bit.
template <typename T, bool UNIT, int R>
__global__ __launch_bounds__(256) void tri_jacobi_strict_kernel(
    int n, const int* __restrict__ srow, const int* __restrict__ scol, const T* __restrict__ sval,
    const T* __restrict__ dval, const T* __restrict__ r, const T* __restrict__ yin, T* __restrict__ yout) {
#pragma clang fp contract(off)
  constexpr int ROWS = JacCap<T, R>::rows, CAP = JacCap<T, R>::cap;
  __shared__ T prod[CAP];
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int r0 = b * ROWS;
  const int r1 = min(r0 + ROWS, n);
  int k0[R], k1[R];
  T rr[R], dg[R], s[R];
#pragma unroll
  for (int q = 0; q < R; q++) {
    const int i = r0 + (int)threadIdx.x + 256 * q;
    k0[q] = k1[q] = 0;
    rr[q] = s[q] = T(0);
    dg[q] = T(1);
    if (i < r1) {
      k0[q] = srow[i];
      k1[q] = srow[i + 1];
      rr[q] = r[i];
      if (!UNIT) dg[q] = dval[i];
    }
  }
  const int c_begin = srow[r0], c_end = srow[r1];
  for (int c0 = c_begin; c0 < c_end; c0 += CAP) {
    const int cn = min(CAP, c_end - c0);
    stage_products<256, T>(scol, sval, yin, c0, cn, prod);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R; q++) {
      const int lo = max(k0[q], c0), hi = min(k1[q], c0 + cn);
      for (int k = lo; k < hi; k++) s[q] = s[q] + prod[k - c0];
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < R; q++) {
    const int i = r0 + (int)threadIdx.x + 256 * q;
    if (i < r1) yout[i] = UNIT ? (rr[q] - s[q]) : (rr[q] - s[q]) / dg[q];
  }
}

// The same sweep on the ELLPACK form: thread per row, the W slots of the row
// read as coalesced slot-major words, padding (column -1) skipped; the sum
// visits the row's entries in CSR order, so results are bit-identical.
template <typename T, bool UNIT, int W>
__global__ __launch_bounds__(256) void tri_jacobi_ell_kernel(int n, const int* __restrict__ ecol,
                                                             const T* __restrict__ eval,
                                                             const T* __restrict__ dval,
                                                             const T* __restrict__ r,
                                                             const T* __restrict__ yin,
                                                             T* __restrict__ yout) {
#pragma clang fp contract(off)
  const int i = xcd_block(blockIdx.x, gridDim.x) * 256 + (int)threadIdx.x;
  if (i >= n) return;
  int c[W];
  T v[W], y[W];
#pragma unroll
  for (int q = 0; q < W; q++) {
    c[q] = ecol[(size_t)q * n + i];
    v[q] = eval[(size_t)q * n + i];
  }
#pragma unroll
  for (int q = 0; q < W; q++) y[q] = c[q] >= 0 ? yin[c[q]] : T(0);
  const T ri = r[i];
  T s = T(0);
#pragma unroll
  for (int q = 0; q < W; q++)
    if (c[q] >= 0) s = s + v[q] * y[q];
  yout[i] = UNIT ? ri - s : (ri - s) / dval[i];
}

template <typename T, bool UNIT>
static void jacobi_ell(const TriPlan& P, const T* r, const T* yin, T* yout) {
  const int nb = ceil_div(P.n, 256);
  const int* ec = P.ecol.as<int>();
  const T* ev = P.eval.as<T>();
  const T* dv = P.dval.as<T>();
  switch (P.ew) {
#define LACE_ELL_CASE(W)                                                                              \
  case W:                                                                                             \
    tri_jacobi_ell_kernel<T, UNIT, W><<<nb, 256, 0, stream()>>>(P.n, ec, ev, dv, r, yin, yout); \
    break;
    LACE_ELL_CASE(1)
    LACE_ELL_CASE(2)
    LACE_ELL_CASE(3)
    LACE_ELL_CASE(4)
    LACE_ELL_CASE(5)
    LACE_ELL_CASE(6)
    LACE_ELL_CASE(7)
    LACE_ELL_CASE(8)
#undef LACE_ELL_CASE
    default:
      break;
  }
}

// The same sweep on the diagonal form: thread per row, the row's value on
// diagonal d read at d*n + i and y at i + off[d] (contiguous across the
// wavefront; rows without an entry there read y[i] and skip it), summed in
// ascending offset = CSR order, so results are bit-identical.  EPI: the
// sweep also writes y2 = yout / d2, the first Jacobi sweep (from zero) of
// the next triangular solve, whose diagonal is d2.
template <typename T, bool UNIT, int D, bool EPI>
__global__ __launch_bounds__(256) void tri_jacobi_dia_kernel(int n, DiaOffsets off,
                                                             const uint8_t* __restrict__ mask,
                                                             const T* __restrict__ dia,
                                                             const T* __restrict__ dval,
                                                             const T* __restrict__ r,
                                                             const T* __restrict__ yin,
                                                             T* __restrict__ yout, T* __restrict__ y2,
                                                             const T* __restrict__ d2) {
#pragma clang fp contract(off)
  const int i = xcd_block(blockIdx.x, gridDim.x) * 256 + (int)threadIdx.x;
  if (i >= n) return;
  const unsigned m = mask[i];
  T v[D], y[D];
#pragma unroll
  for (int q = 0; q < D; q++) v[q] = dia[(size_t)q * n + i];
#pragma unroll
  for (int q = 0; q < D; q++) y[q] = yin[((m >> q) & 1u) ? i + off.o[q] : i];
  const T ri = r[i];
  T s = T(0);
#pragma unroll
  for (int q = 0; q < D; q++)
    if ((m >> q) & 1u) s = s + v[q] * y[q];
  const T o = UNIT ? ri - s : (ri - s) / dval[i];
  yout[i] = o;
  if (EPI) y2[i] = o / d2[i];
}

template <typename T, bool UNIT>
static void jacobi_dia(const TriPlan& P, const T* r, const T* yin, T* yout, T* y2, const T* d2,
                       TimedRegion& tr) {
  const int nb = ceil_div(P.n, 256);
  const uint8_t* mk = P.dia.mask.as<uint8_t>();
  const T* dv = P.dia.val.as<T>();
  const T* dg = P.dval.as<T>();
  T* no_y2 = nullptr;
  const T* no_d2 = nullptr;
  switch (P.dia.nd) {
#define LACE_DIA_CASE(D)                                                                                  \
  case D:                                                                                                 \
    if (y2)                                                                                               \
      tr.launch(tri_jacobi_dia_kernel<T, UNIT, D, true>, dim3(nb), dim3(256), P.n, P.dia.off, mk, dv, dg, \
                r, yin, yout, y2, d2);                                                                    \
    else                                                                                                  \
      tr.launch(tri_jacobi_dia_kernel<T, UNIT, D, false>, dim3(nb), dim3(256), P.n, P.dia.off, mk, dv,    \
                dg, r, yin, yout, no_y2, no_d2);                                                          \
    break;
    LACE_DIA_CASE(1)
    LACE_DIA_CASE(2)
    LACE_DIA_CASE(3)
    LACE_DIA_CASE(4)
    LACE_DIA_CASE(5)
    LACE_DIA_CASE(6)
    LACE_DIA_CASE(7)
    LACE_DIA_CASE(8)
#undef LACE_DIA_CASE
    default:
      break;
  }
}

// First sweep from zero: y = r / diag.
template <typename T>
__global__ void tri_diag_div_kernel(int n, const T* __restrict__ dval, const T* __restrict__ r,
                                    T* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = r[i] / dval[i];
}

template <typename T, bool UNIT, int R>
static void launch_jacobi_strict(const TriPlan& P, const T* r, const T* yin, T* yout) {
  tri_jacobi_strict_kernel<T, UNIT, R><<<ceil_div(P.n, JacCap<T, R>::rows), 256, 0, stream()>>>(
      P.n, P.srow.as<int>(), P.scol.as<int>(), P.sval.as<T>(), P.dval.as<T>(), r, yin, yout);
}

template <typename T, bool UNIT>
static void jacobi_strict(const TriPlan& P, const T* r, const T* yin, T* yout) {
  const int R = get_option("stream_rows_per_thread", sizeof(T) == 8 ? 1 : 2);
  if (R == 1) launch_jacobi_strict<T, UNIT, 1>(P, r, yin, yout);
  else if (R == 4) launch_jacobi_strict<T, UNIT, 4>(P, r, yin, yout);
  else launch_jacobi_strict<T, UNIT, 2>(P, r, yin, yout);
}

template <typename T, bool UPPER, int R>
static void launch_jacobi_stream(const TriPlan& P, const T* val, const T* r, const T* yin, T* yout) {
  tri_jacobi_stream_kernel<T, UPPER, R><<<ceil_div(P.n, JacCap<T, R>::rows), 256, 0, stream()>>>(
      P.n, P.row.as<int>(), P.col.as<int>(), val, r, yin, yout);
}

template <typename T, bool UPPER>
static void jacobi_stream(const TriPlan& P, const T* val, const T* r, const T* yin, T* yout) {
  // measured at 128^3: two rows per thread for FP32 values, one for FP64
  const int R = get_option("stream_rows_per_thread", sizeof(T) == 8 ? 1 : 2);
  if (R == 1) launch_jacobi_stream<T, UPPER, 1>(P, val, r, yin, yout);
  else if (R == 4) launch_jacobi_stream<T, UPPER, 4>(P, val, r, yin, yout);
  else launch_jacobi_stream<T, UPPER, 2>(P, val, r, yin, yout);
}

template <typename T>
__global__ void gather_kernel(int m, const int* __restrict__ map, const T* __restrict__ in,
                              T* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < m) out[p] = in[map[p]];
}

// ---------------------------------------------------------------- host
template <typename T>
int tri_solve_levels(const TriPlan& P, const T* val, const T* r, T* y) {
  TimedRegion tr("trsv_level");
  for (int l = 0; l < P.nlev; l++) {
    const int b = P.h_lptr[l], c = P.h_lptr[l + 1] - b;
    if (P.upper)
      tri_level_kernel<T, true><<<ceil_div(c, 256), 256, 0, stream()>>>(
          b, c, P.perm.as<int>(), P.row.as<int>(), P.col.as<int>(), val, r, y);
    else
      tri_level_kernel<T, false><<<ceil_div(c, 256), 256, 0, stream()>>>(
          b, c, P.perm.as<int>(), P.row.as<int>(), P.col.as<int>(), val, r, y);
  }
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

// Poll budget per dependency wait (about 0.1 s); a solve never legitimately
// waits that long on one row, and the bound keeps every wave finite.
static int g_spin_limit = 1 << 16;

template <typename T>
int tri_solve_syncfree(const TriPlan& P, const T* val, const T* r, T* y, int* err) {
  TimedRegion tr("trsv_syncfree");
  LACE_REQUIRE((const void*)r != (const void*)y, DEV_ERR_ILLEGAL_VALUE,
               "sync-free solve needs distinct r and y");
  fill_sentinel_kernel<T><<<ceil_div(P.n, 256), 256, 0, stream()>>>(P.n, y);
  if (P.upper)
    tri_syncfree_kernel<T, true><<<ceil_div(P.nslots, 256), 256, 0, stream()>>>(
        P.nslots, P.sched.as<int>(), P.row.as<int>(), P.col.as<int>(), val, r, y, err,
        g_spin_limit);
  else
    tri_syncfree_kernel<T, false><<<ceil_div(P.nslots, 256), 256, 0, stream()>>>(
        P.nslots, P.sched.as<int>(), P.row.as<int>(), P.col.as<int>(), val, r, y, err,
        g_spin_limit);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

// One sweep yout = D^{-1}(r - N yin).  With y2 and d2 given, the sweep on
// the diagonal form also writes y2 = yout / d2 and sets *epi.
template <typename T>
static int jacobi_sweep(const TriPlan& P, const T* val, const T* r, const T* yin, T* yout,
                        double* partial, T* y2 = nullptr, const T* d2 = nullptr, bool* epi = nullptr) {
  const int nb = ceil_div(P.n, 256);
  const bool strict = P.strict_src == (const void*)val;  // off-diagonal form is current
  if (!partial && yin && P.n > 0 && P.nnz <= 16LL * P.n) {
    const bool dia = strict && P.dia.nd > 0 && get_option("jacobi_dia", 1) != 0;
    const bool ell = !dia && strict && P.ew > 0 && get_option("jacobi_ell", 1) != 0;
    const bool fuse = dia && y2 && d2;
    // bytes the chosen form reads once each, plus r, the previous iterate
    // and the new one (and the diagonal unless it is a unit one)
    const double es = sizeof(T), n = P.n;
    const double diag = P.unit_diag ? 0.0 : es * n;
    double bytes = 3.0 * es * n + (fuse ? 2.0 * es * n : 0.0);
    if (dia) bytes += P.dia.bytes(sizeof(T)) + diag;
    else if (ell) bytes += (double)P.ew * n * (es + 4.0) + diag;
    else if (strict) bytes += (double)P.snnz * (es + 4.0) + 4.0 * (n + 1) + diag;
    else bytes += (double)P.nnz * (es + 4.0) + 4.0 * (n + 1);
    // the diagonal-form launch carries the events itself
    TimedRegion tr(P.upper ? "trsv_jacobi_sweep_upper" : "trsv_jacobi_sweep_lower", 0.0, bytes, dia);
    if (dia && P.unit_diag) jacobi_dia<T, true>(P, r, yin, yout, fuse ? y2 : nullptr, d2, tr);
    else if (dia) jacobi_dia<T, false>(P, r, yin, yout, fuse ? y2 : nullptr, d2, tr);
    else if (ell && P.unit_diag) jacobi_ell<T, true>(P, r, yin, yout);
    else if (ell) jacobi_ell<T, false>(P, r, yin, yout);
    else if (strict && P.unit_diag) jacobi_strict<T, true>(P, r, yin, yout);
    else if (strict) jacobi_strict<T, false>(P, r, yin, yout);
    else if (P.upper) jacobi_stream<T, true>(P, val, r, yin, yout);
    else jacobi_stream<T, false>(P, val, r, yin, yout);
    LACE_HIP(hipGetLastError());
    if (epi) *epi = fuse;
    return DEV_SUCCESS;
  }
  if (!partial && !yin && strict && P.n > 0) {
    tri_diag_div_kernel<T><<<nb, 256, 0, stream()>>>(P.n, P.dval.as<T>(), r, yout);
    LACE_HIP(hipGetLastError());
    return DEV_SUCCESS;
  }
  if (partial) {
    if (P.upper)
      tri_jacobi_kernel<T, true, true><<<nb, 256, 0, stream()>>>(
          P.n, P.row.as<int>(), P.col.as<int>(), val, r, yin, yout, partial);
    else
      tri_jacobi_kernel<T, false, true><<<nb, 256, 0, stream()>>>(
          P.n, P.row.as<int>(), P.col.as<int>(), val, r, yin, yout, partial);
  } else {
    if (P.upper)
      tri_jacobi_kernel<T, true, false><<<nb, 256, 0, stream()>>>(
          P.n, P.row.as<int>(), P.col.as<int>(), val, r, yin, yout, nullptr);
    else
      tri_jacobi_kernel<T, false, false><<<nb, 256, 0, stream()>>>(
          P.n, P.row.as<int>(), P.col.as<int>(), val, r, yin, yout, nullptr);
  }
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

template <typename T>
int tri_solve_jacobi(const TriPlan& P, const T* val, const T* r, T* y, T* tmp, int sweeps, T* epi_y,
                     const T* epi_d, bool* epi_done, bool first_done) {
  TimedRegion tr("trsv_jacobi");
  if (epi_done) *epi_done = false;
  if (sweeps <= 0) {
    LACE_HIP(hipMemsetAsync(y, 0, sizeof(T) * P.n, stream()));
    return DEV_SUCCESS;
  }
  if (P.unit_diag) {
    // the first sweep from y = 0 returns r / 1 = r exactly, so sweep 2 reads
    // r directly and one launch is saved
    if (sweeps == 1) {
      LACE_HIP(hipMemcpyAsync(y, r, sizeof(T) * P.n, hipMemcpyDeviceToDevice, stream()));
      return DEV_SUCCESS;
    }
    const int rest = sweeps - 1;
    T* a = (rest % 2 == 1) ? y : tmp;
    T* b = (rest % 2 == 1) ? tmp : y;
    // the last sweep may carry the next solve's first sweep
    auto sweep = [&](const T* yin, T* yout, bool last) {
      return last ? jacobi_sweep<T>(P, val, r, yin, yout, nullptr, epi_y, epi_d, epi_done)
                  : jacobi_sweep<T>(P, val, r, yin, yout, nullptr);
    };
    LACE_TRY(sweep(r, a, rest == 1));
    for (int s = 1; s < rest; s++) {
      LACE_TRY(sweep(a, b, s == rest - 1));
      std::swap(a, b);
    }
    return DEV_SUCCESS;
  }
  // ping-pong so that the last sweep lands in y; with first_done the first
  // sweep's result (r / diag) is already in its buffer
  T* a = (sweeps % 2 == 1) ? y : tmp;
  T* b = (sweeps % 2 == 1) ? tmp : y;
  if (!first_done) LACE_TRY(jacobi_sweep<T>(P, val, r, nullptr, a, nullptr));
  for (int s = 1; s < sweeps; s++) {
    LACE_TRY(jacobi_sweep<T>(P, val, r, a, b, nullptr));
    std::swap(a, b);
  }
  return DEV_SUCCESS;
}

template int tri_solve_levels<double>(const TriPlan&, const double*, const double*, double*);
template int tri_solve_levels<float>(const TriPlan&, const float*, const float*, float*);
template int tri_solve_syncfree<double>(const TriPlan&, const double*, const double*, double*, int*);
template int tri_solve_syncfree<float>(const TriPlan&, const float*, const float*, float*, int*);
template int tri_solve_jacobi<double>(const TriPlan&, const double*, const double*, double*, double*, int, double*,
                                       const double*, bool*, bool);
template int tri_solve_jacobi<float>(const TriPlan&, const float*, const float*, float*, float*, int, float*,
                                      const float*, bool*, bool);

// ---------------------------------------------------------------- ParILU solves
int ParILU::build_trsv() {
  LACE_TRY(lower.build(n, h_lrow.data(), h_lcol.data(), false));
  lower.unit_diag = true;  // CSRL with the unit diagonal (never updated by sweeps)
  LACE_TRY(lowerT.build(n, h_ucp.data(), h_uri.data(), false));
  // U in CSR: transpose of the CSC arrays; map CSR slot -> CSC slot.
  {
    std::vector<int> r(n + 1, 0), c(nnzU), map(nnzU);
    for (int p = 0; p < nnzU; p++) r[h_uri[p] + 1]++;
    for (int i = 0; i < n; i++) r[i + 1] += r[i];
    std::vector<int> fill(r.begin(), r.end() - 1);
    for (int j = 0; j < n; j++)
      for (int p = h_ucp[j]; p < h_ucp[j + 1]; p++) {
        int q = fill[h_uri[p]]++;
        c[q] = j;
        map[q] = p;
      }
    LACE_TRY(upper.build(n, r.data(), c.data(), true));
    LACE_TRY(upload(u_csr_from_csc, map.data(), sizeof(int) * nnzU));
  }
  // L^T in CSR (diagonal first): transpose of L's CSR arrays.
  {
    std::vector<int> r(n + 1, 0), c(nnzL), map(nnzL);
    for (int p = 0; p < nnzL; p++) r[h_lcol[p] + 1]++;
    for (int i = 0; i < n; i++) r[i + 1] += r[i];
    std::vector<int> fill(r.begin(), r.end() - 1);
    for (int i = 0; i < n; i++)
      for (int p = h_lrow[i]; p < h_lrow[i + 1]; p++) {
        int q = fill[h_lcol[p]]++;
        c[q] = i;
        map[q] = p;
      }
    LACE_TRY(upperT.build(n, r.data(), c.data(), true));
    upperT.unit_diag = true;
    LACE_TRY(upload(l_t_from_csr, map.data(), sizeof(int) * nnzL));
  }
  const size_t es = elem_size();
  LACE_TRY(uval_csr.alloc(es * nnzU));
  LACE_TRY(lval_t.alloc(es * nnzL));
  LACE_TRY(work.alloc(es * n));
  LACE_TRY(work2.alloc(es * n));
  LACE_TRY(syncfree_err.alloc(sizeof(int) * 4));
  LACE_HIP(hipMemsetAsync(syncfree_err.p, 0, sizeof(int) * 4, stream()));
  csr_dirty = true;
  return DEV_SUCCESS;
}

int ParILU::refresh_csr() {
  if (!csr_dirty) return DEV_SUCCESS;
  if (prec == LACE_F64) {
    gather_kernel<double><<<ceil_div(nnzU, 256), 256, 0, stream()>>>(
        nnzU, u_csr_from_csc.as<int>(), uval[cur].as<double>(), uval_csr.as<double>());
    gather_kernel<double><<<ceil_div(nnzL, 256), 256, 0, stream()>>>(
        nnzL, l_t_from_csr.as<int>(), lval[cur].as<double>(), lval_t.as<double>());
  } else {
    gather_kernel<float><<<ceil_div(nnzU, 256), 256, 0, stream()>>>(
        nnzU, u_csr_from_csc.as<int>(), uval[cur].as<float>(), uval_csr.as<float>());
    gather_kernel<float><<<ceil_div(nnzL, 256), 256, 0, stream()>>>(
        nnzL, l_t_from_csr.as<int>(), lval[cur].as<float>(), lval_t.as<float>());
  }
  LACE_HIP(hipGetLastError());
  // off-diagonal forms for the Jacobi sweeps, from the arrays the solves pass
  const size_t es = elem_size();
  LACE_TRY(lower.refresh_strict(lval[cur].p, es));
  LACE_TRY(upper.refresh_strict(uval_csr.p, es));
  LACE_TRY(lowerT.refresh_strict(uval[cur].p, es));
  LACE_TRY(upperT.refresh_strict(lval_t.p, es));
  csr_dirty = false;
  return DEV_SUCCESS;
}

template <typename T>
static int apply_t(ParILU* P, lace_trsv_kind_t kind, int sweeps, int trans, const T* r, T* z) {
  LACE_TRY(P->refresh_csr());
  T* tmp = P->work.as<T>();
  const TriPlan& first = trans ? P->lowerT : P->lower;
  const TriPlan& second = trans ? P->upperT : P->upper;
  const T* v1 = trans ? P->uval[P->cur].as<T>() : P->lval[P->cur].as<T>();
  const T* v2 = trans ? P->lval_t.as<T>() : P->uval_csr.as<T>();
  if (kind == LACE_TRSV_NONE) {
    LACE_HIP(hipMemcpyAsync(z, r, sizeof(T) * P->n, hipMemcpyDeviceToDevice, stream()));
    return DEV_SUCCESS;
  }
  if (kind == LACE_TRSV_LEVEL) {
    int* err = P->syncfree_err.as<int>();
    LACE_TRY(tri_solve_syncfree<T>(first, v1, r, tmp, err));
    LACE_TRY(tri_solve_syncfree<T>(second, v2, tmp, z, err));
    return DEV_SUCCESS;
  }
  // Jacobi: work holds the intermediate vector, work2 the ping-pong buffer.
  // For an odd sweep count the second solve's first sweep (tmp / diag,
  // written to z) rides on the first solve's last sweep when that runs on
  // the diagonal form; z is then distinct from every buffer that sweep reads.
  T* tmp2 = P->work2.as<T>();
  const bool fuse = first.unit_diag && !second.unit_diag && sweeps % 2 == 1 &&
                    second.strict_src == (const void*)v2 && (const void*)z != (const void*)r &&
                    get_option("jacobi_fuse", 1) != 0;
  bool fused = false;
  LACE_TRY(tri_solve_jacobi<T>(first, v1, r, tmp, tmp2, sweeps, fuse ? z : nullptr,
                               fuse ? second.dval.as<T>() : nullptr, &fused));
  LACE_TRY(tri_solve_jacobi<T>(second, v2, tmp, z, tmp2, sweeps, nullptr, nullptr, nullptr, fused));
  return DEV_SUCCESS;
}

int ParILU::apply(lace_trsv_kind_t kind, int sweeps, int trans, const void* r, void* z) {
  if (prec == LACE_F64)
    return apply_t<double>(this, kind, sweeps, trans, static_cast<const double*>(r),
                           static_cast<double*>(z));
  return apply_t<float>(this, kind, sweeps, trans, static_cast<const float*>(r),
                        static_cast<float*>(z));
}

}  // namespace lace

using namespace lace;

extern "C" {

int lace_parilu_apply(lace_parilu* Ph, lace_trsv_kind_t kind, int jacobi_sweeps, int trans,
                      const void* r, void* z) {
  LACE_TRY(ensure_init());
  return reinterpret_cast<ParILU*>(Ph)->apply(kind, jacobi_sweeps, trans, r, z);
}

int lace_syncfree_error(lace_parilu* Ph, int* err) {
  LACE_TRY(ensure_init());
  auto* P = reinterpret_cast<ParILU*>(Ph);
  LACE_HIP(hipMemcpyAsync(err, P->syncfree_err.p, sizeof(int), hipMemcpyDeviceToHost, stream()));
  LACE_HIP(hipStreamSynchronize(stream()));
  return DEV_SUCCESS;
}

// Iterated synchronous sweeps until sum (y_new - y_old)^2 <= tol or the
// step is not finite (src/trisolve.cpp:274,317), starting from the y
// passed in.  A triangular Jacobi iteration is exact after nlev sweeps, so
// at most nlev+1 sweeps run.  storage/uplo must name CSRL+Lower or
// CSRU+Upper as in the reference (:264-347).
int data_parcsrtrsv(const data_uplo_t uplo, const data_storage_t storage,
                    const data_diagorder_t diag, const int num_rows, const dataType* Aval,
                    const int* row, const int* col, const dataType* rhsval, dataType* yval,
                    const dataType tol, int* iter) {
  LACE_TRY(ensure_init());
  bool upper;
  if (storage == Magma_CSRL && uplo == MagmaLower && diag != Magma_NODIAG) upper = false;
  else if (storage == Magma_CSRU && uplo == MagmaUpper && diag != Magma_NODIAG) upper = true;
  else {
    set_error("data_parcsrtrsv: storage %d / uplo %d must be CSRL+Lower or CSRU+Upper",
              (int)storage, (int)uplo);
    return -1;
  }
  const int n = num_rows;
  TriPlan P;
  LACE_TRY(P.build(n, row, col, upper));
  DevBuf val, r, y[2], partial, sum;
  LACE_TRY(upload(val, Aval, sizeof(double) * std::max(P.nnz, 1)));
  LACE_TRY(upload(r, rhsval, sizeof(double) * n));
  LACE_TRY(upload(y[0], yval, sizeof(double) * n));
  LACE_TRY(y[1].alloc(sizeof(double) * n));
  const int nb = ceil_div(n, 256);
  LACE_TRY(partial.alloc(sizeof(double) * nb));
  std::vector<double> ph(nb);
  int it = 0, cur = 0;
  double step = 1.e8;
  while (step > tol && std::isfinite(step) && it <= P.nlev + 1) {
    LACE_TRY(jacobi_sweep<double>(P, val.as<double>(), r.as<double>(), y[cur].as<double>(),
                                  y[1 - cur].as<double>(), partial.as<double>()));
    LACE_HIP(hipMemcpyAsync(ph.data(), partial.p, sizeof(double) * nb, hipMemcpyDeviceToHost,
                            stream()));
    LACE_HIP(hipStreamSynchronize(stream()));
    step = 0.0;
    for (int b = 0; b < nb; b++) step += ph[b];
    cur = 1 - cur;
    it++;
  }
  LACE_TRY(download(yval, y[cur], sizeof(double) * n));
  *iter = it;
  return 0;
}

// src/trisolve.cpp:20-67
int data_forward_solve(data_d_matrix* L, data_d_matrix* x, data_d_matrix* rhs, const dataType tol,
                       int* iter) {
  if (!(L->storage_type == Magma_CSRL && L->fill_mode == MagmaLower &&
        L->diagorder_type != Magma_NODIAG)) {
    set_error("data_forward_solve: L must be CSRL / lower with a diagonal");
    return -1;
  }
  return data_parcsrtrsv(MagmaLower, Magma_CSRL, L->diagorder_type, L->num_rows, L->val, L->row,
                         L->col, rhs->val, x->val, tol, iter);
}

// src/trisolve.cpp:72-119
int data_backward_solve(data_d_matrix* U, data_d_matrix* x, data_d_matrix* rhs,
                        const dataType tol, int* iter) {
  if (!(U->storage_type == Magma_CSRU && U->fill_mode == MagmaUpper &&
        U->diagorder_type != Magma_NODIAG)) {
    set_error("data_backward_solve: U must be CSRU / upper with a diagonal");
    return -1;
  }
  return data_parcsrtrsv(MagmaUpper, Magma_CSRU, U->diagorder_type, U->num_rows, U->val, U->row,
                         U->col, rhs->val, x->val, tol, iter);
}

}  // extern "C"
