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
  maxoff = wmax;
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
  return build_tiles(row_, col_);
}

// Grid dimensions under which every off-diagonal entry of the triangular
// pattern (row, col) lies at non-negative coordinate steps from its row
// (towards lower rows for a lower factor, higher ones for an upper one),
// with row = x + nx (y + ny z).  The candidates come from the distinct
// distances |i - j| of the off-diagonal entries: {1, a, b} with a | b and
// b | n gives (a, b / a, n / b); {1, a} with a | n gives (a, n / a, 1).
static bool tri_grid(int n, const int* row, const int* col, bool upper, long dims[3]) {
  std::vector<long> offs;
  for (int i = 0; i < n; i++)
    for (int k = row[i]; k < row[i + 1]; k++) {
      const long d = upper ? (long)col[k] - i : (long)i - col[k];
      if (d == 0) continue;
      if (d < 0) return false;
      if (std::find(offs.begin(), offs.end(), d) == offs.end()) {
        if (offs.size() == 3) return false;
        offs.push_back(d);
      }
    }
  std::sort(offs.begin(), offs.end());
  if (offs.size() < 2 || offs[0] != 1) return false;
  long nx, ny, nz;
  if (offs.size() == 3) {
    if (offs[2] % offs[1] != 0 || n % offs[2] != 0) return false;
    nx = offs[1];
    ny = offs[2] / offs[1];
    nz = n / offs[2];
  } else {
    if (n % offs[1] != 0) return false;
    nx = offs[1];
    ny = n / offs[1];
    nz = 1;
  }
  for (int i = 0; i < n; i++)
    for (int k = row[i]; k < row[i + 1]; k++) {
      const long a = upper ? (long)col[k] : i, b = upper ? (long)i : col[k];  // a >= b entrywise
      if (a % nx < b % nx || (a / nx) % ny < (b / nx) % ny || a / (nx * ny) < b / (nx * ny)) return false;
    }
  dims[0] = nx;
  dims[1] = ny;
  dims[2] = nz;
  return true;
}

int TriPlan::build_tiles(const int* row_, const int* col_) {
  tile_dims[0] = tile_dims[1] = tile_dims[2] = 0;
  tval_src = nullptr;
  long g[3];
  if (n < 2 || !tri_grid(n, row_, col_, upper, g)) return DEV_SUCCESS;
  // 512 rows per box: 8 x 8 x 8 in 3-D, 32 x 16 in 2-D (sides clipped to
  // the grid)
  const long want[3] = {g[2] > 1 ? 8 : 32, g[2] > 1 ? 8 : 16, 8};
  long b[3], nbx[3];
  for (int d = 0; d < 3; d++) {
    b[d] = std::min(want[d], g[d]);
    nbx[d] = (g[d] + b[d] - 1) / b[d];
  }
  std::vector<int> pm;
  pm.reserve(n);
  for (long s = 0; s <= nbx[0] + nbx[1] + nbx[2] - 3; s++)
    for (long Z = 0; Z < nbx[2]; Z++)
      for (long Y = 0; Y < nbx[1]; Y++) {
        const long X = s - Y - Z;
        if (X < 0 || X >= nbx[0]) continue;
        for (long z = Z * b[2]; z < std::min(g[2], (Z + 1) * b[2]); z++)
          for (long y = Y * b[1]; y < std::min(g[1], (Y + 1) * b[1]); y++)
            for (long x = X * b[0]; x < std::min(g[0], (X + 1) * b[0]); x++)
              pm.push_back((int)(x + g[0] * (y + g[1] * z)));
      }
  if ((int)pm.size() != n) return DEV_SUCCESS;
  std::vector<int> inv(n);
  for (int p = 0; p < n; p++) inv[pm[p]] = p;
  std::vector<int> tr(n + 1), tc(std::max(nnz, 1)), tm(std::max(nnz, 1));
  tr[0] = 0;
  for (int p = 0; p < n; p++) {
    const int i = pm[p];
    int q = tr[p];
    for (int k = row_[i]; k < row_[i + 1]; k++, q++) {
      tc[q] = inv[col_[k]];
      tm[q] = k;
      // the renumbered factor must stay triangular (true for the order
      // above; checked so that a mistake cannot stall the solve)
      if (upper ? tc[q] < p : tc[q] > p) return DEV_SUCCESS;
    }
    tr[p + 1] = q;
  }
  LACE_TRY(upload(tperm, pm.data(), sizeof(int) * n));
  LACE_TRY(upload(trow, tr.data(), sizeof(int) * (n + 1)));
  LACE_TRY(upload(tcol, tc.data(), sizeof(int) * tc.size()));
  LACE_TRY(upload(tmap, tm.data(), sizeof(int) * tm.size()));
  for (int d = 0; d < 3; d++) tile_dims[d] = (int)b[d];
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
  if (ew > 0 && eval.bytes < es * (size_t)ew * n) LACE_TRY(eval.alloc(es * (size_t)ew * n));
  strict_src = val;
  src_es = es;
  stale = kFormAll;
  tval_src = nullptr;
  return DEV_SUCCESS;
}

int TriPlan::ensure_forms(unsigned forms) {
  const unsigned todo = stale & forms;
  if (!todo || !strict_src) return DEV_SUCCESS;
  const size_t es = src_es;
  const void* val = strict_src;
  // off-diagonal CSR values and the diagonal (one launch: the diagonal
  // alone where the CSR copy is not wanted)
  const int sn = (todo & kFormStrict) ? snnz : 0;
  const int dn = (todo & (kFormStrict | kFormDiag)) ? n : 0;
  const int m = std::max(sn, dn);
  if (m > 0) {
    if (es == 8)
      strict_gather_kernel<double><<<ceil_div(m, 256), 256, 0, stream()>>>(
          sn, dn, smap.as<int>(), dpos.as<int>(), static_cast<const double*>(val), sval.as<double>(),
          dval.as<double>());
    else
      strict_gather_kernel<float><<<ceil_div(m, 256), 256, 0, stream()>>>(
          sn, dn, smap.as<int>(), dpos.as<int>(), static_cast<const float*>(val), sval.as<float>(),
          dval.as<float>());
    LACE_HIP(hipGetLastError());
  }
  if ((todo & kFormEll) && ew > 0) {
    const size_t m2 = (size_t)ew * n;
    if (es == 8)
      ell_gather_kernel<double><<<ceil_div((long)m2, 256), 256, 0, stream()>>>(
          m2, emap.as<int>(), static_cast<const double*>(val), eval.as<double>());
    else
      ell_gather_kernel<float><<<ceil_div((long)m2, 256), 256, 0, stream()>>>(
          m2, emap.as<int>(), static_cast<const float*>(val), eval.as<float>());
    LACE_HIP(hipGetLastError());
  }
  if (todo & kFormDia) LACE_TRY(dia.gather(val, es));
  stale &= ~todo;
  if (dn > 0) stale &= ~kFormDiag;
  return DEV_SUCCESS;
}

// 16-bit copies of the diagonal form and the diagonal, row i relative to
// its scale rs[i] = 2^e (stored as the exponent e).
template <typename TS>
__global__ void dia16_encode_kernel(int n, int nd, const float* __restrict__ dia, const float* __restrict__ dval,
                                    const float* __restrict__ rs, TS* __restrict__ dia16, TS* __restrict__ dval16,
                                    int8_t* __restrict__ dexp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float sc = rs[i];
  const float sinv = 1.0f / sc;
  for (int q = 0; q < nd; q++) dia16[(size_t)q * n + i] = quant<float, TS>(dia[(size_t)q * n + i], sinv);
  if (dval16) dval16[i] = quant<float, TS>(dval[i], sinv);
  dexp[i] = (int8_t)((int)((__float_as_uint(sc) >> 23) & 0xffu) - 127);
}

int TriPlan::refresh_dia16(lace_prec_t fmt, const float* rs) {
  dia16_fmt = LACE_F32;
  dia16_src = nullptr;
  if (dia.nd == 0 || !strict_src) return DEV_SUCCESS;
  LACE_TRY(ensure_forms(kFormDia | kFormDiag));
  LACE_REQUIRE(fmt == LACE_BF16 || fmt == LACE_F16, DEV_ERR_ILLEGAL_VALUE, "16-bit DIA form: format %d", (int)fmt);
  const size_t m = (size_t)dia.nd * n;
  if (dia16.bytes < 2 * m) LACE_TRY(dia16.alloc(2 * m));
  if (dexp.bytes < (size_t)n) LACE_TRY(dexp.alloc(n));
  if (!unit_diag && dval16.bytes < 2 * (size_t)n) LACE_TRY(dval16.alloc(2 * (size_t)n));
  const unsigned nb = (unsigned)ceil_div(n, 256);
  if (fmt == LACE_BF16)
    dia16_encode_kernel<uint16_t><<<nb, 256, 0, stream()>>>(n, dia.nd, dia.val.as<float>(), dval.as<float>(), rs,
                                                             dia16.as<uint16_t>(),
                                                             unit_diag ? nullptr : dval16.as<uint16_t>(),
                                                             dexp.as<int8_t>());
  else
    dia16_encode_kernel<_Float16><<<nb, 256, 0, stream()>>>(n, dia.nd, dia.val.as<float>(), dval.as<float>(), rs,
                                                             dia16.as<_Float16>(),
                                                             unit_diag ? nullptr : dval16.as<_Float16>(),
                                                             dexp.as<int8_t>());
  LACE_HIP(hipGetLastError());
  dia16_fmt = fmt;
  dia16_src = strict_src;
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

// Persistent exact solve in natural row order.  Blocks of BS consecutive
// rows go to workgroups in ticket order (an atomic counter): ascending for a
// lower factor, descending for an upper one, so every row a block reads from
// outside itself belongs to a block handed out earlier.  A workgroup has
// 2 BS lanes: lane t < BS computes row lo + t, lane BS + t fetches that
// row's external operands.
//   - LDS holds the block's solved rows (cells 0..BS-1), the staged external
//     operands (cell BS + q BS + t for slot q of local row t: consecutive
//     lanes touch consecutive cells) and one cell that never holds the
//     sentinel.
//   - The fetch lane polls the solution vector (device-coherent loads; the
//     NaN sentinel marks a row not yet solved) for the row's first R
//     off-diagonal operands that lie outside the block and stages each one
//     as it arrives.
//   - The compute lane keeps, for its row's first R slots, the LDS cell of
//     the operand (a solved row of the block or a staged external value)
//     and the matrix value in registers.  Each round it reads all R cells
//     at once; when none holds the sentinel it sums the products in CSR
//     order and divides.  It never waits on device memory, so a dependency
//     chain inside the block moves at LDS speed.  Operands past the first R
//     (rows longer than R) come straight from memory afterwards.  Waves
//     with nothing to do back off (up to idle_max x 64 cycles) to leave
//     the SIMD to waves that can move.
//   - A solved row goes to LDS for the block and to the solution vector
//     (device-coherent store) for later blocks.
// The earliest unfinished block always has every external operand solved
// (they lie in earlier blocks) and its earliest unsolved row has every
// operand it needs, so every block finishes.  Each lane also gives up after
// spin_limit rounds (setting *err), so every wave exits whatever happens.
// The row sum runs over the row's entries in CSR order and then divides, as
// the level-scheduled kernel does: bit-identical to substitution.
// y must not alias r; y is filled with the sentinel and *ticket zeroed by
// tri_persist_init_kernel first.  With rperm given, the rows are a
// renumbering of the caller's (row i is the caller's row rperm[i]): the
// right-hand side is read as r[rperm[i]] and each solved value is also
// stored to out[rperm[i]].
template <typename T, bool UPPER, int BS, int R>
__global__ __launch_bounds__(2 * BS) void tri_persist_kernel(int n, int nblk, const int* __restrict__ row,
                                                             const int* __restrict__ col,
                                                             const T* __restrict__ val,
                                                             const T* __restrict__ r, T* y, unsigned* ticket,
                                                             int* err, int spin_limit, int poll_sleep,
                                                             const int* __restrict__ rperm,
                                                             T* __restrict__ out, int idle_max) {
#pragma clang fp contract(off)
  using UB = typename Sentinel<T>::U;
  constexpr UB SENT = Sentinel<T>::bits;
  constexpr int kOne = BS + BS * R;  // the cell that never holds the sentinel
  __shared__ UB lv[BS + BS * R + 1];
  __shared__ int s_blk, s_next;
  const int tid = threadIdx.x;
  const bool fetcher = tid >= BS;
  const int t = fetcher ? tid - BS : tid;
  UB* yb = reinterpret_cast<UB*>(y);
  if (tid == 0) {
    s_blk = (int)__hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&lv[kOne], UB(0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();
  int blk = s_blk;
  while (blk < nblk) {
    int lo, hi;
    if (UPPER) {
      hi = n - blk * BS;
      lo = hi - BS > 0 ? hi - BS : 0;
    } else {
      lo = blk * BS;
      hi = lo + BS < n ? lo + BS : n;
    }
    if (!fetcher) {
      __hip_atomic_store(&lv[t], SENT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
      for (int q = 0; q < R; q++)
        __hip_atomic_store(&lv[BS + q * BS + t], SENT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    const int i = lo + t;
    const bool active = i < hi;
    int k0 = 0, len = 0, kd = 0;
    if (active) {
      if (UPPER) {
        kd = row[i];
        k0 = kd + 1;
        len = row[i + 1] - k0;
      } else {
        k0 = row[i];
        kd = row[i + 1] - 1;
        len = kd - k0;
      }
    }
    int cq[R];
#pragma unroll
    for (int q = 0; q < R; q++) cq[q] = q < len ? col[k0 + q] : -1;
    if (fetcher) {
      unsigned pend = 0u;
#pragma unroll
      for (int q = 0; q < R; q++)
        if (q < len && (cq[q] < lo || cq[q] >= hi)) pend |= 1u << q;
      int spins = 0;
      while (true) {
        UB b[R];
#pragma unroll
        for (int q = 0; q < R; q++)
          b[q] = ((pend >> q) & 1u) ? __hip_atomic_load(yb + cq[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                    : SENT;
#pragma unroll
        for (int q = 0; q < R; q++)
          if (((pend >> q) & 1u) && b[q] != SENT) {
            __hip_atomic_store(&lv[BS + q * BS + t], b[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            pend &= ~(1u << q);
          }
        if (pend != 0u && ++spins > spin_limit) {
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          pend = 0u;
        }
        if (!__any(pend != 0u)) break;
        for (int z = 0; z < poll_sleep; z++) __builtin_amdgcn_s_sleep(1);
      }
      // the next block's ticket, fetched while the compute lanes work
      if (t == 0)
        s_next = (int)__hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      int li[R];
      T vq[R];
#pragma unroll
      for (int q = 0; q < R; q++) {
        const int c = cq[q];
        li[q] = q < len ? ((c >= lo && c < hi) ? c - lo : BS + q * BS + t) : kOne;
        vq[q] = q < len ? val[k0 + q] : T(0);
      }
      const int oi = active ? (rperm ? rperm[i] : i) : 0;
      const T ri = active ? r[oi] : T(0);
      const T di = active ? val[kd] : T(1);
      const int head = len < R ? len : R;  // slots held in registers
      T s = T(0);
      int p = 0;  // slots summed so far
      bool done = !active;
      int spins = 0, idle = 0;
      // the row's loads complete here, so no wait inside the loop covers the
      // solved values' stores of earlier rounds
      __builtin_amdgcn_s_waitcnt(0);
      while (true) {
        bool moved = false;
        if (!done) {
          if (p == 0) {
            UB b[R];
#pragma unroll
            for (int q = 0; q < R; q++)
              b[q] = __hip_atomic_load(&lv[li[q]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            bool ready = true;
#pragma unroll
            for (int q = 0; q < R; q++) ready = ready && b[q] != SENT;
            if (ready) {
#pragma unroll
              for (int q = 0; q < R; q++) {
                const T a = s + vq[q] * __builtin_bit_cast(T, b[q]);
                s = q < head ? a : s;
              }
              p = head;
            }
          }
          // operands past the first R slots (rows longer than R)
          while (p >= R && p < len) {
            const int k = k0 + p;
            const int c = col[k];
            const UB b = (c >= lo && c < hi)
                             ? __hip_atomic_load(&lv[c - lo], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                             : __hip_atomic_load(yb + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (b == SENT) break;
            s = s + val[k] * __builtin_bit_cast(T, b);
            p++;
          }
          if (p == len) {
            const T yi = (ri - s) / di;
            const UB ub = __builtin_bit_cast(UB, yi);
            __hip_atomic_store(&lv[t], ub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_store(yb + i, ub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (rperm) out[oi] = yi;
            done = true;
            moved = true;
          } else if (++spins > spin_limit) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            done = true;
          }
        }
        if (__all(done)) break;
        if (__any(moved)) {
          idle = 0;
        } else {
          idle = idle < idle_max ? idle + 1 : idle_max;
          for (int z = 0; z < idle; z++) __builtin_amdgcn_s_sleep(1);
        }
      }
    }
    __syncthreads();
    if (tid == 0) s_blk = s_next;
    __syncthreads();
    blk = s_blk;
  }
}

// y = sentinel everywhere, ticket = 0: the state tri_persist_kernel starts
// from.
template <typename T>
__global__ void tri_persist_init_kernel(int n, T* y, unsigned* ticket) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    typename Sentinel<T>::U b = Sentinel<T>::bits;
    y[i] = __builtin_bit_cast(T, b);
  }
  if (i == 0) *ticket = 0u;
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

// R consecutive elements (16 bytes) at a 16-byte aligned address.
template <typename T>
struct alignas(16) Pack16 {
  static constexpr int R = 16 / (int)sizeof(T);
  T v[R];
};
template <typename T>
__device__ __forceinline__ Pack16<T> ld16(const T* p) {
  return *reinterpret_cast<const Pack16<T>*>(p);
}
template <typename T>
__device__ __forceinline__ void st16(T* p, const Pack16<T>& a) {
  *reinterpret_cast<Pack16<T>*>(p) = a;
}

// y[k] = (lo ++ hi)[k + sh] for k < R, sh in [0, R) the same in every lane.
template <typename T>
__device__ __forceinline__ void window16(const Pack16<T>& lo, const Pack16<T>& hi, int sh,
                                         T (&y)[Pack16<T>::R]) {
  constexpr int R = Pack16<T>::R;
  T c[2 * R];
#pragma unroll
  for (int k = 0; k < R; k++) {
    c[k] = lo.v[k];
    c[R + k] = hi.v[k];
  }
  if constexpr (R == 4) {
    switch (sh) {
      case 0: y[0] = c[0]; y[1] = c[1]; y[2] = c[2]; y[3] = c[3]; break;
      case 1: y[0] = c[1]; y[1] = c[2]; y[2] = c[3]; y[3] = c[4]; break;
      case 2: y[0] = c[2]; y[1] = c[3]; y[2] = c[4]; y[3] = c[5]; break;
      default: y[0] = c[3]; y[1] = c[4]; y[2] = c[5]; y[3] = c[6]; break;
    }
  } else {
    if (sh == 0) {
      y[0] = c[0];
      y[1] = c[1];
    } else {
      y[0] = c[1];
      y[1] = c[2];
    }
  }
}

// The diagonal-form sweep with R = 16 / sizeof(T) consecutive rows per
// thread (4 FP32, 2 FP64): the diagonals, r, the masks, the result and
// the epilogue move as 16-byte words, and y at rows i0 + off + k comes from
// the one or two aligned words that hold them (off mod R is the same in
// every lane).  A word wholly outside [0, n) is not read: no row that reads
// it has an entry there.  Each row's sum is formed exactly as in the kernel
// above, so results are bit-identical.  Needs n % R == 0 and 16-byte
// aligned vectors.
template <typename T, bool UNIT, int D, bool EPI>
__global__ __launch_bounds__(256) void tri_jacobi_dia_vec_kernel(int n, DiaOffsets off,
                                                                 const uint8_t* __restrict__ mask,
                                                                 const T* __restrict__ dia,
                                                                 const T* __restrict__ dval,
                                                                 const T* __restrict__ r,
                                                                 const T* __restrict__ yin,
                                                                 T* __restrict__ yout, T* __restrict__ y2,
                                                                 const T* __restrict__ d2) {
#pragma clang fp contract(off)
  constexpr int R = Pack16<T>::R;
  const int i0 = (xcd_block(blockIdx.x, gridDim.x) * 256 + (int)threadIdx.x) * R;
  if (i0 >= n) return;
  unsigned mk[R];
  if constexpr (R == 4) {
    const uint32_t m4 = *reinterpret_cast<const uint32_t*>(mask + i0);
#pragma unroll
    for (int k = 0; k < R; k++) mk[k] = (m4 >> (8 * k)) & 0xffu;
  } else {
    const uint16_t m2 = *reinterpret_cast<const uint16_t*>(mask + i0);
#pragma unroll
    for (int k = 0; k < R; k++) mk[k] = (m2 >> (8 * k)) & 0xffu;
  }
  Pack16<T> v[D];
#pragma unroll
  for (int q = 0; q < D; q++) v[q] = ld16(dia + (size_t)q * n + i0);
  const Pack16<T> rr = ld16(r + i0);
  T s[R];
#pragma unroll
  for (int k = 0; k < R; k++) s[k] = T(0);
#pragma unroll
  for (int q = 0; q < D; q++) {
    const int sh = ((off.o[q] % R) + R) % R;
    const int a = i0 + off.o[q] - sh;  // aligned word holding row i0 + off
    Pack16<T> lo{}, hi{};
    if (a >= 0 && a < n) lo = ld16(yin + a);
    if (sh != 0 && a + R >= 0 && a + R < n) hi = ld16(yin + a + R);
    T y[R];
    window16<T>(lo, hi, sh, y);
#pragma unroll
    for (int k = 0; k < R; k++)
      if ((mk[k] >> q) & 1u) s[k] = s[k] + v[q].v[k] * y[k];
  }
  Pack16<T> o;
  if constexpr (UNIT) {
#pragma unroll
    for (int k = 0; k < R; k++) o.v[k] = rr.v[k] - s[k];
  } else {
    const Pack16<T> dv = ld16(dval + i0);
#pragma unroll
    for (int k = 0; k < R; k++) o.v[k] = (rr.v[k] - s[k]) / dv.v[k];
  }
  st16(yout + i0, o);
  if constexpr (EPI) {
    const Pack16<T> dd = ld16(d2 + i0);
    Pack16<T> e;
#pragma unroll
    for (int k = 0; k < R; k++) e.v[k] = o.v[k] / dd.v[k];
    st16(y2 + i0, e);
  }
}

// Four 16-bit words from an 8-byte aligned address.
template <typename TS>
__device__ __forceinline__ void ld4x16(const TS* p, TS (&v)[4]) {
  const uint2 w = *reinterpret_cast<const uint2*>(p);
  const uint16_t h[4] = {(uint16_t)(w.x & 0xffffu), (uint16_t)(w.x >> 16), (uint16_t)(w.y & 0xffffu),
                         (uint16_t)(w.y >> 16)};
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = __builtin_bit_cast(TS, h[k]);
}

// The FP32 diagonal-form sweep (four rows per thread) on 16-bit factor
// words: each value is its word times the row's scale 2^e, which is exactly
// the FP32 factor value, and the sums run as in the FP32 kernel, so the
// result is bit-identical to it.  A row moves 2 bytes per diagonal, one
// mask byte and one exponent byte (and 2 bytes of diagonal) instead of 4
// bytes per diagonal (and 4 of diagonal).
template <typename TS, bool UNIT, int D, bool EPI>
__global__ __launch_bounds__(256) void tri_jacobi_dia16_vec_kernel(int n, DiaOffsets off,
                                                                   const uint8_t* __restrict__ mask,
                                                                   const int8_t* __restrict__ dexp,
                                                                   const TS* __restrict__ dia,
                                                                   const TS* __restrict__ dval,
                                                                   const float* __restrict__ r,
                                                                   const float* __restrict__ yin,
                                                                   float* __restrict__ yout, float* __restrict__ y2,
                                                                   const float* __restrict__ d2) {
#pragma clang fp contract(off)
  constexpr int R = 4;
  const int i0 = (xcd_block(blockIdx.x, gridDim.x) * 256 + (int)threadIdx.x) * R;
  if (i0 >= n) return;
  const uint32_t m4 = *reinterpret_cast<const uint32_t*>(mask + i0);
  const uint32_t e4 = *reinterpret_cast<const uint32_t*>(dexp + i0);
  unsigned mk[R];
  float sc[R];
#pragma unroll
  for (int k = 0; k < R; k++) {
    mk[k] = (m4 >> (8 * k)) & 0xffu;
    sc[k] = ldexpf(1.0f, (int)(int8_t)((e4 >> (8 * k)) & 0xffu));
  }
  float v[D][R];
#pragma unroll
  for (int q = 0; q < D; q++) {
    TS h[R];
    ld4x16<TS>(dia + (size_t)q * n + i0, h);
#pragma unroll
    for (int k = 0; k < R; k++) v[q][k] = dequant<float, TS>(h[k], sc[k]);
  }
  const Pack16<float> rr = ld16(r + i0);
  float s[R];
#pragma unroll
  for (int k = 0; k < R; k++) s[k] = 0.0f;
#pragma unroll
  for (int q = 0; q < D; q++) {
    const int sh = ((off.o[q] % R) + R) % R;
    const int a = i0 + off.o[q] - sh;
    Pack16<float> lo{}, hi{};
    if (a >= 0 && a < n) lo = ld16(yin + a);
    if (sh != 0 && a + R >= 0 && a + R < n) hi = ld16(yin + a + R);
    float y[R];
    window16<float>(lo, hi, sh, y);
#pragma unroll
    for (int k = 0; k < R; k++)
      if ((mk[k] >> q) & 1u) s[k] = s[k] + v[q][k] * y[k];
  }
  Pack16<float> o;
  if constexpr (UNIT) {
#pragma unroll
    for (int k = 0; k < R; k++) o.v[k] = rr.v[k] - s[k];
  } else {
    TS h[R];
    ld4x16<TS>(dval + i0, h);
#pragma unroll
    for (int k = 0; k < R; k++) o.v[k] = (rr.v[k] - s[k]) / dequant<float, TS>(h[k], sc[k]);
  }
  st16(yout + i0, o);
  if constexpr (EPI) {
    const Pack16<float> dd = ld16(d2 + i0);
    Pack16<float> e;
#pragma unroll
    for (int k = 0; k < R; k++) e.v[k] = o.v[k] / dd.v[k];
    st16(y2 + i0, e);
  }
}

// Device-coherent 16-byte word: two 8-byte agent-scope accesses, which
// bypass the per-CU and per-XCD caches that other workgroups do not share.
template <typename T>
__device__ __forceinline__ Pack16<T> ld16_coherent(const T* p) {
  const uint64_t* q = reinterpret_cast<const uint64_t*>(p);
  const uint64_t a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint64_t b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  Pack16<T> w;
  __builtin_memcpy(&w.v[0], &a, 8);
  __builtin_memcpy(reinterpret_cast<char*>(&w.v[0]) + 8, &b, 8);
  return w;
}
template <typename T>
__device__ __forceinline__ void st16_coherent(T* p, const Pack16<T>& w) {
  uint64_t a, b;
  __builtin_memcpy(&a, &w.v[0], 8);
  __builtin_memcpy(&b, reinterpret_cast<const char*>(&w.v[0]) + 8, 8);
  uint64_t* q = reinterpret_cast<uint64_t*>(p);
  __hip_atomic_store(q, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(q + 1, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// S Jacobi sweeps of a triangular factor on its diagonal form in one
// launch.  Tiles of 256 R rows (R = 16 / sizeof(T)) go to workgroups in
// ticket order (an atomic counter): ascending for a lower factor, whose
// rows read rows before them, descending for an upper one.  A tile keeps
// its factor rows, masks and r in registers for all S sweeps.  Sweep 1
// reads yin; sweep s > 1 of tile t reads sweep s-1 of its own rows from LDS
// and of tiles t-K..t-1 (t+1..t+K upper; K = ceil(max |offset| / tile
// rows)) from bufs[s-2], published by each tile's flag: every thread's device-coherent stores of
// sweep s-1 complete, then the workgroup barrier, then the flag store.
// The waiting tiles hold smaller tickets, so they started earlier and wait
// only on earlier tickets in turn: every wait ends.  Each wait also gives up
// after spin_limit polls (setting *err), so every wave exits whatever
// happens.  Flags count up across launches (flag_base + s), so nothing is
// reset between launches.  Sweep S writes yout (and, with EPI, y2 =
// yout / d2).  yin must not alias yout, y2 or bufs: tiles finish in any
// order.  Per row the arithmetic is that of the one-sweep kernels:
// bit-identical results.
template <typename T, bool UNIT, int D, bool EPI>
__global__ __launch_bounds__(256) void tri_jacobi_dia_pipe_kernel(
    int n, DiaOffsets off, int ntiles, int K, int upper, int S, const uint8_t* __restrict__ mask,
    const T* __restrict__ dia, const T* __restrict__ dval, const T* __restrict__ r, const T* __restrict__ yin,
    T* bufs, T* __restrict__ yout, T* __restrict__ y2, const T* __restrict__ d2, unsigned* flags,
    unsigned* ctr, unsigned ticket_base, unsigned flag_base, int* err, int spin_limit) {
#pragma clang fp contract(off)
  constexpr int R = Pack16<T>::R;
  constexpr int TR = 256 * R;
  __shared__ int s_tile;
  __shared__ Pack16<T> ys[256];  // this tile's previous sweep, word per thread
  if (threadIdx.x == 0) {
    const int tk = (int)(atomicAdd(ctr, 1u) - ticket_base);
    s_tile = upper ? ntiles - 1 - tk : tk;
  }
  __syncthreads();
  const int t = s_tile;
  const int t0 = t * TR;  // the tile's first row
  const int i0 = t0 + (int)threadIdx.x * R;
  const bool act = t >= 0 && t < ntiles && i0 < n;
  // word at aligned row a of the previous sweep: this tile's from LDS,
  // another tile's through device-coherent loads
  auto prev_word = [&](const T* src, int a) {
    return (a >= t0 && a < t0 + TR) ? ys[(a - t0) / R] : ld16_coherent(src + a);
  };
  unsigned mk[R];
  Pack16<T> v[D], rr{}, dv{};
#pragma unroll
  for (int k = 0; k < R; k++) mk[k] = 0u;
  if (act) {
    if constexpr (R == 4) {
      const uint32_t m4 = *reinterpret_cast<const uint32_t*>(mask + i0);
#pragma unroll
      for (int k = 0; k < R; k++) mk[k] = (m4 >> (8 * k)) & 0xffu;
    } else {
      const uint16_t m2 = *reinterpret_cast<const uint16_t*>(mask + i0);
#pragma unroll
      for (int k = 0; k < R; k++) mk[k] = (m2 >> (8 * k)) & 0xffu;
    }
#pragma unroll
    for (int q = 0; q < D; q++) v[q] = ld16(dia + (size_t)q * n + i0);
    rr = ld16(r + i0);
    if constexpr (!UNIT) dv = ld16(dval + i0);
  }
  for (int s = 1; s <= S; s++) {
    const T* src = s == 1 ? yin : bufs + (size_t)(s - 2) * n;
    if (s > 1) {
      // the K neighbouring tiles on the dependent side have finished s-1
      if (threadIdx.x < 64 && (int)threadIdx.x < K) {
        const int tt = upper ? t + 1 + (int)threadIdx.x : t - 1 - (int)threadIdx.x;
        if (tt >= 0 && tt < ntiles) {
          const unsigned want = flag_base + (unsigned)(s - 1);
          int spins = 0;
          while ((int)(__hip_atomic_load(flags + tt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - want) < 0) {
            if (++spins > spin_limit) {
              __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              break;
            }
            __builtin_amdgcn_s_sleep(2);
          }
        }
      }
      __syncthreads();
    }
    Pack16<T> o{};
    if (act) {
      T sum[R];
#pragma unroll
      for (int k = 0; k < R; k++) sum[k] = T(0);
#pragma unroll
      for (int q = 0; q < D; q++) {
        const int sh = ((off.o[q] % R) + R) % R;
        const int a = i0 + off.o[q] - sh;
        Pack16<T> lo{}, hi{};
        if (s == 1) {
          if (a >= 0 && a < n) lo = ld16(src + a);
          if (sh != 0 && a + R >= 0 && a + R < n) hi = ld16(src + a + R);
        } else {
          if (a >= 0 && a < n) lo = prev_word(src, a);
          if (sh != 0 && a + R >= 0 && a + R < n) hi = prev_word(src, a + R);
        }
        T y[R];
        window16<T>(lo, hi, sh, y);
#pragma unroll
        for (int k = 0; k < R; k++)
          if ((mk[k] >> q) & 1u) sum[k] = sum[k] + v[q].v[k] * y[k];
      }
#pragma unroll
      for (int k = 0; k < R; k++) o.v[k] = UNIT ? rr.v[k] - sum[k] : (rr.v[k] - sum[k]) / dv.v[k];
    }
    if (s < S) __syncthreads();  // every thread has read the previous sweep's LDS words
    if (act) {
      if (s == S) {
        st16(yout + i0, o);
        if constexpr (EPI) {
          const Pack16<T> dd = ld16(d2 + i0);
          Pack16<T> e;
#pragma unroll
          for (int k = 0; k < R; k++) e.v[k] = o.v[k] / dd.v[k];
          st16(y2 + i0, e);
        }
      } else {
        ys[threadIdx.x] = o;
        st16_coherent(bufs + (size_t)(s - 1) * n + i0, o);
      }
    }
    if (s < S) {
      __builtin_amdgcn_s_waitcnt(0);  // this thread's stores have completed
      __syncthreads();
      if (threadIdx.x == 0 && t >= 0 && t < ntiles)
        __hip_atomic_store(flags + t, flag_base + (unsigned)s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

template <typename T, bool UNIT>
static void jacobi_dia(const TriPlan& P, const T* r, const T* yin, T* yout, T* y2, const T* d2,
                       TimedRegion& tr) {
  constexpr int R = Pack16<T>::R;
  const uint8_t* mk = P.dia.mask.as<uint8_t>();
  const T* dv = P.dia.val.as<T>();
  const T* dg = P.dval.as<T>();
  T* no_y2 = nullptr;
  const T* no_d2 = nullptr;
  auto al16 = [](const void* p) { return ((uintptr_t)p & 15u) == 0; };
  // R rows per thread where every vector is 16-byte aligned and R divides n
  const bool vec = get_option("jacobi_dia_vec", 1) != 0 && P.n % R == 0 && al16(r) && al16(yin) &&
                   al16(yout) && al16(dv) && (UNIT || al16(dg)) && (!y2 || (al16(y2) && al16(d2)));
  const int nb = vec ? ceil_div(P.n / R, 256) : ceil_div(P.n, 256);
  switch (P.dia.nd) {
#define LACE_DIA_CASE(D)                                                                                   \
  case D:                                                                                                  \
    if (vec && y2)                                                                                         \
      tr.launch(tri_jacobi_dia_vec_kernel<T, UNIT, D, true>, dim3(nb), dim3(256), P.n, P.dia.off, mk, dv,  \
                dg, r, yin, yout, y2, d2);                                                                 \
    else if (vec)                                                                                          \
      tr.launch(tri_jacobi_dia_vec_kernel<T, UNIT, D, false>, dim3(nb), dim3(256), P.n, P.dia.off, mk, dv, \
                dg, r, yin, yout, no_y2, no_d2);                                                           \
    else if (y2)                                                                                           \
      tr.launch(tri_jacobi_dia_kernel<T, UNIT, D, true>, dim3(nb), dim3(256), P.n, P.dia.off, mk, dv, dg,  \
                r, yin, yout, y2, d2);                                                                     \
    else                                                                                                   \
      tr.launch(tri_jacobi_dia_kernel<T, UNIT, D, false>, dim3(nb), dim3(256), P.n, P.dia.off, mk, dv,     \
                dg, r, yin, yout, no_y2, no_d2);                                                           \
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

// The diagonal-form sweep on the 16-bit words of P (float vectors; four
// rows per thread).  Returns false (launching nothing) where the shape does
// not allow it: the caller then runs the FP32 form, which gives the same
// result.
template <typename TS, bool UNIT>
static bool jacobi_dia16(const TriPlan& P, const float* r, const float* yin, float* yout, float* y2,
                         const float* d2, TimedRegion& tr) {
  auto al16 = [](const void* p) { return ((uintptr_t)p & 15u) == 0; };
  if (P.n % 4 != 0 || !al16(r) || !al16(yin) || !al16(yout) || (y2 && (!al16(y2) || !al16(d2)))) return false;
  const uint8_t* mk = P.dia.mask.as<uint8_t>();
  const int8_t* ex = P.dexp.as<int8_t>();
  const TS* dv = P.dia16.as<TS>();
  const TS* dg = UNIT ? nullptr : P.dval16.as<TS>();
  float* no_y2 = nullptr;
  const float* no_d2 = nullptr;
  const int nb = ceil_div(P.n / 4, 256);
  switch (P.dia.nd) {
#define LACE_DIA16_CASE(D)                                                                                   \
  case D:                                                                                                    \
    if (y2)                                                                                                  \
      tr.launch(tri_jacobi_dia16_vec_kernel<TS, UNIT, D, true>, dim3(nb), dim3(256), P.n, P.dia.off, mk, ex, \
                dv, dg, r, yin, yout, y2, d2);                                                               \
    else                                                                                                     \
      tr.launch(tri_jacobi_dia16_vec_kernel<TS, UNIT, D, false>, dim3(nb), dim3(256), P.n, P.dia.off, mk,    \
                ex, dv, dg, r, yin, yout, no_y2, no_d2);                                                     \
    break;
    LACE_DIA16_CASE(1)
    LACE_DIA16_CASE(2)
    LACE_DIA16_CASE(3)
    LACE_DIA16_CASE(4)
    LACE_DIA16_CASE(5)
    LACE_DIA16_CASE(6)
    LACE_DIA16_CASE(7)
    LACE_DIA16_CASE(8)
#undef LACE_DIA16_CASE
    default:
      return false;
  }
  return true;
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

// Register slots of the persistent exact solve (rows per block: 512 by
// default, 256 on request): 8, or 4 for factors whose rows have at most 4
// off-diagonal entries (a shorter poll of LDS per round).
constexpr int kPersistR = 8;
// Poll rounds before a persistent-solve lane gives up (a round takes a
// memory round trip, or up to trsv_exact_idle x 64 cycles of sleep when
// nothing moved): seconds, far beyond any legitimate wait.
static int g_persist_spin_limit = 1 << 22;

template <typename T, bool UPPER, int BS, int R>
static void persist_launch(int n, const int* row, const int* col, const T* val, const T* r, T* y,
                           unsigned* ticket, int* err, int wgs_per_cu, int poll_sleep, const int* rperm,
                           T* out) {
  static int occ = -1;
  if (occ < 0) {
    int o = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, tri_persist_kernel<T, UPPER, BS, R>, 2 * BS, 0) !=
            hipSuccess ||
        o < 1)
      o = 1;
    occ = o;
  }
  const int nblk = ceil_div(n, BS);
  const int per = wgs_per_cu > 0 ? std::min(wgs_per_cu, occ) : occ;
  const int grid = (int)std::min<long>((long)per * std::max(1, num_cus()), nblk);
  tri_persist_kernel<T, UPPER, BS, R><<<grid, 2 * BS, 0, stream()>>>(
      n, nblk, row, col, val, r, y, ticket, err, g_persist_spin_limit, poll_sleep, rperm, out,
      get_option("trsv_exact_idle", 8));
}

template <typename T>
static int persist_dispatch(int n, bool upper, const int* row, const int* col, const T* val, const T* r, T* y,
                            unsigned* ticket, int* err, bool init, int wgs_per_cu, int poll_sleep, int rows,
                            const int* rperm, T* out, int slots = kPersistR) {
  if (n <= 0) return DEV_SUCCESS;
  LACE_REQUIRE((const void*)r != (const void*)y, DEV_ERR_ILLEGAL_VALUE, "exact solve needs distinct r and y");
  LACE_REQUIRE(rows == 256 || rows == 512, DEV_ERR_ILLEGAL_VALUE, "exact solve: %d rows per block", rows);
  if (init) tri_persist_init_kernel<T><<<ceil_div(n, 256), 256, 0, stream()>>>(n, y, ticket);
#define LACE_PERSIST(UP, BS, R) \
  persist_launch<T, UP, BS, R>(n, row, col, val, r, y, ticket, err, wgs_per_cu, poll_sleep, rperm, out)
  if (slots == 4) {
    if (rows == 512) {
      if (upper) LACE_PERSIST(true, 512, 4);
      else LACE_PERSIST(false, 512, 4);
    } else {
      if (upper) LACE_PERSIST(true, 256, 4);
      else LACE_PERSIST(false, 256, 4);
    }
  } else {
    if (rows == 512) {
      if (upper) LACE_PERSIST(true, 512, kPersistR);
      else LACE_PERSIST(false, 512, kPersistR);
    } else {
      if (upper) LACE_PERSIST(true, 256, kPersistR);
      else LACE_PERSIST(false, 256, kPersistR);
    }
  }
#undef LACE_PERSIST
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

template <typename T>
int tri_solve_persist_raw(int n, bool upper, const int* row, const int* col, const T* val, const T* r, T* y,
                          unsigned* ticket, int* err, bool init, int wgs_per_cu, int poll_sleep, int rows) {
  return persist_dispatch<T>(n, upper, row, col, val, r, y, ticket, err, init, wgs_per_cu, poll_sleep, rows,
                             nullptr, nullptr);
}

// Exact solve with a plan: the persistent kernel (option "trsv_exact" 1,
// the default) with its blocks in box order where the plan has one (option
// "trsv_exact_tiles" 1), else blocks of consecutive rows, or the
// level-ordered sync-free kernel (0).
template <typename T>
int tri_solve_exact(const TriPlan& P, const T* val, const T* r, T* y, unsigned* ticket, int* err) {
  if (get_option("trsv_exact", 1) == 0) return tri_solve_syncfree<T>(P, val, r, y, err);
  const double es = sizeof(T);
  const int wpc = get_option("trsv_exact_wgs_per_cu", 0), sleep = get_option("trsv_exact_poll_sleep", 0),
            rows = get_option("trsv_exact_rows", 512);
  const int slots = P.maxoff <= 4 ? 4 : kPersistR;
  if (P.tile_dims[0] > 0 && P.n > 0 && get_option("trsv_exact_tiles", 1) != 0) {
    // box order: values gathered once per change of the factor values
    if (P.tval_src != (const void*)val) {
      if (P.tval.bytes < sizeof(T) * std::max(P.nnz, 1)) LACE_TRY(P.tval.alloc(sizeof(T) * std::max(P.nnz, 1)));
      if (P.nnz > 0) {
        gather_kernel<T><<<ceil_div(P.nnz, 256), 256, 0, stream()>>>(P.nnz, P.tmap.as<int>(), val, P.tval.as<T>());
        LACE_HIP(hipGetLastError());
      }
      P.tval_src = val;
    }
    if (P.tsol.bytes < sizeof(T) * P.n) LACE_TRY(P.tsol.alloc(sizeof(T) * P.n));
    // + the permutation read and the scattered copy of the solution
    const double bytes = (double)P.nnz * (es + 4.0) + 4.0 * (P.n + 1) + (4.0 + 4.0 * es) * P.n;
    TimedRegion tr(P.upper ? "trsv_exact_upper" : "trsv_exact_lower", 0.0, bytes);
    return persist_dispatch<T>(P.n, P.upper, P.trow.as<int>(), P.tcol.as<int>(), P.tval.as<T>(), r,
                               P.tsol.as<T>(), ticket, err, true, wpc, sleep, rows, P.tperm.as<int>(), y, slots);
  }
  const double bytes = (double)P.nnz * (es + 4.0) + 4.0 * (P.n + 1) + 3.0 * es * P.n;
  TimedRegion tr(P.upper ? "trsv_exact_upper" : "trsv_exact_lower", 0.0, bytes);
  return persist_dispatch<T>(P.n, P.upper, P.row.as<int>(), P.col.as<int>(), val, r, y, ticket, err, true, wpc,
                             sleep, rows, nullptr, nullptr, slots);
}

template int tri_solve_persist_raw<double>(int, bool, const int*, const int*, const double*, const double*,
                                           double*, unsigned*, int*, bool, int, int, int);
template int tri_solve_persist_raw<float>(int, bool, const int*, const int*, const float*, const float*, float*,
                                          unsigned*, int*, bool, int, int, int);

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
    // 16-bit words of the diagonal form where the factors are stored in 16
    // bits (four rows per thread; it needs n % 4 == 0 and aligned vectors)
    const bool d16 = dia && sizeof(T) == 4 && P.dia16_fmt != LACE_F32 && P.dia16_src == (const void*)val &&
                     P.n % 4 == 0 && get_option("jacobi_dia16", 1) != 0;
    // bytes the chosen form reads once each, plus r, the previous iterate
    // and the new one (and the diagonal unless it is a unit one)
    const double es = sizeof(T), n = P.n;
    const double diag = P.unit_diag ? 0.0 : (d16 ? 2.0 : es) * n;
    double bytes = 3.0 * es * n + (fuse ? 2.0 * es * n : 0.0);
    if (d16) bytes += 2.0 * P.dia.nd * n + 2.0 * n + diag;
    else if (dia) bytes += P.dia.bytes(sizeof(T)) + diag;
    else if (ell) bytes += (double)P.ew * n * (es + 4.0) + diag;
    else if (strict) bytes += (double)P.snnz * (es + 4.0) + 4.0 * (n + 1) + diag;
    else bytes += (double)P.nnz * (es + 4.0) + 4.0 * (n + 1);
    // the diagonal-form launch carries the events itself
    TimedRegion tr(P.upper ? "trsv_jacobi_sweep_upper" : "trsv_jacobi_sweep_lower", 0.0, bytes, dia);
    bool done16 = false;
    if constexpr (sizeof(T) == 4) {
      if (d16) {
        float* wy2 = fuse ? (float*)y2 : nullptr;
        if (P.dia16_fmt == LACE_BF16)
          done16 = P.unit_diag ? jacobi_dia16<uint16_t, true>(P, r, yin, yout, wy2, d2, tr)
                               : jacobi_dia16<uint16_t, false>(P, r, yin, yout, wy2, d2, tr);
        else
          done16 = P.unit_diag ? jacobi_dia16<_Float16, true>(P, r, yin, yout, wy2, d2, tr)
                               : jacobi_dia16<_Float16, false>(P, r, yin, yout, wy2, d2, tr);
      }
    }
    if (done16) {
      // launched above
    } else if (dia && P.unit_diag) jacobi_dia<T, true>(P, r, yin, yout, fuse ? y2 : nullptr, d2, tr);
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

// Tile count, dependency reach in tiles, and whether the pipelined sweeps
// apply to P with the value array val: diagonal form current, every offset
// on the factor's side of the diagonal, n a multiple of the 16-byte word,
// reach within one wavefront of flags.
template <typename T>
static bool jacobi_pipe_shape(const TriPlan& P, const void* val, int& ntiles, int& K) {
  constexpr int R = Pack16<T>::R;
  constexpr int TR = 256 * R;
  if (P.dia.nd <= 0 || P.strict_src != val || P.n <= 0 || P.n % R != 0) return false;
  int reach = 0;
  for (int q = 0; q < P.dia.nd; q++) {
    const int o = P.dia.off.o[q];
    if (o == 0 || (o > 0) != P.upper) return false;
    reach = std::max(reach, std::abs(o));
  }
  ntiles = ceil_div(P.n, TR);
  K = ceil_div(reach, TR);
  return K <= 64;
}

// Whether S sweeps of P from yin can run pipelined: the shape above, S in
// 2..32, and every vector 16-byte aligned.
template <typename T>
static bool jacobi_pipe_ready(const TriPlan& P, const void* val, const T* r, const T* yin, const T* yout,
                              int S, const T* y2, const T* d2) {
  int ntiles = 0, K = 0;
  if (S < 2 || S > 32 || !jacobi_pipe_shape<T>(P, val, ntiles, K)) return false;
  auto al16 = [](const void* p) { return ((uintptr_t)p & 15u) == 0; };
  return al16(r) && al16(yin) && al16(yout) && al16(P.dia.val.p) && (P.unit_diag || al16(P.dval.p)) &&
         (!y2 || (al16(y2) && al16(d2)));
}

// S sweeps of P from yin in one launch (tri_jacobi_dia_pipe_kernel); yout
// gets sweep S and, with y2, y2 = yout / d2.  jacobi_pipe_ready must hold.
template <typename T, bool UNIT>
static int jacobi_dia_pipe(const TriPlan& P, const T* val, const T* r, const T* yin, T* yout, int S, T* y2,
                           const T* d2, int* err) {
  int ntiles = 0, K = 0;
  LACE_REQUIRE(jacobi_pipe_ready<T>(P, val, r, yin, yout, S, y2, d2) && jacobi_pipe_shape<T>(P, val, ntiles, K),
               DEV_ERR_ILLEGAL_VALUE, "pipelined Jacobi: shape not supported");
  const T* dv = P.dia.val.as<T>();
  const T* dg = P.dval.as<T>();
  const size_t need = sizeof(T) * (size_t)(S - 1) * P.n;
  if (P.pipe_buf.bytes < need) LACE_TRY(P.pipe_buf.alloc(need));
  if (!P.pipe_flags.p || P.pipe_flags.bytes < sizeof(unsigned) * ntiles) {
    LACE_TRY(P.pipe_flags.alloc(sizeof(unsigned) * ntiles));
    LACE_TRY(P.pipe_ctr.alloc(sizeof(unsigned)));
    LACE_HIP(hipMemsetAsync(P.pipe_flags.p, 0, sizeof(unsigned) * ntiles, stream()));
    LACE_HIP(hipMemsetAsync(P.pipe_ctr.p, 0, sizeof(unsigned), stream()));
    P.pipe_tickets = 0;
    P.pipe_epoch = 0;
  }
  // flag values of this launch: base + 1 .. base + S - 1, above every value
  // an earlier launch left
  P.pipe_epoch += 64;
  const unsigned tb = P.pipe_tickets, fb = P.pipe_epoch;
  P.pipe_tickets += (unsigned)ntiles;
  // algorithmic bytes: the factor's diagonals and masks (and its diagonal),
  // r and yin once, each intermediate sweep written and read once, yout
  // (and the epilogue's d2 and y2)
  const double es = sizeof(T), n = P.n;
  double bytes = P.dia.bytes(sizeof(T)) + (UNIT ? 0.0 : es * n) + 2.0 * es * n + 2.0 * es * n * (S - 1) +
                 es * n + (y2 ? 2.0 * es * n : 0.0);
  TimedRegion tr(P.upper ? "trsv_jacobi_pipe_upper" : "trsv_jacobi_pipe_lower", 0.0, bytes, true);
  const uint8_t* mk = P.dia.mask.as<uint8_t>();
  T* bufs = P.pipe_buf.as<T>();
  unsigned* fl = P.pipe_flags.as<unsigned>();
  unsigned* ct = P.pipe_ctr.as<unsigned>();
  T* no_y2 = nullptr;
  const T* no_d2 = nullptr;
  switch (P.dia.nd) {
#define LACE_PIPE_CASE(D)                                                                                  \
  case D:                                                                                                  \
    if (y2)                                                                                                \
      tr.launch(tri_jacobi_dia_pipe_kernel<T, UNIT, D, true>, dim3(ntiles), dim3(256), P.n, P.dia.off,     \
                ntiles, K, (int)P.upper, S, mk, dv, dg, r, yin, bufs, yout, y2, d2, fl, ct, tb, fb, err,   \
                g_spin_limit);                                                                             \
    else                                                                                                   \
      tr.launch(tri_jacobi_dia_pipe_kernel<T, UNIT, D, false>, dim3(ntiles), dim3(256), P.n, P.dia.off,    \
                ntiles, K, (int)P.upper, S, mk, dv, dg, r, yin, bufs, yout, no_y2, no_d2, fl, ct, tb, fb,  \
                err, g_spin_limit);                                                                        \
    break;
    LACE_PIPE_CASE(1)
    LACE_PIPE_CASE(2)
    LACE_PIPE_CASE(3)
    LACE_PIPE_CASE(4)
    LACE_PIPE_CASE(5)
    LACE_PIPE_CASE(6)
    LACE_PIPE_CASE(7)
    LACE_PIPE_CASE(8)
#undef LACE_PIPE_CASE
    default:
      break;
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
  if (prec == LACE_F64)
    gather_kernel<double><<<ceil_div(nnzU, 256), 256, 0, stream()>>>(
        nnzU, u_csr_from_csc.as<int>(), uval[cur].as<double>(), uval_csr.as<double>());
  else
    gather_kernel<float><<<ceil_div(nnzU, 256), 256, 0, stream()>>>(
        nnzU, u_csr_from_csc.as<int>(), uval[cur].as<float>(), uval_csr.as<float>());
  LACE_HIP(hipGetLastError());
  lval_t_dirty = true;  // gathered by the first transposed solve
  // off-diagonal forms for the Jacobi sweeps, from the arrays the solves
  // pass (gathered on first use)
  const size_t es = elem_size();
  LACE_TRY(lower.refresh_strict(lval[cur].p, es));
  LACE_TRY(upper.refresh_strict(uval_csr.p, es));
  LACE_TRY(lowerT.refresh_strict(uval[cur].p, es));
  LACE_TRY(upperT.refresh_strict(lval_t.p, es));
  if (s16()) {
    // the non-transposed Jacobi sweeps read the 16-bit words (a row of L or
    // U shares its row's scale; the transposed factors do not)
    LACE_TRY(lower.refresh_dia16(prec, rs.as<float>()));
    LACE_TRY(upper.refresh_dia16(prec, rs.as<float>()));
  }
  csr_dirty = false;
  return DEV_SUCCESS;
}

// The off-diagonal forms the Jacobi sweeps read on P under the present
// options: the diagonal always (the first sweep from zero divides by it),
// the diagonal form where there is one (the one-sweep, fused and pipelined
// kernels), else the ELLPACK form, else the off-diagonal CSR; the 32-column
// sweeps read the diagonal form, else the ELLPACK form, else the full CSR.
static unsigned jacobi_forms(const TriPlan& P, bool block) {
  unsigned m = TriPlan::kFormDiag;
  const bool ell = P.ew > 0 && get_option("jacobi_ell", 1) != 0;
  if (block) {
    if (P.dia.nd > 0 && get_option("jacobi_dia", 1) != 0) return m | TriPlan::kFormDia;
    return ell ? m | TriPlan::kFormEll : m;
  }
  if (P.dia.nd > 0) m |= TriPlan::kFormDia;
  if (!(P.dia.nd > 0 && get_option("jacobi_dia", 1) != 0)) m |= ell ? TriPlan::kFormEll : TriPlan::kFormStrict;
  return m;
}

int ParILU::prepare(lace_trsv_kind_t kind, int trans, bool block) {
  LACE_TRY(refresh_csr());
  if (kind == LACE_TRSV_NONE) return DEV_SUCCESS;
  if (trans && lval_t_dirty) {
    if (prec == LACE_F64)
      gather_kernel<double><<<ceil_div(nnzL, 256), 256, 0, stream()>>>(
          nnzL, l_t_from_csr.as<int>(), lval[cur].as<double>(), lval_t.as<double>());
    else
      gather_kernel<float><<<ceil_div(nnzL, 256), 256, 0, stream()>>>(
          nnzL, l_t_from_csr.as<int>(), lval[cur].as<float>(), lval_t.as<float>());
    LACE_HIP(hipGetLastError());
    lval_t_dirty = false;
  }
  if (kind == LACE_TRSV_JACOBI) {
    TriPlan& a = trans ? lowerT : lower;
    TriPlan& b = trans ? upperT : upper;
    LACE_TRY(a.ensure_forms(jacobi_forms(a, block)));
    LACE_TRY(b.ensure_forms(jacobi_forms(b, block)));
  }
  return DEV_SUCCESS;
}

template <typename T>
static int apply_t(ParILU* P, lace_trsv_kind_t kind, int sweeps, int trans, const T* r, T* z) {
  LACE_TRY(P->prepare(kind, trans, false));
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
    unsigned* ticket = reinterpret_cast<unsigned*>(P->syncfree_err.as<int>() + 2);
    LACE_TRY(tri_solve_exact<T>(first, v1, r, tmp, ticket, err));
    LACE_TRY(tri_solve_exact<T>(second, v2, tmp, z, ticket, err));
    return DEV_SUCCESS;
  }
  // Jacobi: work holds the intermediate vector, work2 the ping-pong buffer.
  T* tmp2 = P->work2.as<T>();
  // Pipelined form, on request (unit lower factor, then a non-unit upper
  // one, both on the diagonal form): the first solve's sweeps 2..k from
  // y1 = r in one launch, whose last sweep also writes the second solve's
  // first sweep tmp / diag into work2; then the second solve's sweeps 2..k
  // from work2 into z in one launch.  Off by default: on the 128^3 FP32
  // factors one such launch takes 75 us against 45 us for the four
  // one-sweep launches it replaces (each tile's sweeps wait on a chain of
  // device-coherent store, flag and load round trips).
  const T* d2 = second.dval.as<T>();
  if (sweeps >= 3 && first.unit_diag && !second.unit_diag && get_option("jacobi_pipe", 0) != 0 &&
      jacobi_pipe_ready<T>(first, v1, r, r, tmp, sweeps - 1, tmp2, d2) &&
      jacobi_pipe_ready<T>(second, v2, tmp, tmp2, z, sweeps - 1, nullptr, nullptr)) {
    int* err = P->syncfree_err.as<int>() + 1;
    LACE_TRY((jacobi_dia_pipe<T, true>(first, v1, r, r, tmp, sweeps - 1, tmp2, d2, err)));
    LACE_TRY((jacobi_dia_pipe<T, false>(second, v2, tmp, tmp2, z, sweeps - 1, nullptr, nullptr, err)));
    return DEV_SUCCESS;
  }
  // For an odd sweep count the second solve's first sweep (tmp / diag,
  // written to z) rides on the first solve's last sweep when that runs on
  // the diagonal form; z is then distinct from every buffer that sweep reads.
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
  // slot 0: the sync-free solve, slot 1: the pipelined Jacobi sweeps
  int e[2] = {0, 0};
  LACE_HIP(hipMemcpyAsync(e, P->syncfree_err.p, sizeof(e), hipMemcpyDeviceToHost, stream()));
  LACE_HIP(hipStreamSynchronize(stream()));
  *err = e[0] | e[1];
  return DEV_SUCCESS;
}

}  // extern "C"

namespace {

// The reference loop (src/trisolve.cpp:20-119,256-351) sweeps in place, in
// natural row order for L and reverse order for U, until the sum of squared
// changes is <= tol.  Run by one thread, which is the only schedule with a
// defined result, its first sweep is exact forward / backward substitution
// and its second sweep changes nothing (step 0), so it returns the exact
// solution with iter = 1 when the first step is already <= tol and iter = 2
// otherwise (and no sweep at all when tol >= 1e8, the loop's initial step).
// Here the exact solve runs on the device (level-ordered,
// bit-identical to substitution) and the first step, sum (y0_i - y_i)^2 in
// the reference's row order, is formed on the host from the caller's y0.
// check_finite: data_parcsrtrsv also stops on a non-finite step (:274,317);
// the two solve wrappers do not.  log: the wrappers print each sweep's
// "% iteration = k step = s" line as the reference does (:57,104).
int exact_parcsrtrsv(bool upper, int n, const double* Aval, const int* row, const int* col,
                     const double* rhs, double* y, double tol, bool check_finite, bool log, int* iter) {
  LACE_REQUIRE(n >= 0 && (n == 0 || (Aval && row && col && rhs && y && iter)), DEV_ERR_INVALID_PTR,
               "parcsrtrsv: null argument");
  // the loop starts from step = 1e8: a tol at or above it (or NaN) runs no
  // sweep and leaves y as it is
  if (!(1.e8 > tol)) {
    if (iter) *iter = 0;
    return 0;
  }
  if (n == 0) {
    // an empty sweep has step 0 <= tol
    if (iter) *iter = 1;
    return 0;
  }
  for (int i = 0; i < n; i++) {
    const int kd = upper ? row[i] : row[i + 1] - 1;
    LACE_REQUIRE(row[i + 1] > row[i] && col[kd] == i, DEV_ERR_ILLEGAL_VALUE,
                 "parcsrtrsv: row %d has no diagonal in the %s slot", i, upper ? "first" : "last");
  }
  const int nnz = row[n];
  DevBuf val, drow, dcol, r, yd, err;
  LACE_TRY(upload(val, Aval, sizeof(double) * std::max(nnz, 1)));
  LACE_TRY(upload(drow, row, sizeof(int) * (n + 1)));
  LACE_TRY(upload(dcol, col, sizeof(int) * std::max(nnz, 1)));
  LACE_TRY(upload(r, rhs, sizeof(double) * n));
  LACE_TRY(yd.alloc(sizeof(double) * n));
  // [0] the solve gave up waiting, [1] the block ticket
  LACE_TRY(err.alloc(sizeof(int) * 2));
  LACE_HIP(hipMemsetAsync(err.p, 0, sizeof(int) * 2, stream()));
  LACE_TRY(tri_solve_persist_raw<double>(n, upper, drow.as<int>(), dcol.as<int>(), val.as<double>(),
                                         r.as<double>(), yd.as<double>(),
                                         reinterpret_cast<unsigned*>(err.as<int>() + 1), err.as<int>()));
  std::vector<double> ynew(n);
  int e = 0;
  LACE_HIP(hipMemcpyAsync(ynew.data(), yd.p, sizeof(double) * n, hipMemcpyDeviceToHost, stream()));
  LACE_HIP(hipMemcpyAsync(&e, err.p, sizeof(int), hipMemcpyDeviceToHost, stream()));
  LACE_HIP(hipStreamSynchronize(stream()));
  LACE_REQUIRE(e == 0, DEV_ERR, "parcsrtrsv: the device solve gave up waiting on a row");
  double step = 0.0;
  for (int t = 0; t < n; t++) {
    const int i = upper ? n - 1 - t : t;
    const double d = y[i] - ynew[i];
    step += d * d;
  }
  memcpy(y, ynew.data(), sizeof(double) * n);
  const bool again = check_finite ? (step > tol && std::isfinite(step)) : step > tol;
  if (log) printf("%% iteration = %d step = %e\n", 1, step);
  if (again) {
    // the confirming sweep: every row recomputes the value it holds, so the
    // change is 0, or NaN where a value is not finite
    double step2 = 0.0;
    for (int t = 0; t < n; t++) {
      const int i = upper ? n - 1 - t : t;
      const double d = ynew[i] - ynew[i];
      step2 += d * d;
    }
    if (log) printf("%% iteration = %d step = %e\n", 2, step2);
  }
  *iter = again ? 2 : 1;
  return 0;
}

}  // namespace

extern "C" {

// src/trisolve.cpp:256-351.  storage/uplo must name CSRL+Lower or
// CSRU+Upper as in the reference (:264-347); info -1 otherwise.
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
  return exact_parcsrtrsv(upper, num_rows, Aval, row, col, rhsval, yval, tol, true, false, iter);
}

// src/trisolve.cpp:20-67
int data_forward_solve(data_d_matrix* L, data_d_matrix* x, data_d_matrix* rhs, const dataType tol,
                       int* iter) {
  LACE_TRY(ensure_init());
  if (!(L->storage_type == Magma_CSRL && L->fill_mode == MagmaLower &&
        L->diagorder_type != Magma_NODIAG)) {
    set_error("data_forward_solve: L must be CSRL / lower with a diagonal");
    return -1;
  }
  return exact_parcsrtrsv(false, L->num_rows, L->val, L->row, L->col, rhs->val, x->val, tol, false,
                          get_option("matlab_log", 1) != 0, iter);
}

// src/trisolve.cpp:72-119
int data_backward_solve(data_d_matrix* U, data_d_matrix* x, data_d_matrix* rhs,
                        const dataType tol, int* iter) {
  LACE_TRY(ensure_init());
  if (!(U->storage_type == Magma_CSRU && U->fill_mode == MagmaUpper &&
        U->diagorder_type != Magma_NODIAG)) {
    set_error("data_backward_solve: U must be CSRU / upper with a diagonal");
    return -1;
  }
  return exact_parcsrtrsv(true, U->num_rows, U->val, U->row, U->col, rhs->val, x->val, tol, false,
                          get_option("matlab_log", 1) != 0, iter);
}

}  // extern "C"
