// krylov.hip -- restarted Householder FGMRES on the GPU.
//
// Reference: src/fgmres_householder_restart.cpp:63-788 (Walker's
// Householder Arnoldi with right flexible preconditioning), src/givens.cpp.
//
// Algorithm per restart cycle (n rows, restart length m):
//   r = b - A x, first reflector u_0 from r (:239-308), q = -r/dd, g_0 = -dd
//   for j = 0..m-1:
//     z_j = M^{-1} q                                  (:320-349)
//     w   = A z_j                                     (:356-367)
//     w   = P_j ... P_0 w                             (:375-469)
//     u_{j+1} from w[j+1:], h_{j+1,j} = -dd           (:476-517)
//     q   = P_0 ... P_{j+1} e_{j+1}                   (:519-616)
//     rotate h by the previous Givens rotations, form the new one (:658-701)
//     stop when |g_{j+1}| < rtol                      (:714)
//   y = R^{-1} g, x += Z y                            (:716-745)
//
// MI355X design: the reflector products are applied in compact-WY form,
// P_0 P_1 ... P_j = I - U T U^T with T upper triangular
// (T_{j+1} = [[T_j, -2 T_j U_j^T u_{j+1}], [0, 2]]), so each Arnoldi step
// streams the reflector block U three times (multi-dot; the update fused
// with the multi-dot of the tail; new column + next basis vector) instead
// of 2(j+1) dot/axpy passes.  All O(j^2) work (T, Hessenberg, Givens, least squares) stays on
// the device in one-workgroup kernels; the host only reads the convergence
// flag.  Every kernel after the first of a cycle returns at once when the
// device-side done flag is set, so a caller may queue steps ahead.
#include <cmath>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "krylov.hpp"

namespace lace {

constexpr int kBS = 256;

__device__ __forceinline__ bool is_done(const double* misc) { return misc[MS_DONE] != 0.0; }

// ------------------------------------------------------------ multi-dot
// partial[b][k] = sum_{i in block b, i >= rowstart} U[i,k] * w[i], k < ncols;
// with SELF also partial[b][pstride-1] = sum w[i]^2.
// Sum of NC per-lane values over the 64 lanes of a wave, NC - 1 shuffles
// instead of NC full reductions: stage s pairs lanes across bit 32>>s and
// each lane keeps (and completes) half of its remaining values.  Lanes whose
// low (6 - log2 NC) bits are zero end up holding the total of value index
// idx(lane); the function stores it to out[idx].  Fixed order: deterministic.
template <int NC, typename T>
__device__ __forceinline__ void wave_multi_sum(T (&v)[NC], double* out) {
  static_assert((NC & (NC - 1)) == 0 && NC <= 64, "NC must be a power of two <= 64");
  const int lane = threadIdx.x & 63;
  int idx = 0;
#pragma unroll
  for (int half = NC / 2, o = 32; half >= 1; half >>= 1, o >>= 1) {
    const bool hi = (lane & o) != 0;
#pragma unroll
    for (int i = 0; i < half; i++) {
      const T send = hi ? v[i] : v[i + half];
      const T keep = hi ? v[i + half] : v[i];
      v[i] = keep + __shfl_xor(send, o, 64);
    }
    if (hi) idx += half;
  }
  // the low lane bits not consumed above still split the sum
  constexpr int used = (NC == 1) ? 0 : (NC == 2 ? 1 : NC == 4 ? 2 : NC == 8 ? 3 : NC == 16 ? 4 : NC == 32 ? 5 : 6);
#pragma unroll
  for (int o = 32 >> used; o >= 1; o >>= 1) v[0] += __shfl_xor(v[0], o, 64);
  if ((lane & ((64 >> used) - 1)) == 0) out[idx] = (double)v[0];
}

// Multi-dot over one contiguous row slab per block (rows [b*rpb, (b+1)*rpb)
// intersected with [rowstart, n)).  The four waves of a block split the
// columns: wave q accumulates columns [q*NC/4, (q+1)*NC/4) for all rows of
// the slab (lanes stride the rows), so each lane holds only NC/4
// accumulators and a column's total needs one NC/4-value wave reduction.
// Totals go to partial[b][0..ncols) and, with SELF, sum w^2 (wave 0) to
// partial[b][pstride-1].
template <typename K, int NC, bool SELF>
__global__ __launch_bounds__(kBS) void mdot_kernel(int n, int ld, const K* __restrict__ U, int ncols,
                                                   const K* __restrict__ w, int rowstart, int rpb,
                                                   double* __restrict__ partial, int pstride,
                                                   const double* __restrict__ misc) {
  if (misc && is_done(misc)) return;
  constexpr int NW = kBS / 64;
  constexpr int NCW = NC / NW;  // columns per wave
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int r0 = max(b * rpb, rowstart), r1 = min((b + 1) * rpb, n);
  const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c0 = q * NCW;
  const K* Uq = U + (size_t)c0 * ld;
  K acc[NCW];
  K self = K(0);
#pragma unroll
  for (int k = 0; k < NCW; k++) acc[k] = K(0);
  if (c0 < ncols || (SELF && q == 0)) {
    for (int i = r0 + lane; i < r1; i += 64) {
      const K wi = w[i];
#pragma unroll
      for (int k = 0; k < NCW; k++)
        if (c0 + k < ncols) acc[k] += Uq[(size_t)k * ld + i] * wi;
      if (SELF && q == 0) self += wi * wi;
    }
  }
  __shared__ double red[NC + 1];
  wave_multi_sum<NCW>(acc, red + c0);
  if (SELF && q == 0) {
    const double sv = wave_sum((double)self);
    if (lane == 0) red[NC] = sv;
  }
  __syncthreads();
  for (int k = threadIdx.x; k <= NC; k += kBS) {
    if (k < ncols) partial[(size_t)b * pstride + k] = red[k];
    if (SELF && k == NC) partial[(size_t)b * pstride + pstride - 1] = red[NC];
  }
}

// The same multi-dot with the slab's rows split over the block's threads.
// A whole slab of R * 256 rows (every block but those cut by rowstart or
// the end when R > 0) keeps its w rows in registers, R per thread as 16-byte
// words of V consecutive rows (thread t: words t, t + 256, ...), and walks
// the columns one after another: each column's slab is one contiguous run
// read as 16-byte words, one column at a time per wave, and w is read once
// for all columns.  Measured on FP64 n = 2^21, 50 columns: 6.1 TB/s against
// 3.4 TB/s for rows outer and 16 columns inner with single-row loads.  Other
// slabs take the rows-outer form: thread t takes the groups of V rows t,
// t + 256, ... and loads each column's V rows and w's as one 16-byte word
// (single rows for FP64 chunks of 16 columns, whose two-row form would hold
// 32 doubles of U per thread), the columns in chunks of NC in registers
// with w's slab read again per chunk, from cache.  U, w and ld must be
// 16-byte aligned (the engine's buffers are); groups cut by rowstart or the
// slab end load element by element.
template <typename K, int NC, bool SELF, int R>
__global__ __launch_bounds__(kBS, 4) void mdot_rows_kernel(int n, int ld, const K* __restrict__ U, int ncols,
                                                        const K* __restrict__ w, int rowstart, int rpb,
                                                        double* __restrict__ partial, int pstride,
                                                        const double* __restrict__ misc) {
  if (misc && is_done(misc)) return;
  constexpr int V = (sizeof(K) == 8 && NC > 8) ? 1 : 16 / (int)sizeof(K);
  constexpr int VW = 16 / (int)sizeof(K);  // rows per 16-byte word
  constexpr int RW = R > 0 ? R / VW : 1;   // w words per thread on the register path
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int r0 = max(b * rpb, rowstart), r1 = min((b + 1) * rpb, n);
  const int wave = threadIdx.x >> 6;
  const bool whole = R > 0 && rpb == R * kBS && r0 == b * rpb && r1 == (b + 1) * rpb;
  K wr[RW][VW];
  if (whole) {
#pragma unroll
    for (int i = 0; i < RW; i++) vload<K, VW>(w + r0 + (size_t)(threadIdx.x + kBS * i) * VW, wr[i]);
  }
  __shared__ double red[kBS / 64][NC + 1];
  // at least one pass, for the SELF sum when there are no columns
  for (int c0 = 0; c0 < ncols || (c0 == 0 && SELF); c0 += NC) {
    const int nck = min(NC, ncols - c0);
    const bool self_pass = SELF && c0 == 0;
    K acc[NC];
    K self = K(0);
#pragma unroll
    for (int k = 0; k < NC; k++) acc[k] = K(0);
    if (whole) {
#pragma unroll
      for (int k = 0; k < NC; k++) {
        if (k < nck) {
          const K* col = U + (size_t)(c0 + k) * ld + r0;
#pragma unroll
          for (int i = 0; i < RW; i++) {
            K uv[VW];
            vload<K, VW>(col + (size_t)(threadIdx.x + kBS * i) * VW, uv);
#pragma unroll
            for (int e = 0; e < VW; e++) acc[k] += uv[e] * wr[i][e];
          }
        }
      }
      if (self_pass) {
#pragma unroll
        for (int i = 0; i < RW; i++)
#pragma unroll
          for (int e = 0; e < VW; e++) self += wr[i][e] * wr[i][e];
      }
    } else if (r0 < r1) {
      const int g1 = (r1 + V - 1) / V;
      for (int g = r0 / V + (int)threadIdx.x; g < g1; g += kBS) {
        const int i0 = g * V;
        const bool full = i0 >= r0 && i0 + V <= r1;
        K wv[V];
        if (full) {
          wload<K, V>(w + i0, wv);
        } else {
#pragma unroll
          for (int e = 0; e < V; e++) wv[e] = (i0 + e >= r0 && i0 + e < r1) ? w[i0 + e] : K(0);
        }
#pragma unroll
        for (int k = 0; k < NC; k++) {
          if (k < nck) {
            const K* col = U + (size_t)(c0 + k) * ld;
            K uv[V];
            if (full) {
              wload<K, V>(col + i0, uv);
            } else {
#pragma unroll
              for (int e = 0; e < V; e++) uv[e] = (i0 + e >= r0 && i0 + e < r1) ? col[i0 + e] : K(0);
            }
#pragma unroll
            for (int e = 0; e < V; e++) acc[k] += uv[e] * wv[e];
          }
        }
        if (self_pass) {
#pragma unroll
          for (int e = 0; e < V; e++) self += wv[e] * wv[e];
        }
      }
    }
    wave_multi_sum<NC>(acc, red[wave]);
    if (self_pass) {
      const double sv = wave_sum((double)self);
      if ((threadIdx.x & 63) == 0) red[wave][NC] = sv;
    }
    __syncthreads();
    for (int k = threadIdx.x; k <= NC; k += kBS) {
      if (k < nck)
        partial[(size_t)b * pstride + c0 + k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
      if (self_pass && k == NC)
        partial[(size_t)b * pstride + pstride - 1] = (red[0][NC] + red[1][NC]) + (red[2][NC] + red[3][NC]);
    }
    __syncthreads();
  }
}

// Launches the multi-dot: the register path for whole slabs of 1024 or
// 2048 rows (8 columns per pass), the rows-outer form otherwise (NC columns
// per pass).
// The step's first multi-dot (U_j^T w, no self term) is the timing region
// "hh_mdot" with its algorithmic bytes, every column and w once.
template <typename K, int NC, bool SELF>
static void launch_mdot_rows(const HHEngine<K>& E, const K* U, int ncols, const K* w, int rowstart,
                             double* part, const double* misc) {
  TimedRegion tr(SELF ? "hh_mdot_self" : "hh_mdot", 0.0,
                 (double)sizeof(K) * (double)(E.n - rowstart) * (double)(ncols + 1), true);
  if (E.rpb == 8 * kBS)
    tr.launch(mdot_rows_kernel<K, 8, SELF, 8>, dim3(E.nb), dim3(kBS), E.n, E.ld, U, ncols, w, rowstart, E.rpb,
              part, E.pstride, misc);
  else if (E.rpb == 4 * kBS)
    tr.launch(mdot_rows_kernel<K, 8, SELF, 4>, dim3(E.nb), dim3(kBS), E.n, E.ld, U, ncols, w, rowstart, E.rpb,
              part, E.pstride, misc);
  else
    tr.launch(mdot_rows_kernel<K, NC, SELF, 0>, dim3(E.nb), dim3(kBS), E.n, E.ld, U, ncols, w, rowstart, E.rpb,
              part, E.pstride, misc);
}

// Column totals of the per-block partials when there are many columns: one
// workgroup per column c < cnt (and one more for the self slot when self),
// each summing partial[b][c] over the blocks in a fixed order.  The totals
// land in tot with the partial row layout (self at pstride - 1), so the
// one-workgroup kernels read them as the partials of a single block instead
// of reducing nb rows of many columns on one workgroup.
__global__ __launch_bounds__(256) void reduce_cols_kernel(int nb, int cnt, int self, int pstride,
                                                          const double* __restrict__ partial,
                                                          double* __restrict__ tot,
                                                          const double* __restrict__ misc) {
  if (misc && is_done(misc)) return;
  const int c = blockIdx.x;
  const int col = (self && c == cnt) ? pstride - 1 : c;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += partial[(size_t)b * pstride + col];
  __shared__ double sm[256 / 64];
  const double t = block_sum<256>(s, sm);
  if (threadIdx.x == 0) tot[col] = t;
}

// Sum of the per-block partials in a fixed order: out[k] for k < cnt and,
// if self, out[cnt] = the self slot.  One block of 256 threads, ch columns
// at a time (ch: the column count rounded up to a power of two, at most
// 16): lane t reads column t % ch of the partial rows t / ch, t / ch +
// 256 / ch, ..., so the lanes of a row read its consecutive words and a
// round touches each cache line once, and a thread issues the loads of 64
// of its rows before adding any of them, so the 1024 rows of a 2M-row
// vector's 2048-row slabs take one wait for memory.  The row groups' sums
// then meet in a fixed butterfly per wave and a fixed 4-wave combination.
__device__ __forceinline__ void finalize_partials(int nb, int cnt, bool self, int pstride,
                                                  const double* __restrict__ partial, double* out) {
  constexpr int RB = 64;  // rows in flight per thread
  __shared__ double seg[kBS / 64][16];
  const int tot = cnt + (self ? 1 : 0);
  int ch = 1;
  while (ch < tot && ch < 16) ch <<= 1;
  const int G = kBS / ch;  // row groups
  const int k = threadIdx.x & (ch - 1), g = threadIdx.x / ch;
  const int wv = threadIdx.x >> 6;
  for (int base = 0; base < tot; base += ch) {
    const int c = base + k;
    const unsigned col = (self && c == cnt) ? pstride - 1 : c;
    double acc = 0.0;
    if (c < tot) {
      for (int b0 = g; b0 < nb; b0 += RB * G) {
        double v[RB];
#pragma unroll
        for (int q = 0; q < RB; q++) {
          v[q] = 0.0;
          if (b0 + q * G < nb) v[q] = partial[(unsigned)(b0 + q * G) * (unsigned)pstride + col];
        }
#pragma unroll
        for (int q = 0; q < RB; q++)
          if (b0 + q * G < nb) acc += v[q];
      }
    }
    for (int off = ch; off < 64; off <<= 1) acc += __shfl_xor(acc, off);
    if ((int)(threadIdx.x & 63) < ch) seg[wv][k] = acc;
    __syncthreads();
    if ((int)threadIdx.x < ch && base + (int)threadIdx.x < tot)
      out[base + threadIdx.x] = (seg[0][threadIdx.x] + seg[1][threadIdx.x]) +
                                (seg[2][threadIdx.x] + seg[3][threadIdx.x]);
    __syncthreads();
  }
}

// ------------------------------------------------------------ cycle start
// After the norm pass over r: rnorm, dd = sgn(r_0)*rnorm, scale of u_0,
// g_0 = -dd, T_00 = 2; done if rnorm < rtol or not finite.
template <typename K>
__global__ __launch_bounds__(256) void hh_start_kernel(int nb, int pstride, const double* partial,
                                                       const K* __restrict__ r, double* small,
                                                       SmallLayout lay, int set_rtol, double rtol_rel) {
  double* misc = small + lay.misc();
  __shared__ double nu[1];
  finalize_partials(nb, 0, true, pstride, partial, nu);
  __syncthreads();
  if (threadIdx.x == 0) {
    const double rnorm = sqrt(nu[0]);
    misc[MS_RNORM] = rnorm;
    if (set_rtol) misc[MS_RTOL] = rtol_rel * rnorm;
    const double r0 = (double)r[0];
    const double sg = (r0 >= 0.0) ? 1.0 : -1.0;  // mysgn, include/sparse.h:1131
    const double dd = sg * rnorm;
    const double nv = sqrt(2.0 * rnorm * (rnorm + fabs(r0)));
    misc[MS_DD] = dd;
    misc[MS_SCALE] = nv > 0.0 ? 1.0 / nv : 0.0;
    misc[MS_NCOLS] = 0.0;
    misc[MS_BREAK] = 0.0;
    misc[MS_RES] = rnorm;
    // a zero residual is solved (and has no reflector: dd would be 0)
    const bool done = !(rnorm >= misc[MS_RTOL]) || !isfinite(rnorm) || rnorm == 0.0;
    misc[MS_DONE] = done ? 1.0 : 0.0;
    double* g = small + lay.g();
    for (int t = 0; t < lay.m + 2; t++) g[t] = 0.0;
    g[0] = -dd;
    small[lay.T()] = 2.0;
  }
}

// u_0 = (r + dd e_0) * scale, q = -r / dd.
template <typename K>
__global__ __launch_bounds__(kBS) void hh_start_vec_kernel(int n, const K* __restrict__ r,
                                                           K* __restrict__ U0, K* __restrict__ q,
                                                           const double* __restrict__ misc) {
  if (is_done(misc)) return;
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  const double dd = misc[MS_DD], sc = misc[MS_SCALE];
  const double ri = (double)r[i];
  U0[i] = (K)((i == 0 ? ri + dd : ri) * sc);
  q[i] = (K)(-ri / dd);
}

// ------------------------------------------------------------ step kernels
// The one-workgroup kernels below stage the small dense state they loop
// over (T, the Hessenberg column, the rotations) in LDS with parallel
// loads first, so their O(j) loops never wait on a global load.
constexpr int kMaxT = 65 * 65;  // T for restart lengths up to 63

__device__ __forceinline__ void stage(const double* __restrict__ src, int count, double* dst) {
  for (int e = threadIdx.x; e < count; e += blockDim.x) dst[e] = src[e];
}

// c = U_j^T w (from partials); d = T_j^T c.
__global__ __launch_bounds__(256) void hh_dT_kernel(int nb, int pstride, const double* partial, int j,
                                                    double* small, SmallLayout lay) {
  double* misc = small + lay.misc();
  if (is_done(misc)) return;
  __shared__ double c[72];
  __shared__ double Ts[kMaxT];
  const int ldT = lay.ldT();
  stage(small + lay.T(), (j + 1) * ldT, Ts);
  finalize_partials(nb, j + 1, false, pstride, partial, c);
  __syncthreads();
  double* d = small + lay.d();
  for (int k = threadIdx.x; k <= j; k += 256) {
    double v = 0.0;
    for (int l = 0; l <= k; l++) v += Ts[l + k * ldT] * c[l];
    d[k] = v;
  }
}

// w <- w - U_j d.
template <typename K, int NC>
__global__ __launch_bounds__(kBS) void hh_update_kernel(int n, int ld, const K* __restrict__ U, int ncols,
                                                        const double* __restrict__ small, SmallLayout lay,
                                                        K* __restrict__ w) {
  const double* misc = small + lay.misc();
  if (is_done(misc)) return;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int i = b * kBS + threadIdx.x;
  if (i >= n) return;
  const double* d = small + lay.d();
  K v = w[i];
#pragma unroll
  for (int k = 0; k < NC; k++)
    if (k < ncols) v -= U[(size_t)k * ld + i] * (K)d[k];
  w[i] = v;
}

// w <- w - U_j d and, from the updated rows >= rowstart, the tail
// multi-dot U_j^T w[rowstart:] and ||w[rowstart:]||^2 as per-block partials
// (the slabs of mdot_kernel): one pass over U_j instead of two.  Each
// thread keeps its row's NC reflector entries in registers for both uses.
template <typename K, int NC>
__global__ __launch_bounds__(kBS) void hh_update_mdot_kernel(int n, int ld, const K* __restrict__ U,
                                                             int ncols, K* __restrict__ w, int rowstart,
                                                             int rpb, const double* __restrict__ small,
                                                             SmallLayout lay, double* __restrict__ partial,
                                                             int pstride) {
  const double* misc = small + lay.misc();
  if (is_done(misc)) return;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int r0 = b * rpb, r1 = min(r0 + rpb, n);
  const double* d = small + lay.d();
  K dk[NC], acc[NC];
#pragma unroll
  for (int k = 0; k < NC; k++) {
    dk[k] = k < ncols ? (K)d[k] : K(0);
    acc[k] = K(0);
  }
  K self = K(0);
  // groups of V consecutive rows: 16-byte words up to 8 columns, single
  // rows beyond (a thread holds NC * V reflector entries for both uses;
  // measured at 2M rows, FP32: 16 columns take 24 us with single rows,
  // 26 us with 8-byte and 29 us with 16-byte words, which cost occupancy);
  // slabs start on multiples of 64 rows, and a group cut by the end of the
  // matrix goes element by element
  constexpr int V = NC <= 8 ? 16 / (int)sizeof(K) : 1;
  for (int i0 = r0 + (int)threadIdx.x * V; i0 < r1; i0 += kBS * V) {
    const bool full = i0 + V <= r1;
    K v[V];
    if (full) {
      wload<K, V>(w + i0, v);
    } else {
#pragma unroll
      for (int e = 0; e < V; e++) v[e] = i0 + e < r1 ? w[i0 + e] : K(0);
    }
    K u[NC][V];
#pragma unroll
    for (int k = 0; k < NC; k++) {
      if (k < ncols && full) {
        wload<K, V>(U + (size_t)k * ld + i0, u[k]);
      } else {
#pragma unroll
        for (int e = 0; e < V; e++) u[k][e] = (k < ncols && i0 + e < r1) ? U[(size_t)k * ld + i0 + e] : K(0);
      }
    }
#pragma unroll
    for (int e = 0; e < V; e++) {
#pragma unroll
      for (int k = 0; k < NC; k++) v[e] -= u[k][e] * dk[k];
    }
    if (full) {
      wstore<K, V>(w + i0, v);
    } else {
#pragma unroll
      for (int e = 0; e < V; e++)
        if (i0 + e < r1) w[i0 + e] = v[e];
    }
#pragma unroll
    for (int e = 0; e < V; e++) {
      if (i0 + e >= rowstart && i0 + e < r1) {
#pragma unroll
        for (int k = 0; k < NC; k++) acc[k] += u[k][e] * v[e];
        self += v[e] * v[e];
      }
    }
  }
  __shared__ double red[kBS / 64][NC + 1];
  const int wv = threadIdx.x >> 6;
  wave_multi_sum<NC>(acc, red[wv]);
  const double sv = wave_sum((double)self);
  if ((threadIdx.x & 63) == 0) red[wv][NC] = sv;
  __syncthreads();
  for (int k = threadIdx.x; k <= NC; k += kBS) {
    const double t = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
    if (k < ncols) partial[(size_t)b * pstride + k] = t;
    if (k == NC) partial[(size_t)b * pstride + pstride - 1] = t;
  }
}

// After the tail multi-dot: s = U_j^T w[j+1:], nu = ||w[j+1:]||^2.
// New reflector u_{j+1} = (w[j+1:] + dd e_{j+1}) * scale; T column j+1;
// Hessenberg column j rotated by the previous Givens rotations; new
// rotation (src/givens.cpp:46-67 with tol = the smallest subnormal);
// residual estimate |g_{j+1}|; y' = T_{j+1} U_{j+1}^T e_{j+1} for q.
template <typename K>
__global__ __launch_bounds__(256) void hh_newcol_small_kernel(int nb, int pstride, const double* partial,
                                                              int j, int n, int ld,
                                                              const K* __restrict__ U,
                                                              const K* __restrict__ w, double* small,
                                                              SmallLayout lay) {
  double* misc = small + lay.misc();
  if (is_done(misc)) return;
  __shared__ double s[72];
  __shared__ double e[72];
  __shared__ double h[72];   // Hessenberg column j
  __shared__ double cz[72];  // previous rotations
  __shared__ double sz[72];
  __shared__ double sc[4];
  __shared__ double Ts[kMaxT];
  double* T = small + lay.T();
  const int ldT = lay.ldT();
  const int j1 = j + 1;
  const double* cs = small + lay.cs();
  const double* sn = small + lay.sn();
  // parallel staging: T's first j+1 columns, w'[0..j], the rotations, and
  // row j+1 of the reflectors
  stage(T, j1 * ldT, Ts);
  for (int i = threadIdx.x; i <= j; i += 256) {
    h[i] = (double)w[i];
    e[i] = j1 < n ? (double)U[(size_t)i * ld + j1] : 0.0;
    if (i < j) {
      cz[i] = cs[i];
      sz[i] = sn[i];
    }
  }
  finalize_partials(nb, j + 1, true, pstride, partial, s);  // s[0..j], s[j+1] = nu
  __syncthreads();
  if (threadIdx.x == 0) {
    const double nu = s[j1];
    const double alpha = j1 < n ? (double)w[j1] : 0.0;
    const double sq = sqrt(nu);
    const double dd = (alpha >= 0.0 ? 1.0 : -1.0) * sq;
    const double nv = sqrt(2.0 * sq * (sq + fabs(alpha)));
    const double scale = nv > 0.0 ? 1.0 / nv : 0.0;
    misc[MS_DD] = dd;
    misc[MS_SCALE] = scale;
    misc[MS_ALPHA] = alpha;
    sc[0] = dd;
    sc[1] = scale;
    // e = U_{j+1}^T e_{j+1}: row j+1 of the old columns, then u_{j+1}[j+1]
    e[j1] = (alpha + dd) * scale;
  }
  __syncthreads();
  const double dd = sc[0], scale = sc[1];
  // utu = U_j^T u_{j+1} = scale * (s + dd * e[0..j]); T[0..j, j+1] = -2 T_j utu
  for (int l = threadIdx.x; l <= j1; l += 256) {
    double v = 2.0;
    if (l <= j) {
      v = 0.0;
      for (int k = l; k <= j; k++) v += Ts[l + k * ldT] * (scale * (s[k] + dd * e[k]));
      v = -2.0 * v;
    }
    Ts[l + j1 * ldT] = v;
    T[l + j1 * ldT] = v;
  }
  __syncthreads();
  // y' = T_{j+1} e
  double* yq = small + lay.y();
  for (int l = threadIdx.x; l <= j1; l += 256) {
    double v = 0.0;
    for (int k = l; k <= j1; k++) v += Ts[l + k * ldT] * e[k];
    yq[l] = v;
  }
  if (threadIdx.x == 0) {
    double* csw = small + lay.cs();
    double* snw = small + lay.sn();
    double* g = small + lay.g();
    // Hessenberg column: h_i = w'_i (i <= j), h_{j+1} = -dd; apply the
    // previous rotations
    for (int i = 0; i < j; i++) {
      const double t0 = h[i], t1 = h[i + 1];
      h[i + 1] = -sz[i] * t0 + cz[i] * t1;
      h[i] = cz[i] * t0 + sz[i] * t1;
    }
    const double a = h[j], bb = -dd;
    const double tol = 4.9406564584124654e-324;  // nextafter(0,1)
    double c, sn_;
    if (fabs(bb) <= tol) {
      c = 1.0;
      sn_ = 0.0;
    } else if (fabs(bb) > fabs(a)) {
      const double t = a / bb;
      sn_ = 1.0 / sqrt(1.0 + t * t);
      c = t * sn_;
    } else {
      const double t = bb / a;
      c = 1.0 / sqrt(1.0 + t * t);
      sn_ = t * c;
    }
    csw[j] = c;
    snw[j] = sn_;
    h[j] = c * a - sn_ * dd;
    const double gj = g[j];
    g[j1] = -sn_ * gj;
    g[j] = c * gj;
    const double res = fabs(g[j1]);
    misc[MS_RES] = res;
    misc[MS_NCOLS] = (double)j1;
    if (!(res >= misc[MS_RTOL]) || !isfinite(res)) misc[MS_DONE] = 1.0;
    if (!isfinite(res)) misc[MS_BREAK] = 1.0;
  }
  __syncthreads();
  // rotated Hessenberg column -> R
  double* R = small + lay.R();
  const int ldR = lay.m + 1;
  for (int i = threadIdx.x; i <= j; i += 256) R[i + j * ldR] = h[i];
}

// u_{j+1} = (w + dd e_{j+1}) * scale on rows >= j+1 (0 above);
// q = e_{j+1} - U_{j+1} y'.
template <typename K, int NC>
__global__ __launch_bounds__(kBS) void hh_newcol_vec_kernel(int n, int ld, K* __restrict__ U, int j,
                                                            const K* __restrict__ w,
                                                            const double* __restrict__ small,
                                                            SmallLayout lay, K* __restrict__ q,
                                                            int write_q) {
  const double* misc = small + lay.misc();
  if (is_done(misc)) return;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int i = b * kBS + threadIdx.x;
  if (i >= n) return;
  const int j1 = j + 1;
  const double dd = misc[MS_DD], scale = misc[MS_SCALE];
  K unew = K(0);
  if (i >= j1) unew = (K)(((double)w[i] + (i == j1 ? dd : 0.0)) * scale);
  U[(size_t)j1 * ld + i] = unew;
  if (!write_q) return;
  const double* yq = small + lay.y();
  K v = (i == j1) ? K(1) : K(0);
#pragma unroll
  for (int k = 0; k < NC; k++)
    if (k <= j) v -= U[(size_t)k * ld + i] * (K)yq[k];
  v -= unew * (K)yq[j1];
  q[i] = v;
}

// Least squares y = R^{-1} g by the column-oriented back substitution of
// src/fgmres_householder_restart.cpp:721-729, for the ncols columns done.
// R's leading nc x nc block and g are staged in LDS; each step jj updates
// y_i (i < jj) in parallel, one lane per i, so every y_i receives the same
// operations in the same order as in the serial loop.  One wave (64 lanes):
// nc <= 63.
__global__ __launch_bounds__(64) void hh_lsq_kernel(double* small, SmallLayout lay) {
  __shared__ double Rs[64 * 64];
  __shared__ double ys[64];
  const double* misc = small + lay.misc();
  const int nc = (int)misc[MS_NCOLS];
  const double* R = small + lay.R();
  const double* g = small + lay.g();
  double* y = small + lay.y();
  const int ldR = lay.m + 1;
  const int t = threadIdx.x;
  for (int e = t; e < nc * nc; e += 64) Rs[e] = R[e % nc + (e / nc) * ldR];
  __syncthreads();
  if (t < nc) ys[t] = g[t] / Rs[t + t * nc];
  __syncthreads();
  for (int jj = nc - 1; jj > 0; jj--) {
    const double yj = ys[jj];
    if (t < jj) ys[t] = ys[t] - Rs[t + jj * nc] * yj / Rs[t + t * nc];
    __syncthreads();
  }
  if (t < nc) y[t] = ys[t];
}

// x += Z y over the ncols columns done.
template <typename XT, typename ZT, int NC>
__global__ __launch_bounds__(kBS) void hh_xupdate_kernel(int n, int ld, const ZT* __restrict__ Z,
                                                         const double* __restrict__ small,
                                                         SmallLayout lay, XT* __restrict__ x) {
  const double* misc = small + lay.misc();
  const int nc = (int)misc[MS_NCOLS];
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int i = b * kBS + threadIdx.x;
  if (i >= n) return;
  const double* y = small + lay.y();
  using ACC = typename std::conditional<sizeof(XT) == 8, double, float>::type;
  ACC v = (ACC)x[i];
#pragma unroll
  for (int k = 0; k < NC; k++)
    if (k < nc) v += (ACC)Z[(size_t)k * ld + i] * (ACC)y[k];
  x[i] = (XT)v;
}

// The same update with four consecutive rows per thread, moved as 16-byte
// words (x and each column of Z); each row's sum is the one above, term by
// term, so the result is bit-identical.  Needs n % 4 == 0 and 16-byte
// aligned x, Z and ld (the engine's buffers are).
template <typename XT, typename ZT, int NC>
__global__ __launch_bounds__(kBS) void hh_xupdate4_kernel(int n, int ld, const ZT* __restrict__ Z,
                                                          const double* __restrict__ small,
                                                          SmallLayout lay, XT* __restrict__ x) {
  const double* misc = small + lay.misc();
  const int nc = (int)misc[MS_NCOLS];
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int i0 = (b * kBS + (int)threadIdx.x) * 4;
  if (i0 >= n) return;
  const double* y = small + lay.y();
  using ACC = typename std::conditional<sizeof(XT) == 8, double, float>::type;
  XT xv[4];
  vload<XT, 4>(x + i0, xv);
  ACC v[4];
#pragma unroll
  for (int e = 0; e < 4; e++) v[e] = (ACC)xv[e];
#pragma unroll
  for (int k = 0; k < NC; k++)
    if (k < nc) {
      ZT z[4];
      vload<ZT, 4>(Z + (size_t)k * ld + i0, z);
      const ACC yk = (ACC)y[k];
#pragma unroll
      for (int e = 0; e < 4; e++) v[e] += (ACC)z[e] * yk;
    }
#pragma unroll
  for (int e = 0; e < 4; e++) xv[e] = (XT)v[e];
  vstore<XT, 4>(x + i0, xv);
}

template <typename To, typename From>
__global__ void cast_kernel(int n, const From* __restrict__ in, To* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (To)in[i];
}

__global__ void set_misc_kernel(double* small, SmallLayout lay, int slot, double v) {
  if (threadIdx.x == 0) small[lay.misc() + slot] = v;
}

// ------------------------------------------------------------ engine
template <typename K>
int HHEngine<K>::init(int n_, int m_, bool z_float_) {
  LACE_REQUIRE(m_ >= 1 && m_ <= 63, DEV_ERR_NOT_SUPPORTED,
               "restart length %d outside the supported range 1..63", m_);
  n = n_;
  m = m_;
  z_float = z_float_;
  ld = ((n + 63) / 64) * 64;
  lay = SmallLayout{m};
  // about 1024 row slabs (four per CU), at least 1024 rows each
  rpb = std::max(1024, ((ceil_div(n, 1024) + 63) / 64) * 64);
  nb = ceil_div(n, rpb);
  pstride = m + 3;
  LACE_TRY(U.alloc(sizeof(K) * (size_t)ld * (m + 1)));
  LACE_TRY(Z.alloc((z_float ? 4 : sizeof(K)) * (size_t)ld * m));
  LACE_TRY(w.alloc(sizeof(K) * ld));
  LACE_TRY(q.alloc(sizeof(K) * ld));
  LACE_TRY(r.alloc(sizeof(K) * ld));
  LACE_TRY(small.alloc(sizeof(double) * lay.total()));
  LACE_TRY(partial.alloc(sizeof(double) * (size_t)nb * pstride));
  LACE_TRY(ptot.alloc(sizeof(double) * pstride));
  LACE_HIP(hipMemsetAsync(small.p, 0, sizeof(double) * lay.total(), stream()));
  LACE_HIP(hipMemsetAsync(U.p, 0, sizeof(K) * (size_t)ld * (m + 1), stream()));
  LACE_HIP(hipHostMalloc((void**)&host_misc, sizeof(double) * 32, hipHostMallocDefault));
  return DEV_SUCCESS;
}

template <typename K>
HHEngine<K>::~HHEngine() {
  if (host_misc) (void)hipHostFree(host_misc);
}

template <typename K>
void* HHEngine<K>::zcol(int j) const {
  return (char*)Z.p + (size_t)j * ld * (z_float ? 4 : sizeof(K));
}

template struct HHEngine<double>;
template struct HHEngine<float>;

// Dispatch on the compile-time column bound.
#define LACE_NC_DISPATCH(ncols, BODY) \
  do {                                \
    if ((ncols) <= 8) {               \
      constexpr int NCB = 8;          \
      BODY;                           \
    } else if ((ncols) <= 16) {       \
      constexpr int NCB = 16;         \
      BODY;                           \
    } else if ((ncols) <= 32) {       \
      constexpr int NCB = 32;         \
      BODY;                           \
    } else {                          \
      constexpr int NCB = 64;         \
      BODY;                           \
    }                                 \
  } while (0)

// Norm pass over r and the start of a cycle (u_0, q).  set_rtol: store
// rtol_rel * ||r|| as the absolute tolerance.
template <typename K>
int hh_start(HHEngine<K>& E, int set_rtol, double rtol_rel) {
  TimedRegion tr("hh_start");
  launch_mdot_rows<K, 8, true>(E, E.U.template as<K>(), 0, E.r.template as<K>(), 0,
                               E.partial.template as<double>(), nullptr);
  hh_start_kernel<K><<<1, 256, 0, stream()>>>(E.nb, E.pstride, E.partial.template as<double>(),
                                              E.r.template as<K>(), E.small.template as<double>(),
                                              E.lay, set_rtol, rtol_rel);
  hh_start_vec_kernel<K><<<ceil_div(E.n, kBS), kBS, 0, stream()>>>(
      E.n, E.r.template as<K>(), E.U.template as<K>(), E.q.template as<K>(),
      E.small.template as<double>() + E.lay.misc());
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

// Householder part of step j (after w = A z_j).  Beyond 16 columns the
// per-block partials are first reduced column by column on many
// workgroups, and the one-workgroup kernels read the totals.
template <typename K>
int hh_step(HHEngine<K>& E, int j, bool write_q) {
  TimedRegion tr("householder");
  const int nc = j + 1;
  K* U = E.U.template as<K>();
  K* w = E.w.template as<K>();
  double* small = E.small.template as<double>();
  double* part = E.partial.template as<double>();
  double* ptot = E.ptot.template as<double>();
  const double* misc = small + E.lay.misc();
  const int nbv = ceil_div(E.n, kBS);
  const bool wide = nc > 16;
  if (nc <= 8)
    launch_mdot_rows<K, 8, false>(E, U, nc, w, 0, part, misc);
  else
    launch_mdot_rows<K, 16, false>(E, U, nc, w, 0, part, misc);
  if (wide) reduce_cols_kernel<<<nc, 256, 0, stream()>>>(E.nb, nc, 0, E.pstride, part, ptot, misc);
  hh_dT_kernel<<<1, 256, 0, stream()>>>(wide ? 1 : E.nb, E.pstride, wide ? ptot : part, j, small, E.lay);
  // fused update + tail multi-dot while the per-thread reflector row fits
  // in registers (up to 16 columns; 32 in FP32)
  if (nc <= 16 || (sizeof(K) == 4 && nc <= 32)) {
    if (nc <= 8)
      hh_update_mdot_kernel<K, 8><<<E.nb, kBS, 0, stream()>>>(E.n, E.ld, U, nc, w, j + 1, E.rpb,
                                                             small, E.lay, part, E.pstride);
    else if (nc <= 16)
      hh_update_mdot_kernel<K, 16><<<E.nb, kBS, 0, stream()>>>(E.n, E.ld, U, nc, w, j + 1, E.rpb,
                                                              small, E.lay, part, E.pstride);
    else
      hh_update_mdot_kernel<K, 32><<<E.nb, kBS, 0, stream()>>>(E.n, E.ld, U, nc, w, j + 1, E.rpb,
                                                              small, E.lay, part, E.pstride);
  } else {
    LACE_NC_DISPATCH(nc, (hh_update_kernel<K, NCB><<<nbv, kBS, 0, stream()>>>(E.n, E.ld, U, nc, small,
                                                                              E.lay, w)));
    launch_mdot_rows<K, 16, true>(E, U, nc, w, j + 1, part, misc);
  }
  if (wide) reduce_cols_kernel<<<nc + 1, 256, 0, stream()>>>(E.nb, nc, 1, E.pstride, part, ptot, misc);
  hh_newcol_small_kernel<K><<<1, 256, 0, stream()>>>(wide ? 1 : E.nb, E.pstride, wide ? ptot : part, j,
                                                     E.n, E.ld, U, w, small, E.lay);
  LACE_NC_DISPATCH(nc, (hh_newcol_vec_kernel<K, NCB><<<nbv, kBS, 0, stream()>>>(
                           E.n, E.ld, U, j, w, small, E.lay, E.q.template as<K>(), write_q ? 1 : 0)));
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

// y = R^{-1} g and x += Z y.
template <typename K, typename XT>
int hh_finish(HHEngine<K>& E, XT* x) {
  TimedRegion tr("hh_finish");
  double* small = E.small.template as<double>();
  hh_lsq_kernel<<<1, 64, 0, stream()>>>(small, E.lay);
  const int nbv = ceil_div(E.n, kBS);
  const bool vec = E.n % 4 == 0 && reinterpret_cast<uintptr_t>(x) % 16 == 0;
  const int nb4 = ceil_div(E.n / 4, kBS);
  if (E.z_float) {
    if (vec)
      LACE_NC_DISPATCH(E.m, (hh_xupdate4_kernel<XT, float, NCB><<<nb4, kBS, 0, stream()>>>(
                                E.n, E.ld, E.Z.template as<float>(), small, E.lay, x)));
    else
      LACE_NC_DISPATCH(E.m, (hh_xupdate_kernel<XT, float, NCB><<<nbv, kBS, 0, stream()>>>(
                                E.n, E.ld, E.Z.template as<float>(), small, E.lay, x)));
  } else {
    if (vec)
      LACE_NC_DISPATCH(E.m, (hh_xupdate4_kernel<XT, K, NCB><<<nb4, kBS, 0, stream()>>>(
                                E.n, E.ld, E.Z.template as<K>(), small, E.lay, x)));
    else
      LACE_NC_DISPATCH(E.m, (hh_xupdate_kernel<XT, K, NCB><<<nbv, kBS, 0, stream()>>>(
                                E.n, E.ld, E.Z.template as<K>(), small, E.lay, x)));
  }
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

template <typename K>
int read_misc(HHEngine<K>& E) {
  LACE_HIP(hipMemcpyAsync(E.host_misc, E.small.template as<double>() + E.lay.misc(),
                          sizeof(double) * 16, hipMemcpyDeviceToHost, stream()));
  LACE_HIP(hipStreamSynchronize(stream()));
  return DEV_SUCCESS;
}

template int hh_start<double>(HHEngine<double>&, int, double);
template int hh_start<float>(HHEngine<float>&, int, double);
template int hh_step<double>(HHEngine<double>&, int, bool);
template int hh_step<float>(HHEngine<float>&, int, bool);
template int hh_finish<double, double>(HHEngine<double>&, double*);
template int hh_finish<float, float>(HHEngine<float>&, float*);
template int read_misc<double>(HHEngine<double>&);
template int read_misc<float>(HHEngine<float>&);

// ------------------------------------------------------------ CGS inner GMRES
// GMRES with classical Gram-Schmidt on a basis kept unnormalised: column k
// of U holds W_k, the basis vector v_k = W_k / ||W_k|| times nothing (the
// scales s_k = 1 / ||W_k|| live in the small state), and Z column k holds
// M^{-1} W_k.  Step j: w = A Z_j; one pass forms the dots W_k . w (k <= j);
// one one-workgroup kernel turns them into the Hessenberg column of the
// basis Z (h_kj = s_k (W_k . w)) and, since ||W_j||^2 arrived with the same
// partials, finishes column j - 1 (its subdiagonal h_{j,j-1} = ||W_j||, the
// Givens rotations, the residual estimate); one pass writes W_{j+1} = w -
// sum_k h_kj s_k W_k and the partial sums of its squared norm.  With Z_j =
// M^{-1} W_j the relation A Z_j = sum_{k <= j+1} h_kj v_k holds for the
// vectors as stored, so x = Z y with y the least-squares solution is a GMRES
// iterate of the preconditioned operator.  Two vector passes and one small
// launch per step (the Householder engine takes three and two), at the
// price of classical Gram-Schmidt's orthogonality, which a fixed short
// inner solve inside a flexible outer FGMRES can afford.  The scales go in
// the T slot of the small state, the raw Hessenberg column in R's column.
template <typename K, int NC>
__global__ __launch_bounds__(kBS) void cgs_update_kernel(int n, int ld, K* __restrict__ U, int ncols,
                                                         const K* __restrict__ w, int store, int rpb,
                                                         const double* __restrict__ small, SmallLayout lay,
                                                         double* __restrict__ partial, int pstride) {
  const double* misc = small + lay.misc();
  if (is_done(misc)) return;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int r0 = b * rpb, r1 = min(r0 + rpb, n);
  const double* d = small + lay.d();
  K dk[NC];
#pragma unroll
  for (int k = 0; k < NC; k++) dk[k] = k < ncols ? (K)d[k] : K(0);
  // 16-byte words of rows per thread up to 16 columns (the coefficients
  // stay in registers, the columns stream through one word at a time)
  constexpr int V = NC <= 16 ? 16 / (int)sizeof(K) : 1;
  K ss = K(0);
  K* Wn = U + (size_t)ncols * ld;
  for (int i0 = r0 + (int)threadIdx.x * V; i0 < r1; i0 += kBS * V) {
    if (i0 + V <= r1) {
      K v[V];
      wload<K, V>(w + i0, v);
#pragma unroll
      for (int k = 0; k < NC; k++) {
        if (k < ncols) {
          K u[V];
          wload<K, V>(U + (size_t)k * ld + i0, u);
#pragma unroll
          for (int e = 0; e < V; e++) v[e] -= u[e] * dk[k];
        }
      }
#pragma unroll
      for (int e = 0; e < V; e++) ss += v[e] * v[e];
      if (store) {
#pragma unroll
        for (int e = 0; e < V; e++) Wn[i0 + e] = v[e];
      }
    } else {
      for (int i = i0; i < r1; i++) {
        K v = w[i];
#pragma unroll
        for (int k = 0; k < NC; k++)
          if (k < ncols) v -= U[(size_t)k * ld + i] * dk[k];
        ss += v * v;
        if (store) Wn[i] = v;
      }
    }
  }
  __shared__ double sm[kBS / 64];
  const double tot = block_sum<kBS>((double)ss, sm);
  if (threadIdx.x == 0) partial[(size_t)b * pstride + pstride - 1] = tot;
}

__global__ void cgs_reset_kernel(double* small, SmallLayout lay) {
  double* misc = small + lay.misc();
  misc[MS_DONE] = 0.0;
  misc[MS_BREAK] = 0.0;
  misc[MS_NCOLS] = 0.0;
}

// Column j - 1 finished with its subdiagonal h = ||W_j|| (j >= 1): the
// previous rotations, a new one (src/givens.cpp:46-67), g, the residual;
// returns whether the cycle is done (also recorded in misc).
// hs: column jc of R staged in LDS (rotated in place, written back by the
// caller); css, sns: the earlier rotations, gj = g[jc] and rtol staged the
// same way, so the serial chain below waits on LDS, not on memory.
__device__ bool cgs_close_column(double* small, SmallLayout lay, int jc, double sub, double* hs,
                                 const double* css, const double* sns, double gj, double rtol) {
  double* misc = small + lay.misc();
  double* cs = small + lay.cs();
  double* sn = small + lay.sn();
  double* g = small + lay.g();
  double t0 = hs[0];  // h_i as rotated by the steps before i
  for (int i = 0; i < jc; i++) {
    const double t1 = hs[i + 1];
    const double h1 = -sns[i] * t0 + css[i] * t1;
    hs[i] = css[i] * t0 + sns[i] * t1;
    t0 = h1;
  }
  const double a = t0, bb = sub;
  const double tol = 4.9406564584124654e-324;  // nextafter(0,1)
  double c, s;
  if (fabs(bb) <= tol) {
    c = 1.0;
    s = 0.0;
  } else if (fabs(bb) > fabs(a)) {
    const double t = a / bb;
    s = 1.0 / sqrt(1.0 + t * t);
    c = t * s;
  } else {
    const double t = bb / a;
    c = 1.0 / sqrt(1.0 + t * t);
    s = t * c;
  }
  cs[jc] = c;
  sn[jc] = s;
  hs[jc] = c * a + s * bb;
  g[jc + 1] = -s * gj;
  g[jc] = c * gj;
  const double res = fabs(-s * gj);
  misc[MS_RES] = res;
  misc[MS_NCOLS] = (double)(jc + 1);
  const bool done = !(res >= rtol) || !isfinite(res) || !(sub > 0.0);
  if (done) misc[MS_DONE] = 1.0;
  if (!isfinite(res)) misc[MS_BREAK] = 1.0;
  return done;
}

// Step j's small part (j = m: the closing of column m - 1 only).  The state
// the closing of column j - 1 reads is staged in LDS by every lane while
// the partials are summed.
__global__ __launch_bounds__(256) void cgs_small_kernel(int nb, int pstride, const double* partial, int j,
                                                        double* small, SmallLayout lay, int set_rtol,
                                                        double rtol_rel) {
  double* misc = small + lay.misc();
  // the flag is tested once the partials are summed, so that its load and
  // theirs are in flight together (a finished cycle's partials are stale
  // but in bounds, and nothing is written before the test)
  const bool done = is_done(misc);
  __shared__ double s[72];
  __shared__ double hs[64], css[64], sns[64], pre[2], scs[72];
  __shared__ int fin;
  double* R = small + lay.R();
  double* sc = small + lay.T();  // 1 / ||W_k||, staged in scs
  for (int k = threadIdx.x; k < j; k += 256) scs[k] = sc[k];
  const int ldR = lay.m + 1;
  const int jc = j - 1;
  if (j >= 1) {
    const double* cs = small + lay.cs();
    const double* sn = small + lay.sn();
    for (int k = threadIdx.x; k <= jc; k += 256) hs[k] = R[k + (size_t)jc * ldR];
    for (int k = threadIdx.x; k < jc; k += 256) {
      css[k] = cs[k];
      sns[k] = sn[k];
    }
    if (threadIdx.x == 0) pre[0] = small[lay.g() + jc];
  }
  if (threadIdx.x == 0) pre[1] = misc[MS_RTOL];
  const int cnt = j < lay.m ? j + 1 : 0;
  finalize_partials(nb, cnt, true, pstride, partial, s);  // s[0..cnt), s[cnt] = ||W_j||^2
  __syncthreads();
  if (done) return;
  if (threadIdx.x == 0) {
    const double nrm = sqrt(s[cnt]);
    if (j == 0) {
      misc[MS_RNORM] = nrm;
      misc[MS_RES] = nrm;
      const double rtol = set_rtol ? rtol_rel * nrm : pre[1];
      if (set_rtol) misc[MS_RTOL] = rtol;
      double* g = small + lay.g();
      for (int t = 0; t < lay.m + 2; t++) g[t] = 0.0;
      g[0] = nrm;
      fin = !(nrm >= rtol) || !isfinite(nrm) || nrm == 0.0;
      if (fin) misc[MS_DONE] = 1.0;
    } else {
      fin = cgs_close_column(small, lay, jc, nrm, hs, css, sns, pre[0], pre[1]);
    }
    scs[j] = nrm > 0.0 ? 1.0 / nrm : 0.0;
    sc[j] = scs[j];
  }
  __syncthreads();
  if (j >= 1)
    for (int k = threadIdx.x; k <= jc; k += 256) R[k + (size_t)jc * ldR] = hs[k];
  if (j >= lay.m || fin) return;
  // column j: h_kj = s_k (W_k . w) into R's column (rotated when it closes)
  // and the update's coefficients d_k = h_kj s_k
  double* d = small + lay.d();
  for (int k = threadIdx.x; k <= j; k += 256) {
    const double h = scs[k] * s[k];
    R[k + (size_t)j * ldR] = h;
    d[k] = h * scs[k];
  }
}

int cast_d2f(int n, const double* in, float* out);

// W_0 = r (cast), the partial sums of ||r||^2, and a fresh cycle state.
template <typename K>
int cgs_start(HHEngine<K>& E, const double* r) {
  TimedRegion tr("cgs_start");
  K* U = E.U.template as<K>();
  if constexpr (sizeof(K) == 4) {
    LACE_TRY(cast_d2f(E.n, r, U));
  } else {
    LACE_HIP(hipMemcpyAsync(U, r, sizeof(double) * E.n, hipMemcpyDeviceToDevice, stream()));
  }
  cgs_reset_kernel<<<1, 1, 0, stream()>>>(E.small.template as<double>(), E.lay);
  launch_mdot_rows<K, 8, true>(E, U, 0, U, 0, E.partial.template as<double>(), nullptr);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

// Step j after w = A Z_j (Z_j = M^{-1} W_j).  last: the step's new vector
// is not stored, only its norm (for closing column j), and the closing
// follows.
template <typename K>
int cgs_step(HHEngine<K>& E, int j, bool last, double rtol_rel) {
  TimedRegion tr("cgs_step");
  const int nc = j + 1;
  K* U = E.U.template as<K>();
  K* w = E.w.template as<K>();
  double* small = E.small.template as<double>();
  double* part = E.partial.template as<double>();
  const double* misc = small + E.lay.misc();
  if (nc <= 8)
    launch_mdot_rows<K, 8, false>(E, U, nc, w, 0, part, misc);
  else
    launch_mdot_rows<K, 16, false>(E, U, nc, w, 0, part, misc);
  cgs_small_kernel<<<1, 256, 0, stream()>>>(E.nb, E.pstride, part, j, small, E.lay, 1, rtol_rel);
  LACE_NC_DISPATCH(nc, (cgs_update_kernel<K, NCB><<<E.nb, kBS, 0, stream()>>>(
                           E.n, E.ld, U, nc, w, last ? 0 : 1, E.rpb, small, E.lay, part, E.pstride)));
  if (last) cgs_small_kernel<<<1, 256, 0, stream()>>>(E.nb, E.pstride, part, E.m, small, E.lay, 1, rtol_rel);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

template int cgs_start<float>(HHEngine<float>&, const double*);
template int cgs_start<double>(HHEngine<double>&, const double*);
template int cgs_step<float>(HHEngine<float>&, int, bool, double);
template int cgs_step<double>(HHEngine<double>&, int, bool, double);

// Helpers used by the solver for norms and casts.
int norm2_sq_partials(int n, int rpb, const double* x, double* partial, int pstride, int nb) {
  mdot_kernel<double, 8, true><<<nb, kBS, 0, stream()>>>(n, n, x, 0, x, 0, rpb, partial, pstride,
                                                        nullptr);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

__global__ __launch_bounds__(256) void sum_slot_kernel(int nb, int pstride, const double* partial, double* out) {
  __shared__ double v[1];
  finalize_partials(nb, 0, true, pstride, partial, v);
  __syncthreads();
  if (threadIdx.x == 0) *out = v[0];
}

int sum_slot(int nb, int pstride, const double* partial, double* out) {
  sum_slot_kernel<<<1, 256, 0, stream()>>>(nb, pstride, partial, out);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

int cast_d2f(int n, const double* in, float* out) {
  cast_kernel<float, double><<<ceil_div(n, 256), 256, 0, stream()>>>(n, in, out);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}
int cast_f2d(int n, const float* in, double* out) {
  cast_kernel<double, float><<<ceil_div(n, 256), 256, 0, stream()>>>(n, in, out);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

int set_misc_d(double* small, SmallLayout lay, int slot, double v) {
  set_misc_kernel<<<1, 64, 0, stream()>>>(small, lay, slot, v);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

}  // namespace lace
