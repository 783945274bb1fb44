// parilu.hip -- ParILU(0) fixed-point sweeps on the GPU.
//
// Reference: src/parilu_v0_3.cpp:16-132 (the sweep body :70-102),
// src/parilu_v0_3_prescribedSweeps.cpp:17-132, src/parilu_v0_3_gpu.cpp:17-242.
//
// Layout in HBM (one set per handle):
//   A      CSR (row, col) + COO row index per nonzero + values in the
//          iteration precision T (double for F64, float for F32/BF16)
//   L      CSRL: strict-lower entries of A's pattern then the unit diagonal,
//          two value buffers (Jacobi double buffering)
//   U      CSCU: U by columns, rows ascending, diagonal last; two buffers
//   dst    per nonzero of A, its slot in L (>= 0) or U (-1 - slot); used to
//          re-initialise the factors from A on the device
//
// Sweep semantics: synchronous (Jacobi).  Every update of sweep k reads
// iterate k-1 only, so the result does not depend on scheduling.  The
// per-nonzero arithmetic is the reference's: the merge subtracts each
// rounded product l_ik*u_kj in ascending k, adds the last product back
// (src/parilu_v0_3.cpp:90) and divides by u_jj for L entries.  Contraction
// is disabled in the kernel so the FP64 and FP32 results equal the CPU
// restatement bit for bit.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "parilu.hpp"

namespace lace {

constexpr int kSweepBS = 256;

// One thread per nonzero of A.  STEP: also accumulate sum (old-new)^2 into
// a per-block partial.
template <typename T, bool STEP>
__global__ __launch_bounds__(kSweepBS) void parilu_sweep_kernel(
    int nnz, const int* __restrict__ rowidx, const int* __restrict__ acol,
    const T* __restrict__ aval, const int* __restrict__ lrow, const int* __restrict__ lcol,
    const T* __restrict__ Lin, T* __restrict__ Lout, const int* __restrict__ ucp,
    const int* __restrict__ uri, const T* __restrict__ Uin, T* __restrict__ Uout,
    double* __restrict__ partial) {
#pragma clang fp contract(off)
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int k = b * kSweepBS + threadIdx.x;
  double d2 = 0.0;
  if (k < nnz) {
    const int i = rowidx[k], j = acol[k];
    T s = aval[k];
    T sp = T(0);
    int il = lrow[i];
    const int ile = lrow[i + 1];
    int iu = ucp[j];
    const int iue = ucp[j + 1];
    while (il < ile && iu < iue) {
      sp = T(0);
      const int jl = lcol[il], ju = uri[iu];
      if (jl == ju) {
        sp = Lin[il] * Uin[iu];
        s = s - sp;
      }
      il += (jl <= ju);
      iu += (jl >= ju);
    }
    s = s + sp;
    if (i > j) {
      const T t = s / Uin[iue - 1];
      if (STEP) {
        double d = (double)Lin[il - 1] - (double)t;
        d2 = d * d;
      }
      Lout[il - 1] = t;
    } else {
      if (STEP) {
        double d = (double)Uin[iu - 1] - (double)s;
        d2 = d * d;
      }
      Uout[iu - 1] = s;
    }
  }
  if (STEP) {
    __shared__ double sm[kSweepBS / 64];
    double tot = block_sum<kSweepBS>(d2, sm);
    if (threadIdx.x == 0) partial[b] = tot;
  }
}

// Symmetric-pattern sweep as a stream over factor positions: thread p reads
// a_ij, a_ji, l_ij, u_ji at its own position (consecutive threads,
// consecutive addresses) and the merge partners recorded at set-up -- u_jj
// for a strict position; the earlier positions of its own row for a
// diagonal one (u_ii = a_ii - sum_k l_ik u_ki); and, when the pattern has
// any, the strict position's matches k < j.  Same arithmetic and order as
// the per-position kernel above.
template <typename T, bool STEP, bool MATCH>
__global__ __launch_bounds__(kSweepBS) void parilu_stream_sweep_kernel(
    int np, const int* __restrict__ pmeta, const int* __restrict__ mptr,
    const int* __restrict__ mpairs, const T* __restrict__ AL, const T* __restrict__ AU,
    const T* __restrict__ VLin, const T* __restrict__ VUin, T* __restrict__ VLout,
    T* __restrict__ VUout, double* __restrict__ partial) {
#pragma clang fp contract(off)
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int p = b * kSweepBS + threadIdx.x;
  double d2 = 0.0;
  if (p < np) {
    const int m = pmeta[p];
    const T uold = VUin[p];
    const T pu = T(1) * uold;  // l_jj * u_ji (unit diagonal)
    if (m < 0) {
      // diagonal: u_ii over the row's strict positions [i0, p)
      T s = AU[p];
      for (int q = -m - 1; q < p; q++) s = s - VLin[q] * VUin[q];
      s = s - pu;
      s = s + pu;
      VUout[p] = s;
      if (STEP) {
        const double d = (double)uold - (double)s;
        d2 = d * d;
      }
    } else {
      T sL = AL[p], sU = AU[p];
      const T lold = VLin[p];
      const T ujj = VUin[m];
      if (MATCH) {
        for (int t = mptr[p]; t < mptr[p + 1]; t++) {
          const int qa = mpairs[2 * t], qc = mpairs[2 * t + 1];
          sL = sL - VLin[qa] * VUin[qc];
          sU = sU - VLin[qc] * VUin[qa];
        }
      }
      sU = sU - pu;
      sU = sU + pu;
      VUout[p] = sU;
      const T pl = lold * ujj;
      sL = sL - pl;
      sL = sL + pl;
      const T t = sL / ujj;
      VLout[p] = t;
      if (STEP) {
        const double d1 = (double)uold - (double)sU;
        const double d = (double)lold - (double)t;
        d2 = d1 * d1 + d * d;
      }
    }
  }
  if (STEP) {
    __shared__ double sm[kSweepBS / 64];
    double tot = block_sum<kSweepBS>(d2, sm);
    if (threadIdx.x == 0) partial[b] = tot;
  }
}

// The streaming sweep with P consecutive positions per thread (4 in FP32,
// 2 in FP64): pmeta, a_ij, a_ji, l_ij and u_ji come in as 16-byte words
// (pmeta as a P * 4-byte word) and both outputs go out as 16-byte words;
// each position's arithmetic is the one above (a diagonal position also
// stores its unchanged unit l_ii).  np must be a multiple of P for the
// vector path; the last partial group takes the scalar form.  A diagonal
// position's products l_iq u_qi run over the strict positions of its own
// row, which precede it: those among the thread's positions and the four
// before them (the previous lanes', shuffled in before any branch) come
// from registers, only earlier ones (rows longer than the window, the
// first lanes of a wave) from memory, in the same ascending order.
template <typename T, bool STEP, bool MATCH, int P>
__global__ __launch_bounds__(kSweepBS) void parilu_stream4_sweep_kernel(
    int np, const int* __restrict__ pmeta, const int* __restrict__ mptr,
    const int* __restrict__ mpairs, const T* __restrict__ AL, const T* __restrict__ AU,
    const T* __restrict__ VLin, const T* __restrict__ VUin, T* __restrict__ VLout,
    T* __restrict__ VUout, double* __restrict__ partial) {
#pragma clang fp contract(off)
  static_assert(P == 2 || P == 4, "two or four positions per thread");
  static_assert(P * sizeof(T) % 16 == 0, "16-byte words of P values");
  constexpr int W = 4;  // positions of the window before the thread's first
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int p0 = (b * kSweepBS + (int)threadIdx.x) * P;
  const int lane = threadIdx.x & 63;
  const bool full = p0 + P - 1 < np;
  int mv[P];
  T al[P], au[P], lo[P], uo[P], lout[P], uout[P], ujj[P];
  // u_jj of each strict position is gathered as soon as pmeta is in, in
  // the branch that loads it (every index valid: a diagonal position reads
  // entry 0 and ignores it)
  if (full) {
    wload<int, P>(pmeta + p0, mv);
    vload<T, P>(AL + p0, al);
    vload<T, P>(AU + p0, au);
    vload<T, P>(VLin + p0, lo);
    vload<T, P>(VUin + p0, uo);
#pragma unroll
    for (int e = 0; e < P; e++) ujj[e] = VUin[max(mv[e], 0)];
  } else {
#pragma unroll
    for (int e = 0; e < P; e++) {
      const int p = max(0, min(p0 + e, np - 1));
      mv[e] = pmeta[p];
      al[e] = AL[p];
      au[e] = AU[p];
      lo[e] = VLin[p];
      uo[e] = VUin[p];
    }
#pragma unroll
    for (int e = 0; e < P; e++) ujj[e] = VUin[max(mv[e], 0)];
  }
  // old l and u of positions p0 - W .. p0 - 1 from the W / P previous lanes,
  // every lane taking part in the shuffles; plo[k] holds position
  // p0 - W + k, valid when that lane is in this wave (lane >= its distance)
  T plo[W], puo[W];
#pragma unroll
  for (int k = 0; k < W; k++) {
    const int back = (W - k + P - 1) / P;  // lanes back
    const int e = k - (W - back * P);      // position within that lane
    plo[k] = __shfl_up(lo[e], back);
    puo[k] = __shfl_up(uo[e], back);
  }
  double d2 = 0.0;
  if (p0 < np) {
#pragma unroll
    for (int e = 0; e < P; e++) {
      const int p = p0 + e;
      const int m = mv[e];
      const T pu = T(1) * uo[e];
      if (p >= np) {
        uout[e] = lout[e] = T(0);
      } else if (m < 0) {
        const int i0 = -m - 1;
        T s = au[e];
        for (int q = i0; q < min(p, p0 - W); q++) s = s - VLin[q] * VUin[q];
#pragma unroll
        for (int k = -W; k < e; k++) {
          const int q = p0 + k;
          if (q >= i0) {
            T lq, uq;
            const int back = (-k + P - 1) / P;  // lanes back for k < 0
            if (k >= 0) {
              lq = lo[k < 0 ? 0 : k];
              uq = uo[k < 0 ? 0 : k];
            } else if (lane >= back) {
              lq = plo[k < 0 ? k + W : 0];
              uq = puo[k < 0 ? k + W : 0];
            } else {
              lq = VLin[q];
              uq = VUin[q];
            }
            s = s - lq * uq;
          }
        }
        s = s - pu;
        s = s + pu;
        uout[e] = s;
        lout[e] = lo[e];
        if (STEP) {
          const double d = (double)uo[e] - (double)s;
          d2 += d * d;
        }
      } else {
        T sL = al[e], sU = au[e];
        if (MATCH) {
          for (int t = mptr[p]; t < mptr[p + 1]; t++) {
            const int qa = mpairs[2 * t], qc = mpairs[2 * t + 1];
            sL = sL - VLin[qa] * VUin[qc];
            sU = sU - VLin[qc] * VUin[qa];
          }
        }
        sU = sU - pu;
        sU = sU + pu;
        uout[e] = sU;
        const T pl = lo[e] * ujj[e];
        sL = sL - pl;
        sL = sL + pl;
        const T t = sL / ujj[e];
        lout[e] = t;
        if (STEP) {
          const double d1 = (double)uo[e] - (double)sU;
          const double d = (double)lo[e] - (double)t;
          d2 += d1 * d1 + d * d;
        }
      }
    }
    if (full) {
      vstore<T, P>(VLout + p0, lout);
      vstore<T, P>(VUout + p0, uout);
    } else {
#pragma unroll
      for (int e = 0; e < P; e++)
        if (p0 + e < np) {
          VLout[p0 + e] = lout[e];
          VUout[p0 + e] = uout[e];
        }
    }
  }
  if (STEP) {
    __shared__ double sm[kSweepBS / 64];
    double tot = block_sum<kSweepBS>(d2, sm);
    if (threadIdx.x == 0) partial[b] = tot;
  }
}

// Sum of per-block partials in a fixed order (one block).
__global__ __launch_bounds__(256) void sum_partials_kernel(int nb, const double* __restrict__ partial,
                                                           double scale, double* __restrict__ out) {
  __shared__ double sm[4];
  double v = 0.0;
  for (int t = threadIdx.x; t < nb; t += 256) v += partial[t];
  double tot = block_sum<256>(v, sm);
  if (threadIdx.x == 0) *out = tot * scale;
}

// Factors from A: L0 = strict-lower(A) + I, U0 = upper(A).
template <typename T>
__global__ void parilu_reset_kernel(int nnz, const int* __restrict__ dst, const T* __restrict__ aval,
                                    T* __restrict__ L0, T* __restrict__ L1, T* __restrict__ U0,
                                    T* __restrict__ U1) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const int d = dst[k];
  const T v = aval[k];
  if (d >= 0) {
    L0[d] = v;
    L1[d] = v;
  } else {
    U0[-1 - d] = v;
    U1[-1 - d] = v;
  }
}

// a_ij into the symmetric kernels' position-ordered copies: a strict-lower
// entry to AL at its L position, an upper entry a_ij (j >= i) to AU at the
// L position (j, i), a diagonal entry to both.
template <typename T>
__global__ void parilu_scatter_sym_kernel(int nnz, const int* __restrict__ dst, const int* __restrict__ acol,
                                          const int* __restrict__ arowidx, const T* __restrict__ aval,
                                          T* __restrict__ AL, T* __restrict__ AU) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const int d = dst[k];
  const T v = aval[k];
  if (d >= 0) {
    AL[d] = v;
  } else {
    AU[-1 - d] = v;
    if (acol[k] == arowidx[k]) AL[-1 - d] = v;
  }
}

// The same initial factors for a structurally symmetric pattern, from the
// position-ordered copies of A the streaming sweep reads: L's and U's
// positions coincide (L in CSR, U in CSC), AL holds a_ij and AU a_ji at the
// position of (i, j), so L0 = AL with a unit diagonal and U0 = AU; reads and
// writes are coalesced (the scatter from A's CSR order is not).
template <typename T>
__global__ void parilu_reset_sym_kernel(int np, const int* __restrict__ pmeta, const T* __restrict__ AL,
                                        const T* __restrict__ AU, T* __restrict__ L0, T* __restrict__ L1,
                                        T* __restrict__ U0, T* __restrict__ U1) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const T l = pmeta[p] < 0 ? T(1) : AL[p];
  const T u = AU[p];
  L0[p] = l;
  L1[p] = l;
  U0[p] = u;
  U1[p] = u;
}

template <typename T>
__global__ void parilu_unitdiag_kernel(int n, const int* __restrict__ lrow, const int* __restrict__ lcol,
                                       T* __restrict__ L0, T* __restrict__ L1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int e = lrow[i + 1] - 1;
  if (e >= lrow[i] && lcol[e] == i) {
    L0[e] = T(1);
    L1[e] = T(1);
  }
}

// 16-bit storage: each factor value becomes its 16-bit word relative to
// the scale s of its row, times s (exact: s is a power of two), in place;
// L's diagonal (the unit) is left alone.  STEP: per-block partial sums of
// (old - rounded new)^2 over both factors.
template <typename TS, bool STEP>
__global__ __launch_bounds__(256) void parilu_round16_kernel(int nnzL, int nnzU, const int* __restrict__ lrowof,
                                                             const int* __restrict__ lcol,
                                                             const int* __restrict__ uri,
                                                             const float* __restrict__ rs, float* __restrict__ L,
                                                             float* __restrict__ U, const float* __restrict__ Lold,
                                                             const float* __restrict__ Uold,
                                                             double* __restrict__ partial) {
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int t = b * 256 + threadIdx.x;
  double d2 = 0.0;
  if (t < nnzL) {
    const int i = lrowof[t];
    if (lcol[t] != i) {
      const float sc = rs[i];
      const float v = dequant<float, TS>(quant<float, TS>(L[t], 1.0f / sc), sc);
      L[t] = v;
      if (STEP) {
        const double d = (double)Lold[t] - (double)v;
        d2 += d * d;
      }
    }
  }
  if (t < nnzU) {
    const float sc = rs[uri[t]];
    const float v = dequant<float, TS>(quant<float, TS>(U[t], 1.0f / sc), sc);
    U[t] = v;
    if (STEP) {
      const double d = (double)Uold[t] - (double)v;
      d2 += d * d;
    }
  }
  if (STEP) {
    __shared__ double sm[4];
    const double tot = block_sum<256>(d2, sm);
    if (threadIdx.x == 0) partial[b] = tot;
  }
}

// Residual of the fixed point on A's pattern: (LU)_ij - a_ij with the full
// merge (no add-back), summed as squares per block.
template <typename T>
__global__ __launch_bounds__(kSweepBS) void parilu_nonlinres_kernel(
    int nnz, const int* __restrict__ rowidx, const int* __restrict__ acol,
    const T* __restrict__ aval, const int* __restrict__ lrow, const int* __restrict__ lcol,
    const T* __restrict__ L, const int* __restrict__ ucp, const int* __restrict__ uri,
    const T* __restrict__ U, double* __restrict__ partial) {
  const int k = blockIdx.x * kSweepBS + threadIdx.x;
  double d2 = 0.0;
  if (k < nnz) {
    const int i = rowidx[k], j = acol[k];
    double s = 0.0;
    int il = lrow[i], iu = ucp[j];
    const int ile = lrow[i + 1], iue = ucp[j + 1];
    while (il < ile && iu < iue) {
      const int jl = lcol[il], ju = uri[iu];
      if (jl == ju) s += (double)L[il] * (double)U[iu];
      il += (jl <= ju);
      iu += (jl >= ju);
    }
    double d = s - (double)aval[k];
    d2 = d * d;
  }
  __shared__ double sm[kSweepBS / 64];
  double tot = block_sum<kSweepBS>(d2, sm);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Structurally symmetric sweep: one thread per L position p = (i, j),
// j <= i, updating l_ij (j < i) and u_ji from one merge, so both writes land
// at p (coalesced).  With a symmetric pattern the U column j holds rows k of
// row j's strict-lower pattern plus j, stored at row j's L positions, so
// every operand sits in row i or row j of the shared pattern:
//   l_ij = (a_ij - sum_{k in Lr(i) & Lr(j)} l_ik u_kj) / u_jj
//   u_ji =  a_ji - sum_{k in Lr(i) & Lr(j)} l_jk u_ki
//   u_ii =  a_ii - sum_{k in Lr(i)} l_ik u_ki                      (j == i)
// Matches are taken in ascending k, each rounded product subtracted, then
// the reference's closing term (k = min(i,j)) is subtracted and added back
// (src/parilu_v0_3.cpp:84-90), so results equal the general kernel bit for
// bit.  Rows with at most MAXL strict-lower entries keep both rows in
// registers (all loads issued at once); longer rows fall back to a loop.
template <typename T, bool STEP, int MAXL>
__global__ __launch_bounds__(kSweepBS) void parilu_sym_sweep_kernel(
    int np, const int* __restrict__ prow, const int* __restrict__ pcol,
    const int* __restrict__ lptr, const T* __restrict__ AL, const T* __restrict__ AU,
    const T* __restrict__ VLin, const T* __restrict__ VUin, T* __restrict__ VLout,
    T* __restrict__ VUout, double* __restrict__ partial) {
#pragma clang fp contract(off)
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int p = b * kSweepBS + threadIdx.x;
  double d2 = 0.0;
  if (p < np) {
    const int i = prow[p], j = pcol[p];
    const int i0 = lptr[i], i1 = lptr[i + 1] - 1;  // row i strict-lower [i0,i1), diag i1
    const int j0 = lptr[j], j1 = lptr[j + 1] - 1;
    const int ni = i1 - i0, nj = j1 - j0;
    T sL = AL[p];
    T sU = AU[p];
    if (ni <= MAXL && nj <= MAXL) {
      int ci[MAXL], cj[MAXL];
      T li[MAXL], ui[MAXL], lj[MAXL], uj[MAXL];
#pragma unroll
      for (int t = 0; t < MAXL; t++) {
        const bool oi = t < ni, oj = t < nj;
        ci[t] = oi ? pcol[i0 + t] : -1;
        li[t] = oi ? VLin[i0 + t] : T(0);
        ui[t] = oi ? VUin[i0 + t] : T(0);
        cj[t] = oj ? pcol[j0 + t] : -2;
        lj[t] = oj ? VLin[j0 + t] : T(0);
        uj[t] = oj ? VUin[j0 + t] : T(0);
      }
      // row i's columns ascend, so matches come out in ascending k
#pragma unroll
      for (int a = 0; a < MAXL; a++) {
#pragma unroll
        for (int c = 0; c < MAXL; c++) {
          if (ci[a] == cj[c]) {
            sL = sL - li[a] * uj[c];
            sU = sU - lj[c] * ui[a];
          }
        }
      }
    } else {
      int a = i0, c = j0;
      while (a < i1 && c < j1) {
        const int ka = pcol[a], kc = pcol[c];
        if (ka == kc) {
          sL = sL - VLin[a] * VUin[c];
          sU = sU - VLin[c] * VUin[a];
        }
        a += ka <= kc;
        c += ka >= kc;
      }
    }
    const T uold = VUin[p];
    const T pu = T(1) * uold;  // l_jj * u_ji (unit diagonal)
    sU = sU - pu;
    sU = sU + pu;
    VUout[p] = sU;
    if (STEP) {
      const double d = (double)uold - (double)sU;
      d2 = d * d;
    }
    if (j < i) {
      const T ujj = VUin[j1];
      const T lold = VLin[p];
      const T pl = lold * ujj;
      sL = sL - pl;
      sL = sL + pl;
      const T t = sL / ujj;
      VLout[p] = t;
      if (STEP) {
        const double d = (double)lold - (double)t;
        d2 += d * d;
      }
    }
  }
  if (STEP) {
    __shared__ double sm[kSweepBS / 64];
    double tot = block_sum<kSweepBS>(d2, sm);
    if (threadIdx.x == 0) partial[b] = tot;
  }
}

// Same update as parilu_sym_sweep_kernel with one thread per row: row i's
// positions (strict-lower entries and the diagonal) are loaded once into
// registers and every position of the row is updated from them, so only
// the partner rows j are gathered.  Requires every row to have at most MAXL
// strict-lower entries (checked on the host).  Identical arithmetic.
template <typename T, bool STEP, int MAXL>
__global__ __launch_bounds__(kSweepBS) void parilu_symrow_sweep_kernel(
    int n, const int* __restrict__ pcol, const int* __restrict__ lptr,
    const T* __restrict__ AL, const T* __restrict__ AU, const T* __restrict__ VLin,
    const T* __restrict__ VUin, T* __restrict__ VLout, T* __restrict__ VUout,
    double* __restrict__ partial) {
#pragma clang fp contract(off)
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int i = b * kSweepBS + threadIdx.x;
  double d2 = 0.0;
  if (i < n) {
    const int i0 = lptr[i], i1 = lptr[i + 1] - 1;  // strict-lower [i0,i1), diagonal at i1
    const int ni = i1 - i0;
    int ci[MAXL];
    T li[MAXL], ui[MAXL];
#pragma unroll
    for (int t = 0; t < MAXL; t++) {
      const bool o = t < ni;
      ci[t] = o ? pcol[i0 + t] : -1;
      li[t] = o ? VLin[i0 + t] : T(0);
      ui[t] = o ? VUin[i0 + t] : T(0);
    }
    // diagonal: u_ii = a_ii - sum_k l_ik u_ki, closing term 1*u_ii
    {
      T s = AU[i1];
#pragma unroll
      for (int t = 0; t < MAXL; t++)
        if (t < ni) s = s - li[t] * ui[t];
      const T uold = VUin[i1];
      const T pu = T(1) * uold;
      s = s - pu;
      s = s + pu;
      VUout[i1] = s;
      if (STEP) {
        const double d = (double)uold - (double)s;
        d2 += d * d;
      }
    }
    // strict-lower positions (i, j): l_ij and u_ji
#pragma unroll
    for (int t = 0; t < MAXL; t++) {
      if (t < ni) {
        const int p = i0 + t, j = ci[t];
        const int j0 = lptr[j], j1 = lptr[j + 1] - 1;
        const int nj = j1 - j0;
        T sL = AL[p], sU = AU[p];
        // partner row j: its strict-lower columns are all < j, so only row
        // i's entries before t can match
#pragma unroll
        for (int c = 0; c < MAXL; c++) {
          if (c < nj) {
            const int kc = pcol[j0 + c];
            const T ljc = VLin[j0 + c], ujc = VUin[j0 + c];
#pragma unroll
            for (int a = 0; a < MAXL; a++) {
              if (a < t && ci[a] == kc) {
                sL = sL - li[a] * ujc;
                sU = sU - ljc * ui[a];
              }
            }
          }
        }
        const T uold = ui[t];
        const T pu = T(1) * uold;
        sU = sU - pu;
        sU = sU + pu;
        VUout[p] = sU;
        const T ujj = VUin[j1];
        const T lold = li[t];
        const T pl = lold * ujj;
        sL = sL - pl;
        sL = sL + pl;
        const T tv = sL / ujj;
        VLout[p] = tv;
        if (STEP) {
          const double d = (double)uold - (double)sU, e = (double)lold - (double)tv;
          d2 += d * d + e * e;
        }
      }
    }
  }
  if (STEP) {
    __shared__ double sm[kSweepBS / 64];
    double tot = block_sum<kSweepBS>(d2, sm);
    if (threadIdx.x == 0) partial[b] = tot;
  }
}

template <typename To, typename From>
__global__ void convert_kernel(long n, const From* __restrict__ in, To* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = static_cast<To>(in[i]);
}

// ---------------------------------------------------------------- host side
// Power-of-two scale of each row for 16-bit factor storage: 2^e with
// 2^e <= max_j |a_ij| < 2^(e+1) (1 for a zero row), e clamped to the FP32
// normal range.
static void scalar_row_scales(int n, const int* row, const double* val, std::vector<float>& rs) {
  rs.assign(n, 1.0f);
  for (int i = 0; i < n; i++) {
    double m = 0.0;
    for (int k = row[i]; k < row[i + 1]; k++) m = std::max(m, std::fabs(val[k]));
    if (m > 0.0 && std::isfinite(m)) {
      int e = 0;
      (void)std::frexp(m, &e);  // m = f 2^e, f in [0.5, 1)
      e = std::max(-126, std::min(127, e - 1));
      rs[i] = std::ldexp(1.0f, e);
    }
  }
}

static int build_pattern(const data_d_matrix* A, ParILU* P) {
  const int n = A->num_rows;
  LACE_REQUIRE(A->num_rows == A->num_cols, DEV_ERR_ILLEGAL_VALUE, "ParILU needs a square matrix");
  LACE_REQUIRE(A->row && A->col && A->val, DEV_ERR_INVALID_PTR, "ParILU: A has null arrays");
  const int nnz = A->row[n];
  // every row must be sorted and hold its diagonal (the merge relies on the
  // diagonal closing each L row and U column)
  for (int i = 0; i < n; i++) {
    bool diag = false;
    for (int k = A->row[i]; k < A->row[i + 1]; k++) {
      if (k > A->row[i]) LACE_REQUIRE(A->col[k] > A->col[k - 1], DEV_ERR_ILLEGAL_VALUE,
                                      "ParILU: row %d is not sorted / has duplicates", i);
      diag |= A->col[k] == i;
    }
    LACE_REQUIRE(diag, DEV_ERR_ILLEGAL_VALUE, "ParILU: row %d has no diagonal entry", i);
  }
  P->n = n;
  P->nnz = nnz;
  // L (CSRL): per row, strict-lower then diagonal
  P->h_lrow.assign(n + 1, 0);
  P->h_lcol.clear();
  std::vector<int> dst(nnz);
  std::vector<int> ucount(n + 1, 0);
  for (int i = 0; i < n; i++) {
    P->h_lrow[i] = (int)P->h_lcol.size();
    for (int k = A->row[i]; k < A->row[i + 1]; k++) {
      int j = A->col[k];
      if (j < i) {
        dst[k] = (int)P->h_lcol.size();
        P->h_lcol.push_back(j);
      } else {
        ucount[j + 1]++;
      }
    }
    P->h_lcol.push_back(i);
  }
  P->h_lrow[n] = (int)P->h_lcol.size();
  P->nnzL = P->h_lrow[n];
  // U (CSCU): counting sort of upper entries by column, rows ascending
  for (int j = 0; j < n; j++) ucount[j + 1] += ucount[j];
  P->h_ucp = ucount;
  P->nnzU = ucount[n];
  P->h_uri.assign(P->nnzU, 0);
  std::vector<int> fill(ucount.begin(), ucount.end() - 1);
  for (int i = 0; i < n; i++)
    for (int k = A->row[i]; k < A->row[i + 1]; k++) {
      int j = A->col[k];
      if (j >= i) {
        int p = fill[j]++;
        P->h_uri[p] = i;
        dst[k] = -1 - p;
      }
    }
  std::vector<int> rowidx(nnz);
  for (int i = 0; i < n; i++)
    for (int k = A->row[i]; k < A->row[i + 1]; k++) rowidx[k] = i;
  double s = 0.0;
  for (int k = 0; k < nnz; k++) s += A->val[k] * A->val[k];
  P->anorm = std::sqrt(s);
  if (P->s16()) {
    std::vector<float> rs;
    scalar_row_scales(n, A->row, A->val, rs);
    LACE_TRY(upload(P->rs, rs.data(), sizeof(float) * n));
    std::vector<int> lr(P->nnzL);
    for (int i = 0; i < n; i++)
      for (int p = P->h_lrow[i]; p < P->h_lrow[i + 1]; p++) lr[p] = i;
    LACE_TRY(upload(P->lrowof, lr.data(), sizeof(int) * P->nnzL));
  }
  // symmetric pattern: the CSCU arrays of U equal the CSRL arrays of L
  P->sym = P->h_lrow == P->h_ucp && P->h_lcol == P->h_uri;
  if (P->sym) {
    std::vector<int> prow(P->nnzL);
    P->maxl = 0;
    for (int i = 0; i < n; i++) {
      for (int p = P->h_lrow[i]; p < P->h_lrow[i + 1]; p++) prow[p] = i;
      P->maxl = std::max(P->maxl, P->h_lrow[i + 1] - P->h_lrow[i] - 1);
    }
    std::vector<double> al(P->nnzL), au(P->nnzL);
    for (int i = 0; i < n; i++)
      for (int k = A->row[i]; k < A->row[i + 1]; k++) {
        const int d = dst[k], j = A->col[k];
        if (d >= 0) {
          al[d] = A->val[k];                    // a_ij, j < i, at L position (i,j)
        } else {
          au[-1 - d] = A->val[k];               // a_ij, j >= i, at L position (j,i)
          if (j == i) al[-1 - d] = A->val[k];
        }
      }
    LACE_TRY(upload(P->prow, prow.data(), sizeof(int) * P->nnzL));
    // static merge data for the streaming kernel: per position the index of
    // u_jj (strict) or -(row start + 1) (diagonal); the strict positions'
    // own matches k < j in rows i and j, ascending k, as (q_i, q_j) pairs
    {
      std::vector<int> meta(P->nnzL), mptr(P->nnzL + 1, 0), pairs;
      for (int i = 0; i < n; i++) {
        const int i0 = P->h_lrow[i], i1 = P->h_lrow[i + 1] - 1;
        for (int p = i0; p <= i1; p++) {
          mptr[p] = (int)(pairs.size() / 2);
          if (p == i1) {
            meta[p] = -(i0 + 1);
            continue;
          }
          const int j = P->h_lcol[p];
          const int j0 = P->h_lrow[j], j1 = P->h_lrow[j + 1] - 1;
          meta[p] = j1;
          int a = i0, c = j0;
          while (a < p && c < j1) {
            const int ka = P->h_lcol[a], kc = P->h_lcol[c];
            if (ka == kc) {
              pairs.push_back(a);
              pairs.push_back(c);
            }
            a += ka <= kc;
            c += ka >= kc;
          }
        }
      }
      mptr[P->nnzL] = (int)(pairs.size() / 2);
      LACE_TRY(upload(P->pmeta, meta.data(), sizeof(int) * P->nnzL));
      P->strict_matches = !pairs.empty();
      if (P->strict_matches) {
        LACE_TRY(upload(P->mptr, mptr.data(), sizeof(int) * (P->nnzL + 1)));
        LACE_TRY(upload(P->mpairs, pairs.data(), sizeof(int) * pairs.size()));
      }
    }
    const size_t esz = P->elem_size();
    if (P->prec == LACE_F64) {
      LACE_TRY(upload(P->AL, al.data(), sizeof(double) * P->nnzL));
      LACE_TRY(upload(P->AU, au.data(), sizeof(double) * P->nnzL));
    } else {
      std::vector<float> alf(al.begin(), al.end()), auf(au.begin(), au.end());
      LACE_TRY(upload(P->AL, alf.data(), esz * P->nnzL));
      LACE_TRY(upload(P->AU, auf.data(), esz * P->nnzL));
    }
  }

  LACE_TRY(upload(P->arow, A->row, sizeof(int) * (n + 1)));
  LACE_TRY(upload(P->acol, A->col, sizeof(int) * nnz));
  LACE_TRY(upload(P->arowidx, rowidx.data(), sizeof(int) * nnz));
  LACE_TRY(upload(P->dst, dst.data(), sizeof(int) * nnz));
  LACE_TRY(upload(P->lrow, P->h_lrow.data(), sizeof(int) * (n + 1)));
  LACE_TRY(upload(P->lcol, P->h_lcol.data(), sizeof(int) * P->nnzL));
  LACE_TRY(upload(P->ucp, P->h_ucp.data(), sizeof(int) * (n + 1)));
  LACE_TRY(upload(P->uri, P->h_uri.data(), sizeof(int) * P->nnzU));
  // A values in the iteration precision
  DevBuf a64;
  LACE_TRY(upload(a64, A->val, sizeof(double) * nnz));
  const size_t es = P->elem_size();
  LACE_TRY(P->aval.alloc(es * nnz));
  if (P->prec == LACE_F64) {
    LACE_HIP(hipMemcpyAsync(P->aval.p, a64.p, sizeof(double) * nnz, hipMemcpyDeviceToDevice, stream()));
  } else {
    convert_kernel<float, double><<<ceil_div(nnz, 256), 256, 0, stream()>>>(nnz, a64.as<double>(),
                                                                          P->aval.as<float>());
    LACE_HIP(hipGetLastError());
  }
  for (int b = 0; b < 2; b++) {
    LACE_TRY(P->lval[b].alloc(es * P->nnzL));
    LACE_TRY(P->uval[b].alloc(es * P->nnzU));
  }
  const int nb = ceil_div(nnz, kSweepBS);
  LACE_TRY(P->partial.alloc(sizeof(double) * (nb + 1)));
  LACE_TRY(P->scalar.alloc(sizeof(double) * 4));
  LACE_HIP(hipStreamSynchronize(stream()));
  return DEV_SUCCESS;
}

int ParILU::reset() {
  const int nb = ceil_div(nnz, 256);
  if (sym && pmeta.p && AL.p && AU.p) {
    const int nbl = ceil_div(nnzL, 256);
    if (prec == LACE_F64)
      parilu_reset_sym_kernel<double><<<nbl, 256, 0, stream()>>>(
          nnzL, pmeta.as<int>(), AL.as<double>(), AU.as<double>(), lval[0].as<double>(), lval[1].as<double>(),
          uval[0].as<double>(), uval[1].as<double>());
    else
      parilu_reset_sym_kernel<float><<<nbl, 256, 0, stream()>>>(
          nnzL, pmeta.as<int>(), AL.as<float>(), AU.as<float>(), lval[0].as<float>(), lval[1].as<float>(),
          uval[0].as<float>(), uval[1].as<float>());
  } else if (prec == LACE_F64) {
    parilu_reset_kernel<double><<<nb, 256, 0, stream()>>>(nnz, dst.as<int>(), aval.as<double>(),
                                                          lval[0].as<double>(), lval[1].as<double>(),
                                                          uval[0].as<double>(), uval[1].as<double>());
    parilu_unitdiag_kernel<double><<<ceil_div(n, 256), 256, 0, stream()>>>(
        n, lrow.as<int>(), lcol.as<int>(), lval[0].as<double>(), lval[1].as<double>());
  } else {
    parilu_reset_kernel<float><<<nb, 256, 0, stream()>>>(nnz, dst.as<int>(), aval.as<float>(),
                                                         lval[0].as<float>(), lval[1].as<float>(),
                                                         uval[0].as<float>(), uval[1].as<float>());
    parilu_unitdiag_kernel<float><<<ceil_div(n, 256), 256, 0, stream()>>>(
        n, lrow.as<int>(), lcol.as<int>(), lval[0].as<float>(), lval[1].as<float>());
  }
  LACE_HIP(hipGetLastError());
  cur = 0;
  csr_dirty = true;
  if (s16()) LACE_TRY(round16(true, false));
  return DEV_SUCCESS;
}

int ParILU::set_values(const double* vals) {
  LACE_REQUIRE(vals, DEV_ERR_INVALID_PTR, "ParILU set_values: null values");
  double s = 0.0;
  for (int k = 0; k < nnz; k++) s += vals[k] * vals[k];
  anorm = std::sqrt(s);
  if (s16()) {
    std::vector<int> arow(n + 1);
    LACE_HIP(hipMemcpyAsync(arow.data(), this->arow.p, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, stream()));
    LACE_HIP(hipStreamSynchronize(stream()));
    std::vector<float> h_rs;
    scalar_row_scales(n, arow.data(), vals, h_rs);
    LACE_HIP(hipMemcpyAsync(rs.p, h_rs.data(), sizeof(float) * n, hipMemcpyHostToDevice, stream()));
    LACE_HIP(hipStreamSynchronize(stream()));
  }
  DevBuf a64;
  LACE_TRY(upload(a64, vals, sizeof(double) * nnz));
  const int nb = ceil_div(nnz, 256);
  if (prec == LACE_F64) {
    LACE_HIP(hipMemcpyAsync(aval.p, a64.p, sizeof(double) * nnz, hipMemcpyDeviceToDevice, stream()));
    if (sym)
      parilu_scatter_sym_kernel<double><<<nb, 256, 0, stream()>>>(nnz, dst.as<int>(), acol.as<int>(),
                                                                   arowidx.as<int>(), aval.as<double>(),
                                                                   AL.as<double>(), AU.as<double>());
  } else {
    convert_kernel<float, double><<<nb, 256, 0, stream()>>>(nnz, a64.as<double>(), aval.as<float>());
    if (sym)
      parilu_scatter_sym_kernel<float><<<nb, 256, 0, stream()>>>(nnz, dst.as<int>(), acol.as<int>(),
                                                                  arowidx.as<int>(), aval.as<float>(),
                                                                  AL.as<float>(), AU.as<float>());
  }
  LACE_HIP(hipGetLastError());
  // a64 is released on return: finish the copies that read it first
  LACE_HIP(hipStreamSynchronize(stream()));
  return DEV_SUCCESS;
}

template <typename T, int MAXL>
static void launch_sym(ParILU* P, bool with_step, int in, int out, int nb) {
  if (with_step)
    parilu_sym_sweep_kernel<T, true, MAXL><<<nb, kSweepBS, 0, stream()>>>(
        P->nnzL, P->prow.as<int>(), P->lcol.as<int>(), P->lrow.as<int>(), P->AL.as<T>(),
        P->AU.as<T>(), P->lval[in].as<T>(), P->uval[in].as<T>(), P->lval[out].as<T>(),
        P->uval[out].as<T>(), P->partial.as<double>());
  else
    parilu_sym_sweep_kernel<T, false, MAXL><<<nb, kSweepBS, 0, stream()>>>(
        P->nnzL, P->prow.as<int>(), P->lcol.as<int>(), P->lrow.as<int>(), P->AL.as<T>(),
        P->AU.as<T>(), P->lval[in].as<T>(), P->uval[in].as<T>(), P->lval[out].as<T>(),
        P->uval[out].as<T>(), nullptr);
}

template <typename T>
static int sweep_t(ParILU* P, bool with_step) {
  const int in = P->cur, out = 1 - P->cur;
  const bool stream4 = P->sym && (P->variant == 0 || P->variant == 5);
  // the four-position kernel's launch carries the events itself
  TimedRegion tr("parilu_sweep", 0.0, P->sweep_bytes(), stream4);
  // symmetric patterns default to the four-position streaming kernel
  // (128^3: FP32 38.8 us vs 43.7 us for one position per thread)
  if (stream4) {
    // four positions per thread in FP32, two in FP64 (one 16-byte word per
    // array either way)
    constexpr int PP = sizeof(T) == 8 ? 2 : 4;
    const int nb = ceil_div(P->nnzL, kSweepBS * PP);
    auto go = [&](auto kern) {
      tr.launch(kern, dim3(nb), dim3(kSweepBS), P->nnzL, P->pmeta.as<int>(), P->mptr.as<int>(),
                P->mpairs.as<int>(), P->AL.as<T>(), P->AU.as<T>(), P->lval[in].as<T>(),
                P->uval[in].as<T>(), P->lval[out].as<T>(), P->uval[out].as<T>(), P->partial.as<double>());
    };
    if (P->strict_matches) {
      if (with_step) go(parilu_stream4_sweep_kernel<T, true, true, PP>);
      else go(parilu_stream4_sweep_kernel<T, false, true, PP>);
    } else {
      if (with_step) go(parilu_stream4_sweep_kernel<T, true, false, PP>);
      else go(parilu_stream4_sweep_kernel<T, false, false, PP>);
    }
    LACE_HIP(hipGetLastError());
    P->last_nb = nb;
    P->cur = out;
    P->csr_dirty = true;
    return DEV_SUCCESS;
  }
  if (P->sym && P->variant == 4) {
    const int nb = ceil_div(P->nnzL, kSweepBS);
    auto go = [&](auto kern) {
      kern<<<nb, kSweepBS, 0, stream()>>>(P->nnzL, P->pmeta.as<int>(), P->mptr.as<int>(),
                                          P->mpairs.as<int>(), P->AL.as<T>(), P->AU.as<T>(),
                                          P->lval[in].as<T>(), P->uval[in].as<T>(),
                                          P->lval[out].as<T>(), P->uval[out].as<T>(),
                                          P->partial.as<double>());
    };
    if (P->strict_matches) {
      if (with_step) go(parilu_stream_sweep_kernel<T, true, true>);
      else go(parilu_stream_sweep_kernel<T, false, true>);
    } else {
      if (with_step) go(parilu_stream_sweep_kernel<T, true, false>);
      else go(parilu_stream_sweep_kernel<T, false, false>);
    }
    LACE_HIP(hipGetLastError());
    P->last_nb = nb;
    P->cur = out;
    P->csr_dirty = true;
    return DEV_SUCCESS;
  }
  if (P->sym && P->maxl <= 4 && P->variant == 3) {
    const int nb = ceil_div(P->n, kSweepBS);
    if (with_step)
      parilu_symrow_sweep_kernel<T, true, 4><<<nb, kSweepBS, 0, stream()>>>(
          P->n, P->lcol.as<int>(), P->lrow.as<int>(), P->AL.as<T>(), P->AU.as<T>(),
          P->lval[in].as<T>(), P->uval[in].as<T>(), P->lval[out].as<T>(), P->uval[out].as<T>(),
          P->partial.as<double>());
    else
      parilu_symrow_sweep_kernel<T, false, 4><<<nb, kSweepBS, 0, stream()>>>(
          P->n, P->lcol.as<int>(), P->lrow.as<int>(), P->AL.as<T>(), P->AU.as<T>(),
          P->lval[in].as<T>(), P->uval[in].as<T>(), P->lval[out].as<T>(), P->uval[out].as<T>(),
          nullptr);
    LACE_HIP(hipGetLastError());
    P->last_nb = nb;
    P->cur = out;
    P->csr_dirty = true;
    return DEV_SUCCESS;
  }
  if (P->sym && P->variant != 1) {
    const int nb = ceil_div(P->nnzL, kSweepBS);
    if (P->maxl <= 4) launch_sym<T, 4>(P, with_step, in, out, nb);
    else launch_sym<T, 8>(P, with_step, in, out, nb);
    LACE_HIP(hipGetLastError());
    P->last_nb = nb;
    P->cur = out;
    P->csr_dirty = true;
    return DEV_SUCCESS;
  }
  const int nb = ceil_div(P->nnz, kSweepBS);
  P->last_nb = nb;
  P->csr_dirty = true;
  if (with_step)
    parilu_sweep_kernel<T, true><<<nb, kSweepBS, 0, stream()>>>(
        P->nnz, P->arowidx.as<int>(), P->acol.as<int>(), P->aval.as<T>(), P->lrow.as<int>(),
        P->lcol.as<int>(), P->lval[in].as<T>(), P->lval[out].as<T>(), P->ucp.as<int>(),
        P->uri.as<int>(), P->uval[in].as<T>(), P->uval[out].as<T>(), P->partial.as<double>());
  else
    parilu_sweep_kernel<T, false><<<nb, kSweepBS, 0, stream()>>>(
        P->nnz, P->arowidx.as<int>(), P->acol.as<int>(), P->aval.as<T>(), P->lrow.as<int>(),
        P->lcol.as<int>(), P->lval[in].as<T>(), P->lval[out].as<T>(), P->ucp.as<int>(),
        P->uri.as<int>(), P->uval[in].as<T>(), P->uval[out].as<T>(), nullptr);
  LACE_HIP(hipGetLastError());
  P->cur = out;
  return DEV_SUCCESS;
}

int ParILU::round16(bool both_buffers, bool with_step) {
  const int nb = ceil_div(std::max(nnzL, nnzU), 256);
  for (int bsel = 0; bsel < (both_buffers ? 2 : 1); bsel++) {
    const int buf = both_buffers ? bsel : cur;
    float* Lv = lval[buf].as<float>();
    float* Uv = uval[buf].as<float>();
    const float* Lo = lval[1 - buf].as<float>();
    const float* Uo = uval[1 - buf].as<float>();
    const int* lr = lrowof.as<int>();
    const int* lc = lcol.as<int>();
    const int* ur = uri.as<int>();
    const float* r = rs.as<float>();
    double* pt = partial.as<double>();
    if (prec == LACE_BF16) {
      if (with_step)
        parilu_round16_kernel<uint16_t, true><<<nb, 256, 0, stream()>>>(nnzL, nnzU, lr, lc, ur, r, Lv, Uv, Lo, Uo, pt);
      else
        parilu_round16_kernel<uint16_t, false><<<nb, 256, 0, stream()>>>(nnzL, nnzU, lr, lc, ur, r, Lv, Uv, Lo, Uo,
                                                                       pt);
    } else {
      if (with_step)
        parilu_round16_kernel<_Float16, true><<<nb, 256, 0, stream()>>>(nnzL, nnzU, lr, lc, ur, r, Lv, Uv, Lo, Uo, pt);
      else
        parilu_round16_kernel<_Float16, false><<<nb, 256, 0, stream()>>>(nnzL, nnzU, lr, lc, ur, r, Lv, Uv, Lo, Uo,
                                                                       pt);
    }
  }
  LACE_HIP(hipGetLastError());
  if (with_step) last_nb = nb;
  return DEV_SUCCESS;
}

int ParILU::sweep(bool with_step) {
  if (prec == LACE_F64) return sweep_t<double>(this, with_step);
  if (!s16()) return sweep_t<float>(this, with_step);
  // FP32 sweep, then the rounding to 16-bit storage (which forms the step
  // from the stored values)
  LACE_TRY(sweep_t<float>(this, false));
  return round16(false, with_step);
}

int ParILU::read_step(double* step) {
  const int nb = last_nb;
  sum_partials_kernel<<<1, 256, 0, stream()>>>(nb, partial.as<double>(), 1.0 / anorm,
                                               scalar.as<double>());
  LACE_HIP(hipGetLastError());
  LACE_HIP(hipMemcpyAsync(step, scalar.p, sizeof(double), hipMemcpyDeviceToHost, stream()));
  LACE_HIP(hipStreamSynchronize(stream()));
  return DEV_SUCCESS;
}

int ParILU::nonlinres(double* out) {
  const int nb = ceil_div(nnz, kSweepBS);
  if (prec == LACE_F64)
    parilu_nonlinres_kernel<double><<<nb, kSweepBS, 0, stream()>>>(
        nnz, arowidx.as<int>(), acol.as<int>(), aval.as<double>(), lrow.as<int>(), lcol.as<int>(),
        lval[cur].as<double>(), ucp.as<int>(), uri.as<int>(), uval[cur].as<double>(),
        partial.as<double>());
  else
    parilu_nonlinres_kernel<float><<<nb, kSweepBS, 0, stream()>>>(
        nnz, arowidx.as<int>(), acol.as<int>(), aval.as<float>(), lrow.as<int>(), lcol.as<int>(),
        lval[cur].as<float>(), ucp.as<int>(), uri.as<int>(), uval[cur].as<float>(),
        partial.as<double>());
  LACE_HIP(hipGetLastError());
  double s = 0.0;
  sum_partials_kernel<<<1, 256, 0, stream()>>>(nb, partial.as<double>(), 1.0, scalar.as<double>());
  LACE_HIP(hipMemcpyAsync(&s, scalar.p, sizeof(double), hipMemcpyDeviceToHost, stream()));
  LACE_HIP(hipStreamSynchronize(stream()));
  *out = std::sqrt(s);
  return DEV_SUCCESS;
}

double ParILU::sweep_bytes() const {
  // Algorithmic traffic of one sweep: per nonzero of A its value and column
  // index, the previous and the new factor value; plus the row pointers:
  // nnz*(s_A + 4 + 2 s_F) + 4(n+1).
  const double es = (double)elem_size();
  return (double)nnz * (es + 4.0 + 2.0 * es) + 4.0 * (n + 1);
}

int ParILU::get(data_d_matrix* L, data_d_matrix* U) {
  // L as CSRL with unit diagonal, U as CSCU (src/parilu_v0_3.cpp:23-27)
  std::vector<double> lv(nnzL), uv(nnzU);
  LACE_TRY(values_f64(lv.data(), uv.data()));
  data_zmfree(L);
  data_zmfree(U);
  L->storage_type = Magma_CSRL;
  L->diagorder_type = Magma_UNITY;
  L->fill_mode = MagmaLower;
  L->major = MagmaRowMajor;
  L->num_rows = L->num_cols = n;
  L->nnz = L->true_nnz = nnzL;
  LACE_TRY(host_calloc(&L->row, n + 1));
  LACE_TRY(host_calloc(&L->col, nnzL));
  LACE_TRY(host_calloc(&L->val, nnzL));
  memcpy(L->row, h_lrow.data(), sizeof(int) * (n + 1));
  memcpy(L->col, h_lcol.data(), sizeof(int) * nnzL);
  memcpy(L->val, lv.data(), sizeof(double) * nnzL);
  U->storage_type = Magma_CSCU;
  U->diagorder_type = Magma_VALUE;
  U->fill_mode = MagmaUpper;
  U->major = MagmaColMajor;
  U->num_rows = U->num_cols = n;
  U->nnz = U->true_nnz = nnzU;
  LACE_TRY(host_calloc(&U->row, n + 1));
  LACE_TRY(host_calloc(&U->col, nnzU));
  LACE_TRY(host_calloc(&U->val, nnzU));
  memcpy(U->row, h_ucp.data(), sizeof(int) * (n + 1));
  memcpy(U->col, h_uri.data(), sizeof(int) * nnzU);
  memcpy(U->val, uv.data(), sizeof(double) * nnzU);
  return DEV_SUCCESS;
}

int ParILU::values_f64(double* lv, double* uv) {
  if (prec == LACE_F64) {
    LACE_HIP(hipMemcpyAsync(lv, lval[cur].p, sizeof(double) * nnzL, hipMemcpyDeviceToHost, stream()));
    LACE_HIP(hipMemcpyAsync(uv, uval[cur].p, sizeof(double) * nnzU, hipMemcpyDeviceToHost, stream()));
    LACE_HIP(hipStreamSynchronize(stream()));
  } else {
    std::vector<float> lf(nnzL), uf(nnzU);
    LACE_HIP(hipMemcpyAsync(lf.data(), lval[cur].p, sizeof(float) * nnzL, hipMemcpyDeviceToHost, stream()));
    LACE_HIP(hipMemcpyAsync(uf.data(), uval[cur].p, sizeof(float) * nnzU, hipMemcpyDeviceToHost, stream()));
    LACE_HIP(hipStreamSynchronize(stream()));
    for (int k = 0; k < nnzL; k++) lv[k] = lf[k];
    for (int k = 0; k < nnzU; k++) uv[k] = uf[k];
  }
  return DEV_SUCCESS;
}

int parilu_create(const data_d_matrix* A, lace_prec_t prec, ParILU** out) {
  LACE_TRY(ensure_init());
  LACE_REQUIRE(prec == LACE_F64 || prec == LACE_F32 || prec == LACE_BF16 || prec == LACE_F16,
               DEV_ERR_NOT_SUPPORTED, "ParILU precision %d not supported", (int)prec);
  auto* P = new ParILU();
  P->prec = prec;
  int rc = build_pattern(A, P);
  if (rc == DEV_SUCCESS) rc = P->reset();
  if (rc == DEV_SUCCESS) rc = P->build_trsv();
  if (rc != DEV_SUCCESS) {
    delete P;
    return rc;
  }
  *out = P;
  return DEV_SUCCESS;
}

}  // namespace lace

using namespace lace;

// ---------------------------------------------------------------- C ABI
extern "C" {

int lace_parilu_create(const data_d_matrix* A, lace_prec_t prec, lace_parilu** out) {
  ParILU* P = nullptr;
  LACE_TRY(parilu_create(A, prec, &P));
  *out = reinterpret_cast<lace_parilu*>(P);
  return DEV_SUCCESS;
}

int lace_parilu_destroy(lace_parilu* P) {
  delete reinterpret_cast<ParILU*>(P);
  return DEV_SUCCESS;
}

int lace_parilu_reset(lace_parilu* P) {
  LACE_TRY(ensure_init());
  return reinterpret_cast<ParILU*>(P)->reset();
}

int lace_parilu_set_values(lace_parilu* P, const double* vals) {
  LACE_TRY(ensure_init());
  LACE_REQUIRE(P, DEV_ERR_INVALID_PTR, "lace_parilu_set_values: null handle");
  return reinterpret_cast<ParILU*>(P)->set_values(vals);
}

int lace_parilu_sweeps(lace_parilu* Ph, int nsweeps, double* step) {
  LACE_TRY(ensure_init());
  auto* P = reinterpret_cast<ParILU*>(Ph);
  for (int s = 0; s < nsweeps; s++) LACE_TRY(P->sweep(step != nullptr && s == nsweeps - 1));
  if (step) {
    if (nsweeps > 0) LACE_TRY(P->read_step(step));
    else *step = 0.0;
  }
  return DEV_SUCCESS;
}

int lace_parilu_get(lace_parilu* P, data_d_matrix* L, data_d_matrix* U) {
  LACE_TRY(ensure_init());
  return reinterpret_cast<ParILU*>(P)->get(L, U);
}

int lace_parilu_nonlinres(lace_parilu* P, double* nonlinres) {
  LACE_TRY(ensure_init());
  return reinterpret_cast<ParILU*>(P)->nonlinres(nonlinres);
}

double lace_parilu_sweep_bytes(const lace_parilu* P) {
  return reinterpret_cast<const ParILU*>(P)->sweep_bytes();
}

int lace_parilu_set_variant(lace_parilu* Ph, int general) {
  LACE_REQUIRE(general >= 0 && general <= 5, DEV_ERR_ILLEGAL_VALUE, "variant %d", general);
  reinterpret_cast<ParILU*>(Ph)->variant = general;
  return DEV_SUCCESS;
}

int lace_parilu_levels(const lace_parilu* Ph, int* nlev_L, int* nlev_U) {
  auto* P = reinterpret_cast<const ParILU*>(Ph);
  *nlev_L = P->lower.nlev;
  *nlev_U = P->upper.nlev;
  return DEV_SUCCESS;
}

int lace_parilu_forms(const lace_parilu* Ph, int upper, int* ndiag, int* ell_width) {
  LACE_REQUIRE(Ph && ndiag && ell_width, DEV_ERR_INVALID_PTR, "lace_parilu_forms: null argument");
  auto* P = reinterpret_cast<const ParILU*>(Ph);
  const TriPlan& T = upper ? P->upper : P->lower;
  *ndiag = T.dia.nd;
  *ell_width = T.ew;
  return DEV_SUCCESS;
}

int lace_parilu_tiles(const lace_parilu* Ph, int which, int* dims) {
  LACE_REQUIRE(Ph && dims, DEV_ERR_INVALID_PTR, "lace_parilu_tiles: null argument");
  LACE_REQUIRE(which >= 0 && which <= 3, DEV_ERR_ILLEGAL_VALUE, "lace_parilu_tiles: which = %d", which);
  auto* P = reinterpret_cast<const ParILU*>(Ph);
  const TriPlan* T[4] = {&P->lower, &P->upper, &P->lowerT, &P->upperT};
  for (int d = 0; d < 3; d++) dims[d] = T[which]->tile_dims[d];
  return DEV_SUCCESS;
}

// control/norms.cpp:301-405: res = ||A - LU||_F over LU's pattern,
// nonlinres = the same restricted to A's pattern.  Host evaluation (a
// reporting routine called once per factorization).  L: CSRL with unit
// diagonal (or strictly lower), U: CSR/CSRU or CSC/CSCU.
int data_zilures(data_d_matrix A, data_d_matrix L, data_d_matrix U, data_d_matrix* LU,
                 dataType* res, dataType* nonlinres) {
  data_d_matrix Lr, Ur;
  memset(&Lr, 0, sizeof(Lr));
  memset(&Ur, 0, sizeof(Ur));
  const int n = A.num_rows;
  // unit-diagonal L in CSR
  bool strict = true;
  for (int i = 0; i < n && strict; i++)
    for (int k = L.row[i]; k < L.row[i + 1]; k++)
      if (L.col[k] == i) strict = false;
  if (strict) {
    const int z = L.row[n] + n;
    Lr.storage_type = Magma_CSR;
    Lr.num_rows = Lr.num_cols = n;
    LACE_TRY(host_calloc(&Lr.row, n + 1));
    LACE_TRY(host_calloc(&Lr.col, z));
    LACE_TRY(host_calloc(&Lr.val, z));
    int p = 0;
    for (int i = 0; i < n; i++) {
      Lr.row[i] = p;
      for (int k = L.row[i]; k < L.row[i + 1]; k++) {
        Lr.col[p] = L.col[k];
        Lr.val[p++] = L.val[k];
      }
      Lr.col[p] = i;
      Lr.val[p++] = 1.0;
    }
    Lr.row[n] = p;
    Lr.nnz = p;
  } else {
    LACE_TRY(data_zmconvert(L, &Lr, Magma_CSR, Magma_CSR));
  }
  if (U.storage_type == Magma_CSC || U.storage_type == Magma_CSCU)
    LACE_TRY(data_zmconvert(U, &Ur, Magma_CSC, Magma_CSR));
  else
    LACE_TRY(data_zmconvert(U, &Ur, Magma_CSR, Magma_CSR));
  int rc = data_z_spmm(1.0, Lr, Ur, LU);
  data_zmfree(&Lr);
  data_zmfree(&Ur);
  LACE_TRY(rc);
  double r = 0.0, nl = 0.0;
  for (int i = 0; i < n; i++) {
    for (int k = A.row[i]; k < A.row[i + 1]; k++) {
      // LU rows are sorted: binary search for A's column
      int lo = LU->row[i], hi = LU->row[i + 1];
      int c = A.col[k];
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (LU->col[mid] < c) lo = mid + 1;
        else hi = mid;
      }
      if (lo < LU->row[i + 1] && LU->col[lo] == c) {
        double t = LU->val[lo] - A.val[k];
        LU->val[lo] = t;
        nl += t * t;
      }
    }
  }
  for (int k = 0; k < LU->row[n]; k++) r += LU->val[k] * LU->val[k];
  *res = std::sqrt(r);
  *nonlinres = std::sqrt(nl);
  return DEV_SUCCESS;
}

}  // extern "C"

// ---------------------------------------------------------------- drop-ins
namespace {

// Shared body of the ParILU drop-ins.  mode 0: stop when step <=
// reduction*Ares or after 20 sweeps (src/parilu_v0_3.cpp:44-45,62);
// mode 1: stop after log->maxSweeps sweeps or when step <= 4 eps
// (src/parilu_v0_3_prescribedSweeps.cpp:62); mode 2: stop when step <=
// reduction*Ares, the only rule of src/parilu_v0_2.cpp:39,61 (which has
// no sweep cap; this one stops at 1000).
int parilu_dropin(data_d_matrix* A, data_d_matrix* L, data_d_matrix* U, double reduction,
                  data_d_preconditioner_log* log, int mode) {
  LACE_TRY(ensure_init());
  L->diagorder_type = Magma_UNITY;
  LACE_TRY(data_zmconvert(*A, L, Magma_CSR, Magma_CSRL));
  U->diagorder_type = Magma_VALUE;
  LACE_TRY(data_zmconvert(*A, U, Magma_CSR, Magma_CSCU));
  double Ares = 0.0, Anl = 0.0;
  data_d_matrix LU;
  memset(&LU, 0, sizeof(LU));
  LACE_TRY(data_zilures(*A, *L, *U, &LU, &Ares, &Anl));
  data_zmfree(&LU);
  double anorm = 0.0;
  LACE_TRY(data_zfrobenius(*A, &anorm));
  LACE_TRY(data_rowindex(A, &A->rowidx));
  ParILU* P = nullptr;
  auto t0 = std::chrono::steady_clock::now();
  LACE_TRY(parilu_create(A, LACE_F64, &P));
  const double tol = reduction * Ares;
  int iter = 0;
  double step = 3.402823466e+38;  // FLT_MAX, as the reference starts
  int rc = DEV_SUCCESS;
  while (rc == DEV_SUCCESS) {
    if (mode == 0 && !(step > tol && iter < 20)) break;
    if (mode == 1 && !(iter < log->maxSweeps && step > 4.0 * 2.220446049250313e-16)) break;
    if (mode == 2 && !(step > tol && iter < 1000)) break;
    rc = P->sweep(true);
    if (rc == DEV_SUCCESS) rc = P->read_step(&step);
    iter++;
  }
  if (rc == DEV_SUCCESS) rc = P->get(L, U);
  auto t1 = std::chrono::steady_clock::now();
  delete P;
  LACE_TRY(rc);
  log->sweeps = iter;
  log->tol = tol;
  log->finalStep = step;
  log->A_Frobenius = anorm;
  log->precond_generation_time = std::chrono::duration<double>(t1 - t0).count();
  log->initial_residual = Ares;
  log->initial_nonlinear_residual = Anl;
  log->omp_num_threads = 1;
  return DEV_SUCCESS;
}

}  // namespace

extern "C" {

// src/parilu_v0_2.cpp:16-110: the same sweep, stop rule only, no log.
void data_PariLU_v0_2(data_d_matrix* A, data_d_matrix* L, data_d_matrix* U, dataType reduction) {
  data_d_preconditioner_log log;
  memset(&log, 0, sizeof(log));
  (void)parilu_dropin(A, L, U, reduction, &log, 2);
}

void data_PariLU_v0_3(data_d_matrix* A, data_d_matrix* L, data_d_matrix* U, dataType reduction,
                      data_d_preconditioner_log* log) {
  if (parilu_dropin(A, L, U, reduction, log, 0) != DEV_SUCCESS) log->sweeps = -1;
}

void data_PariLU_v0_3_prescribedSweeps(data_d_matrix* A, data_d_matrix* L, data_d_matrix* U,
                                       dataType reduction, data_d_preconditioner_log* log) {
  if (parilu_dropin(A, L, U, reduction, log, 1) != DEV_SUCCESS) log->sweeps = -1;
}

// nthreads was the CUDA block size of the single-block reference kernel
// (src/parilu_v0_3_gpu.cpp:203-209); the grid here is sized by nnz.
void data_PariLU_v0_3_gpu(data_d_matrix* A, data_d_matrix* L, data_d_matrix* U, dataType reduction,
                          data_d_preconditioner_log* log, int nthreads) {
  (void)nthreads;
  if (parilu_dropin(A, L, U, reduction, log, 0) != DEV_SUCCESS) log->sweeps = -1;
}

// The thread count is an input here (src/parilu_v0_3_gpu_prescribedSweeps.cpp:
// 140 reads log->omp_num_threads as the CUDA block size and leaves it), and
// drivers name their factor files after it, so it is kept.
void data_PariLU_v0_3_gpu_prescribedSweeps(data_d_matrix* A, data_d_matrix* L, data_d_matrix* U,
                                           dataType reduction, data_d_preconditioner_log* log) {
  const int threads = log->omp_num_threads;
  if (parilu_dropin(A, L, U, reduction, log, 1) != DEV_SUCCESS) log->sweeps = -1;
  log->omp_num_threads = threads;
}

}  // extern "C"
