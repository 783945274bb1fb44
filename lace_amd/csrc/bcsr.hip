// bcsr.hip -- block-sparse (BCSR) SpMV, block ParILU(0) sweeps and block
// triangular solves on the GPU.
//
// Reference: src/parilu_v0_3_bcsr.cpp:16-276 (block ParILU: S = A_IJ minus
// the merged products L_IK U_KJ, the last product added back, L_IJ = S
// inv(U_JJ), U_IJ = S; absolute stop tolerance), src/inverse.cpp:17 (block
// inverse), control/magma_zmconverter.cpp:1248-1445 (BCSRL with identity
// diagonal blocks, BCSCU = block transpose of BCSRU with block interiors
// kept as they are).
//
// Thread mapping: a group of BS consecutive lanes owns one block row of a
// block (lane r of the group computes row r), 64/BS groups per wave, groups
// never straddle a wave.  Full blocks a lane needs (U_KJ, U_JJ) are read by
// every lane of the group from the same addresses, which the texture path
// merges.  Arithmetic follows the CPU restatement operation by operation
// (products accumulated from zero in ascending index order, no fused
// multiply-add), so FP64 and FP32 factors are bit-identical to it.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "bcsr.hpp"
#include "common.hpp"
#include "device.hpp"

#pragma clang fp contract(off)

namespace lace {

constexpr int kBBS = 256;  // threads per block for every kernel here

template <typename TS>
constexpr bool is16() {
  return sizeof(TS) == 2;
}

// Gauss-Jordan inverse without pivoting, in the fixed order of the CPU
// restatement: pivot p scales its row, then every other row eliminates
// column p.
template <int BS, typename T>
__device__ __forceinline__ void binv(const T (&M)[BS * BS], T (&inv)[BS * BS]) {
  T a[BS * BS];
#pragma unroll
  for (int e = 0; e < BS * BS; e++) {
    a[e] = M[e];
    inv[e] = T(0);
  }
#pragma unroll
  for (int r = 0; r < BS; r++) inv[r * BS + r] = T(1);
#pragma unroll
  for (int p = 0; p < BS; p++) {
    const T iv = T(1) / a[p * BS + p];
#pragma unroll
    for (int c = 0; c < BS; c++) {
      a[p * BS + c] = a[p * BS + c] * iv;
      inv[p * BS + c] = inv[p * BS + c] * iv;
    }
#pragma unroll
    for (int r = 0; r < BS; r++) {
      if (r == p) continue;
      const T f = a[r * BS + p];
#pragma unroll
      for (int c = 0; c < BS; c++) {
        a[r * BS + c] = a[r * BS + c] - f * a[p * BS + c];
        inv[r * BS + c] = inv[r * BS + c] - f * inv[p * BS + c];
      }
    }
  }
}

// Full block q (row scales rs[kb*BS + m] for 16-bit storage).
template <int BS, typename TC, typename TS>
__device__ __forceinline__ void load_block(const TS* __restrict__ V, int q, const float* __restrict__ rs,
                                           int kb, TC (&out)[BS * BS]) {
#pragma unroll
  for (int m = 0; m < BS; m++) {
    const float s = is16<TS>() ? rs[kb * BS + m] : 1.0f;
#pragma unroll
    for (int c = 0; c < BS; c++) out[m * BS + c] = ldv<TC>(V, (size_t)q * BS * BS + m * BS + c, s);
  }
}

template <int BS, typename TC, typename TS>
__device__ __forceinline__ void load_row(const TS* __restrict__ V, int q, int r, float s, TC (&out)[BS]) {
#pragma unroll
  for (int c = 0; c < BS; c++) out[c] = ldv<TC>(V, (size_t)q * BS * BS + r * BS + c, s);
}

// Lane-group coordinates: group index (over the grid) and row in the block.
template <int BS>
struct Group {
  int g, r;
  bool valid;
  __device__ __forceinline__ Group(int count) {
    constexpr int PPW = 64 / BS;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x & 63;
    const int slot = lane / BS;
    r = lane - slot * BS;
    g = (blk * (kBBS / 64) + (int)(threadIdx.x >> 6)) * PPW + slot;
    valid = slot < PPW && g < count;
  }
};

template <int BS>
inline int group_blocks(long count) {
  const long per = (long)(kBBS / 64) * (64 / BS);
  return (int)std::max(1L, (count + per - 1) / per);
}

// ---------------------------------------------------------------- SpMV
// MODE 0: y = A x; MODE 1: y = b - A x.  Lane (I, r) sums block row I's
// row r over blocks in order, entries in column order.
template <typename TV, typename TX, typename TY, int BS, int MODE>
__global__ __launch_bounds__(kBBS) void bcsr_spmv_kernel(int nbr, const int* __restrict__ row,
                                                         const int* __restrict__ col,
                                                         const TV* __restrict__ val,
                                                         const TX* __restrict__ x, TY* __restrict__ y,
                                                         const TY* __restrict__ b) {
  using ACC = typename std::conditional<sizeof(TV) == 8 || sizeof(TY) == 8, double, float>::type;
  const Group<BS> G(nbr);
  if (!G.valid) return;
  const int I = G.g, r = G.r;
  const int k0 = row[I], k1 = row[I + 1];
  ACC acc = ACC(0);
  for (int k = k0; k < k1; k++) {
    const int K = col[k];
    const TV* v = val + (size_t)k * BS * BS + r * BS;
#pragma unroll
    for (int c = 0; c < BS; c++) acc = acc + (ACC)v[c] * (ACC)x[(size_t)K * BS + c];
  }
  const size_t i = (size_t)I * BS + r;
  if (MODE == 1) y[i] = (TY)((ACC)b[i] - acc);
  else y[i] = (TY)acc;
}

template <typename TV, typename TX, typename TY, int MODE>
static int launch_bspmv(const DevBCSR& A, const TX* x, TY* y, const TY* b) {
  const int* row = A.row.as<int>();
  const int* col = A.col.as<int>();
  const TV* val = A.val.as<TV>();
#define LACE_BSPMV(BSV)                                                                    \
  case BSV:                                                                                \
    bcsr_spmv_kernel<TV, TX, TY, BSV, MODE>                                                \
        <<<group_blocks<BSV>(A.nbr), kBBS, 0, stream()>>>(A.nbr, row, col, val, x, y, b); \
    break;
  switch (A.bs) {
    LACE_BSPMV(1) LACE_BSPMV(2) LACE_BSPMV(3) LACE_BSPMV(4)
    LACE_BSPMV(5) LACE_BSPMV(6) LACE_BSPMV(7) LACE_BSPMV(8)
    default: LACE_REQUIRE(false, DEV_ERR_NOT_SUPPORTED, "BCSR block size %d", A.bs);
  }
#undef LACE_BSPMV
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

int DevBCSR::create(const data_d_matrix* A, lace_prec_t p) {
  LACE_TRY(ensure_init());
  LACE_REQUIRE(A && A->row && A->col && A->val, DEV_ERR_INVALID_PTR, "DevBCSR: null arrays");
  LACE_REQUIRE(A->storage_type == Magma_BCSR, DEV_ERR_NOT_SUPPORTED, "DevBCSR: storage %d",
               (int)A->storage_type);
  LACE_REQUIRE(A->blocksize >= 1 && A->blocksize <= kMaxBlock, DEV_ERR_NOT_SUPPORTED,
               "DevBCSR: block size %d (1..%d)", A->blocksize, kMaxBlock);
  LACE_REQUIRE(p == LACE_F64 || p == LACE_F32, DEV_ERR_NOT_SUPPORTED, "DevBCSR: precision %d", (int)p);
  nbr = A->num_rows;
  bs = A->blocksize;
  nblocks = A->row[nbr];
  prec = p;
  const size_t lb = (size_t)bs * bs, nv = (size_t)std::max(nblocks, 1) * lb;
  LACE_TRY(upload(row, A->row, sizeof(int) * (nbr + 1)));
  LACE_TRY(upload(col, A->col, sizeof(int) * std::max(nblocks, 1)));
  if (p == LACE_F64) {
    LACE_TRY(upload(val, A->val, sizeof(double) * nv));
  } else {
    std::vector<float> f(nv, 0.0f);
    for (size_t e = 0; e < (size_t)nblocks * lb; e++) f[e] = (float)A->val[e];
    LACE_TRY(upload(val, f.data(), sizeof(float) * nv));
  }
  return DEV_SUCCESS;
}

double DevBCSR::bytes(int sx, int sy, bool residual) const {
  const double sv = prec == LACE_F64 ? 8.0 : 4.0;
  const double n = (double)nbr * bs;
  return (double)nblocks * bs * bs * sv + 4.0 * nblocks + 4.0 * (nbr + 1) + (double)sx * n +
         (double)sy * n + (residual ? (double)sy * n : 0.0);
}

template <typename TX, typename TY>
int bcsr_spmv(const DevBCSR& A, const TX* x, TY* y) {
  TimedRegion tr("bcsr_spmv");
  if (A.prec == LACE_F64) return launch_bspmv<double, TX, TY, 0>(A, x, y, nullptr);
  return launch_bspmv<float, TX, TY, 0>(A, x, y, nullptr);
}

template <typename TX, typename TY>
int bcsr_residual(const DevBCSR& A, const TX* x, const TY* b, TY* r) {
  TimedRegion tr("bcsr_spmv_residual");
  if (A.prec == LACE_F64) return launch_bspmv<double, TX, TY, 1>(A, x, r, b);
  return launch_bspmv<float, TX, TY, 1>(A, x, r, b);
}

template int bcsr_spmv<double, double>(const DevBCSR&, const double*, double*);
template int bcsr_spmv<float, double>(const DevBCSR&, const float*, double*);
template int bcsr_spmv<float, float>(const DevBCSR&, const float*, float*);
template int bcsr_residual<double, double>(const DevBCSR&, const double*, const double*, double*);
template int bcsr_residual<float, float>(const DevBCSR&, const float*, const float*, float*);

// ---------------------------------------------------------------- ParILU
// Factors from A: L = strict-lower blocks of A (the identity diagonal
// blocks are implicit and stored as zeros), U = upper blocks of A.  One
// lane per stored entry; lfrom/ufrom name the source block in src_l/src_u.
template <typename TC, typename TS>
__global__ __launch_bounds__(kBBS) void bilu_reset_kernel(long countL, long countU, int bs,
                                                          const int* __restrict__ lrowv,
                                                          const int* __restrict__ uri,
                                                          const int* __restrict__ lfrom,
                                                          const int* __restrict__ ufrom,
                                                          const TC* __restrict__ src_l,
                                                          const TC* __restrict__ src_u,
                                                          const float* __restrict__ rs,
                                                          TS* __restrict__ L0, TS* __restrict__ L1,
                                                          TS* __restrict__ U0, TS* __restrict__ U1) {
  const long e = (long)blockIdx.x * kBBS + threadIdx.x;
  const int lb = bs * bs;
  if (e < countL) {
    const int q = (int)(e / lb), w = (int)(e - (long)q * lb), r = w / bs;
    const int f = lfrom[q];
    const float s = is16<TS>() ? rs[lrowv[q] * bs + r] : 1.0f;
    const TS v = f >= 0 ? quant<TC, TS>(src_l[(size_t)f * lb + w], 1.0f / s) : TS(0);
    L0[e] = v;
    L1[e] = v;
  }
  if (e < countU) {
    const int q = (int)(e / lb), w = (int)(e - (long)q * lb), r = w / bs;
    const float s = is16<TS>() ? rs[uri[q] * bs + r] : 1.0f;
    const TS v = quant<TC, TS>(src_u[(size_t)ufrom[q] * lb + w], 1.0f / s);
    U0[e] = v;
    U1[e] = v;
  }
}

// Symmetric-pattern sweep: group (q, r) updates row r of L_IJ and of U_JI
// at position q = (I, J), reading (Lin, Uin) and writing (Lout, Uout).
template <int BS, typename TC, typename TS, bool STEP>
__global__ __launch_bounds__(kBBS) void bilu_sym_sweep_kernel(
    int np, const int* __restrict__ prow, const int* __restrict__ pcol, const int* __restrict__ lptr,
    const TC* __restrict__ AL, const TC* __restrict__ AU, const TS* __restrict__ Lin,
    const TS* __restrict__ Uin, TS* __restrict__ Lout, TS* __restrict__ Uout,
    const float* __restrict__ rs, double* __restrict__ partial) {
  constexpr int LB = BS * BS;
  const Group<BS> G(np);
  double d2 = 0.0;
  if (G.valid) {
    const int p = G.g, r = G.r;
    const int I = prow[p], J = pcol[p];
    const bool strict = I != J;
    const float sI = is16<TS>() ? rs[I * BS + r] : 1.0f;
    const float sJ = is16<TS>() ? rs[J * BS + r] : 1.0f;
    TC sL[BS], sU[BS];
#pragma unroll
    for (int c = 0; c < BS; c++) {
      sU[c] = AU[(size_t)p * LB + r * BS + c];
      sL[c] = strict ? AL[(size_t)p * LB + r * BS + c] : TC(0);
    }
    // products over K < J: row I positions before p, row J positions
    // before its diagonal, merged by block column
    int qi = lptr[I], qj = lptr[J];
    const int qj_end = lptr[J + 1] - 1;
    while (qi < p && qj < qj_end) {
      const int ki = pcol[qi], kj = pcol[qj];
      if (ki == kj) {
        TC ub[LB], lr[BS];
        if (strict) {
          // L_IJ row r -= L_IK(r,:) U_KJ;  U_KJ sits at position (J, K)
          load_block<BS, TC, TS>(Uin, qj, rs, ki, ub);
          load_row<BS, TC, TS>(Lin, qi, r, sI, lr);
#pragma unroll
          for (int c = 0; c < BS; c++) {
            TC acc = TC(0);
#pragma unroll
            for (int m = 0; m < BS; m++) acc = acc + lr[m] * ub[m * BS + c];
            sL[c] = sL[c] - acc;
          }
        }
        // U_JI row r -= L_JK(r,:) U_KI;  U_KI sits at position (I, K)
        load_block<BS, TC, TS>(Uin, qi, rs, ki, ub);
        load_row<BS, TC, TS>(Lin, qj, r, sJ, lr);
#pragma unroll
        for (int c = 0; c < BS; c++) {
          TC acc = TC(0);
#pragma unroll
          for (int m = 0; m < BS; m++) acc = acc + lr[m] * ub[m * BS + c];
          sU[c] = sU[c] - acc;
        }
      }
      qi += ki <= kj;
      qj += ki >= kj;
    }
    // the closing product K = J, subtracted and added back
    if (strict) {
      TC ujj[LB], lold[BS], inv[LB];
      load_block<BS, TC, TS>(Uin, qj_end, rs, J, ujj);
      load_row<BS, TC, TS>(Lin, p, r, sI, lold);
#pragma unroll
      for (int c = 0; c < BS; c++) {
        TC acc = TC(0);
#pragma unroll
        for (int m = 0; m < BS; m++) acc = acc + lold[m] * ujj[m * BS + c];
        sL[c] = sL[c] - acc;
        sL[c] = sL[c] + acc;
      }
      binv<BS, TC>(ujj, inv);
      const float sinv = 1.0f / sI;
#pragma unroll
      for (int c = 0; c < BS; c++) {
        TC acc = TC(0);
#pragma unroll
        for (int m = 0; m < BS; m++) acc = acc + sL[m] * inv[m * BS + c];
        const TS q = quant<TC, TS>(acc, sinv);
        Lout[(size_t)p * LB + r * BS + c] = q;
        if (STEP) {
          const double d = (double)lold[c] - (double)dequant<TC, TS>(q, sI);
          d2 += d * d;
        }
      }
    }
    {
      TC uold[BS];
      load_row<BS, TC, TS>(Uin, p, r, sJ, uold);
      const float sinv = 1.0f / sJ;
#pragma unroll
      for (int c = 0; c < BS; c++) {
        // L_JJ is the identity: the closing product is U_JI's own row
        TC v = sU[c] - uold[c];
        v = v + uold[c];
        const TS q = quant<TC, TS>(v, sinv);
        Uout[(size_t)p * LB + r * BS + c] = q;
        if (STEP) {
          const double d = (double)uold[c] - (double)dequant<TC, TS>(q, sJ);
          d2 += d * d;
        }
      }
    }
  }
  if (STEP) {
    __shared__ double sm[kBBS / 64];
    const double tot = block_sum<kBBS>(d2, sm);
    if (threadIdx.x == 0) partial[xcd_block(blockIdx.x, gridDim.x)] = tot;
  }
}

// ne consecutive elements from a 16-byte aligned global range to LDS, as
// 16-byte words where whole, element by element for the tail.
template <typename T>
__device__ __forceinline__ void stage_range(const T* __restrict__ src, T* dst, int ne) {
  constexpr int W = 16 / (int)sizeof(T);
  const int nw = ne / W;
  for (int k = threadIdx.x; k < nw; k += kBBS)
    reinterpret_cast<uint4*>(dst)[k] = reinterpret_cast<const uint4*>(src)[k];
  for (int e = nw * W + (int)threadIdx.x; e < ne; e += kBBS) dst[e] = src[e];
}
template <typename T>
__device__ __forceinline__ void unstage_range(const T* src, T* __restrict__ dst, int ne) {
  constexpr int W = 16 / (int)sizeof(T);
  const int nw = ne / W;
  for (int k = threadIdx.x; k < nw; k += kBBS)
    reinterpret_cast<uint4*>(dst)[k] = reinterpret_cast<const uint4*>(src)[k];
  for (int e = nw * W + (int)threadIdx.x; e < ne; e += kBBS) dst[e] = src[e];
}

// The symmetric sweep driven by per-position tables built at set-up:
// meta[p] = (I, J, position of u_jj, first product pair) and pend[p], and
// pairs[k] = (q_IK, q_JK, K) for every K < J where row I and row J both
// hold a block, in increasing K (the merge of the kernel above, done once on
// the host).  The lane's loads then depend on p alone or on meta[p], never
// on a chain through lptr and pcol.  With STAGE the workgroup's consecutive
// positions (48 for 5x5 blocks) go through LDS: their a_ij, a_ji, l_ij and
// u_ji blocks are contiguous in all four arrays and come in as 16-byte
// words; each group's BS lanes load one row each of its u_jj block and share
// it; each lane writes its new rows over the ones it read, and the blocks go
// out as 16-byte words, so the stride-BS row accesses become LDS accesses.
// Same arithmetic in the same order as the kernel above: identical results.
template <int BS, typename TC, typename TS, bool STEP, int STAGE>
__global__ __launch_bounds__(kBBS) void bilu_sym_sweep_meta_kernel(
    int np, const int4* __restrict__ meta, const int* __restrict__ pend, const int4* __restrict__ pairs,
    const TC* __restrict__ AL, const TC* __restrict__ AU, const TS* __restrict__ Lin,
    const TS* __restrict__ Uin, TS* __restrict__ Lout, TS* __restrict__ Uout,
    const float* __restrict__ rs, double* __restrict__ partial) {
  constexpr int LB = BS * BS;
  constexpr int PPB = (kBBS / 64) * (64 / BS);  // positions per workgroup
  // STAGE 2: all four block sets and U_JJ through LDS; 1: L, U and U_JJ
  // (A's blocks read directly); 0: none
  constexpr int SE = STAGE ? PPB * LB : 1;
  constexpr int SEA = STAGE == 2 ? PPB * LB : 1;
  constexpr bool SA = STAGE == 2;
  __shared__ __align__(16) TC shAL[SEA];
  __shared__ __align__(16) TC shAU[SEA];
  __shared__ __align__(16) TS shL[SE];
  __shared__ __align__(16) TS shU[SE];
  __shared__ TC shUjj[SE];
  const int p0 = xcd_block(blockIdx.x, gridDim.x) * PPB;
  const int ne = min(PPB, np - p0) * LB;
  if constexpr (SA) {
    stage_range<TC>(AL + (size_t)p0 * LB, shAL, ne);
    stage_range<TC>(AU + (size_t)p0 * LB, shAU, ne);
  }
  if constexpr (STAGE) {
    stage_range<TS>(Lin + (size_t)p0 * LB, shL, ne);
    stage_range<TS>(Uin + (size_t)p0 * LB, shU, ne);
  }
  const Group<BS> G(np);
  const int lp = G.g - p0;  // this group's position within the workgroup
  int4 m = make_int4(0, 0, 0, 0);
  int pe = 0;
  if (G.valid) {
    m = meta[G.g];
    pe = pend[G.g];
  }
  const int I = m.x, J = m.y, qd = m.z;
  if constexpr (STAGE) {
    if (G.valid) {
      const float s = is16<TS>() ? rs[J * BS + G.r] : 1.0f;
#pragma unroll
      for (int c = 0; c < BS; c++)
        shUjj[lp * LB + G.r * BS + c] = ldv<TC>(Uin, (size_t)qd * LB + G.r * BS + c, s);
    }
    __syncthreads();
  }
  double d2 = 0.0;
  if (G.valid) {
    const int p = G.g, r = G.r;
    const bool strict = I != J;
    const float sI = is16<TS>() ? rs[I * BS + r] : 1.0f;
    const float sJ = is16<TS>() ? rs[J * BS + r] : 1.0f;
    // this lane's row: in LDS when staged, else in the global arrays
    const int row0 = STAGE ? lp * LB + r * BS : 0;
    const size_t grow = (size_t)p * LB + r * BS;
    TC aL[BS], aU[BS], lold[BS], uold[BS];
#pragma unroll
    for (int c = 0; c < BS; c++) {
      aU[c] = SA ? shAU[row0 + c] : AU[grow + c];
      aL[c] = strict ? (SA ? shAL[row0 + c] : AL[grow + c]) : TC(0);
      uold[c] = dequant<TC, TS>(STAGE ? shU[row0 + c] : Uin[grow + c], sJ);
      if (strict) lold[c] = dequant<TC, TS>(STAGE ? shL[row0 + c] : Lin[grow + c], sI);
    }
    // products over K < J (none on a 7-point block pattern)
    for (int k = m.w; k < pe; k++) {
      const int4 t = pairs[k];  // (q_IK, q_JK, K)
      TC ub[LB], lr[BS];
      if (strict) {
        // L_IJ row r -= L_IK(r,:) U_KJ;  U_KJ sits at position (J, K)
        load_block<BS, TC, TS>(Uin, t.y, rs, t.z, ub);
        load_row<BS, TC, TS>(Lin, t.x, r, sI, lr);
#pragma unroll
        for (int c = 0; c < BS; c++) {
          TC acc = TC(0);
#pragma unroll
          for (int mm = 0; mm < BS; mm++) acc = acc + lr[mm] * ub[mm * BS + c];
          aL[c] = aL[c] - acc;
        }
      }
      // U_JI row r -= L_JK(r,:) U_KI;  U_KI sits at position (I, K)
      load_block<BS, TC, TS>(Uin, t.x, rs, t.z, ub);
      load_row<BS, TC, TS>(Lin, t.y, r, sJ, lr);
#pragma unroll
      for (int c = 0; c < BS; c++) {
        TC acc = TC(0);
#pragma unroll
        for (int mm = 0; mm < BS; mm++) acc = acc + lr[mm] * ub[mm * BS + c];
        aU[c] = aU[c] - acc;
      }
    }
    // the closing product K = J, subtracted and added back
    if (strict) {
      TC ujj[LB], inv[LB];
      if constexpr (STAGE) {
#pragma unroll
        for (int e = 0; e < LB; e++) ujj[e] = shUjj[lp * LB + e];
      } else {
        load_block<BS, TC, TS>(Uin, qd, rs, J, ujj);
      }
#pragma unroll
      for (int c = 0; c < BS; c++) {
        TC acc = TC(0);
#pragma unroll
        for (int mm = 0; mm < BS; mm++) acc = acc + lold[mm] * ujj[mm * BS + c];
        aL[c] = aL[c] - acc;
        aL[c] = aL[c] + acc;
      }
      binv<BS, TC>(ujj, inv);
      const float sinv = 1.0f / sI;
#pragma unroll
      for (int c = 0; c < BS; c++) {
        TC acc = TC(0);
#pragma unroll
        for (int mm = 0; mm < BS; mm++) acc = acc + aL[mm] * inv[mm * BS + c];
        const TS q = quant<TC, TS>(acc, sinv);
        if constexpr (STAGE) shL[row0 + c] = q;  // this lane alone reads and writes its row
        else Lout[grow + c] = q;
        if (STEP) {
          const double d = (double)lold[c] - (double)dequant<TC, TS>(q, sI);
          d2 += d * d;
        }
      }
    }
    {
      const float sinv = 1.0f / sJ;
#pragma unroll
      for (int c = 0; c < BS; c++) {
        // L_JJ is the identity: the closing product is U_JI's own row
        TC v = aU[c] - uold[c];
        v = v + uold[c];
        const TS q = quant<TC, TS>(v, sinv);
        if constexpr (STAGE) shU[row0 + c] = q;
        else Uout[grow + c] = q;
        if (STEP) {
          const double d = (double)uold[c] - (double)dequant<TC, TS>(q, sJ);
          d2 += d * d;
        }
      }
    }
  }
  if constexpr (STAGE) {
    __syncthreads();
    // diagonal positions keep their (unchanged) staged L block
    unstage_range<TS>(shL, Lout + (size_t)p0 * LB, ne);
    unstage_range<TS>(shU, Uout + (size_t)p0 * LB, ne);
  }
  if (STEP) {
    __shared__ double sm[kBBS / 64];
    const double tot = block_sum<kBBS>(d2, sm);
    if (threadIdx.x == 0) partial[xcd_block(blockIdx.x, gridDim.x)] = tot;
  }
}

// Staging applies when every workgroup's range starts on a 16-byte boundary
// in all four arrays and the five staged sets fit 64 KiB of LDS.
template <int BS, typename TC, typename TS>
constexpr bool sym_lds_ok() {
  constexpr int e = (kBBS / 64) * (64 / BS) * BS * BS;
  return (e * sizeof(TC)) % 16 == 0 && (e * sizeof(TS)) % 16 == 0 &&
         e * (3 * sizeof(TC) + 2 * sizeof(TS)) <= 65536;
}
template <int BS, typename TC, typename TS>
constexpr bool sym_lds_lu_ok() {
  constexpr int e = (kBBS / 64) * (64 / BS) * BS * BS;
  return (e * sizeof(TS)) % 16 == 0 && e * (sizeof(TC) + 2 * sizeof(TS)) <= 65536;
}

// ---------------------------------------------------------------- wave-run sweep
// A wave's PPW consecutive positions occupy one contiguous run of CE = PPW
// BS^2 elements in each block array (the arrays are padded to whole
// workgroups of runs, so every run is complete).  WaveRun moves such a run
// between global memory and LDS as the widest words that tile it (16 bytes
// for FP32 and FP64 runs of 5x5 blocks, 8 for 16-bit ones): consecutive
// lanes take consecutive words, so each load or store instruction of the
// wave covers one contiguous stretch of memory.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
template <typename T, int CE>
struct WaveRun {
  static constexpr int kBytes = CE * (int)sizeof(T);
  using W = std::conditional_t<
      kBytes % 16 == 0, u32x4,
      std::conditional_t<kBytes % 8 == 0, u32x2,
                         std::conditional_t<kBytes % 4 == 0, unsigned, unsigned short>>>;
  static constexpr int kWords = kBytes / (int)sizeof(W);
  static constexpr int kR = (kWords + 63) / 64;  // words per lane
  W w[kR];
  __device__ __forceinline__ void load(const T* __restrict__ src, int lane) {
#pragma unroll
    for (int k = 0; k < kR; k++)
      if (lane + 64 * k < kWords) w[k] = reinterpret_cast<const W*>(src)[lane + 64 * k];
  }
  __device__ __forceinline__ void to_lds(T* dst, int lane) const {
#pragma unroll
    for (int k = 0; k < kR; k++)
      if (lane + 64 * k < kWords) reinterpret_cast<W*>(dst)[lane + 64 * k] = w[k];
  }
  static __device__ __forceinline__ void from_lds(const T* src, T* __restrict__ dst, int lane) {
#pragma unroll
    for (int k = 0; k < kR; k++)
      if (lane + 64 * k < kWords)
        reinterpret_cast<W*>(dst)[lane + 64 * k] = reinterpret_cast<const W*>(src)[lane + 64 * k];
  }
};

// Positions per workgroup of the wave-run form, and a position count
// rounded up to whole workgroups (the length the block arrays and tables
// are padded to).
inline int wave_ppb(int bs) { return (kBBS / 64) * (64 / bs); }
inline int wave_pad(int bs, int np) { return (np + wave_ppb(bs) - 1) / wave_ppb(bs) * wave_ppb(bs); }

// Positions staged before a wave's run (the halo): a diagonal position's
// products come from the earlier positions of its block row, three of them
// on a 7-point block pattern, and the row may begin in the previous run.
constexpr int kHalo = 3;

// Where the products of the wave-run sweep find their operands: the old L
// and U blocks of positions [pw - kHalo, pw + PPW) staged in LDS (sL, sU
// indexed from the run's first position, the halo at negative offsets, the
// halo empty in the first run) with the scales of U's rows (sS, 16-bit
// storage), else memory.
template <int BS, int PPW, typename TS>
struct StagedFactors {
  const TS* sL;
  const TS* sU;
  const float* sS;
  const TS* __restrict__ L;
  const TS* __restrict__ U;
  const float* __restrict__ rs;
  int pw;
  __device__ __forceinline__ bool staged(int q) const {
    return (unsigned)(q - pw + kHalo) < (unsigned)(PPW + kHalo);
  }
};

// Row e0 / BS of factor block q (row scale s), staged or from memory.
template <int BS, int PPW, typename TC, typename TS>
__device__ __forceinline__ void staged_row(bool in_lds, const TS* sh, const TS* __restrict__ g, int pw, int q,
                                           int e0, float s, TC (&out)[BS]) {
  if (in_lds) {
#pragma unroll
    for (int c = 0; c < BS; c++) out[c] = dequant<TC, TS>(sh[(q - pw) * BS * BS + e0 + c], s);
  } else {
#pragma unroll
    for (int c = 0; c < BS; c++) out[c] = ldv<TC>(g, (size_t)q * BS * BS + e0 + c, s);
  }
}

// s -= L(qr, row r) U(qb) (U_qb's rows are block row kb's): each entry of
// the product summed from zero over the inner index in ascending order (as
// the products of load_row and load_block are), U's rows fetched one at a
// time.
template <int BS, int PPW, typename TC, typename TS, bool LOCAL>
__device__ __forceinline__ void staged_row_times_block_sub(const StagedFactors<BS, PPW, TS>& F, int qr, int r,
                                                           float sr, int qb, int kb, TC (&s)[BS]) {
  TC lr[BS], acc[BS];
  staged_row<BS, PPW, TC, TS>(LOCAL || F.staged(qr), F.sL, F.L, F.pw, qr, r * BS, sr, lr);
  const bool ub = LOCAL || F.staged(qb);
#pragma unroll
  for (int c = 0; c < BS; c++) acc[c] = TC(0);
#pragma unroll
  for (int mm = 0; mm < BS; mm++) {
    TC ur[BS];
    const float sm = !is16<TS>() ? 1.0f : (ub ? F.sS[(qb - F.pw) * BS + mm] : F.rs[kb * BS + mm]);
    staged_row<BS, PPW, TC, TS>(ub, F.sU, F.U, F.pw, qb, mm * BS, sm, ur);
#pragma unroll
    for (int c = 0; c < BS; c++) acc[c] = acc[c] + lr[mm] * ur[c];
  }
#pragma unroll
  for (int c = 0; c < BS; c++) s[c] = s[c] - acc[c];
}

// Orders this wave's LDS accesses: the LDS executes one wave's operations
// in issue order, so a wave-private buffer needs no workgroup barrier, only
// the LDS counter drained and the compiler kept from moving accesses across.
__device__ __forceinline__ void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Column col of inv(M) by the elimination of binv: the operations on that
// column's entries are binv's, in binv's order (binv updates each column
// of the inverse from its own entries and the multipliers alone), so the
// column is bit-identical to binv's.  The multipliers are the same in
// every lane; work on columns of M that no later step reads is dead code.
template <int BS, typename T>
__device__ __forceinline__ void binv_col(const T (&M)[BS * BS], int col, T (&x)[BS]) {
  T a[BS * BS];
#pragma unroll
  for (int e = 0; e < BS * BS; e++) a[e] = M[e];
#pragma unroll
  for (int r = 0; r < BS; r++) x[r] = r == col ? T(1) : T(0);
#pragma unroll
  for (int p = 0; p < BS; p++) {
    const T iv = T(1) / a[p * BS + p];
#pragma unroll
    for (int c = 0; c < BS; c++) a[p * BS + c] = a[p * BS + c] * iv;
    x[p] = x[p] * iv;
#pragma unroll
    for (int r = 0; r < BS; r++) {
      if (r == p) continue;
      const T f = a[r * BS + p];
#pragma unroll
      for (int c = 0; c < BS; c++) a[r * BS + c] = a[r * BS + c] - f * a[p * BS + c];
      x[r] = x[r] - f * x[p];
    }
  }
}

// The table-driven symmetric sweep with every block array moved as
// whole-word wave runs.  Each wave owns PPW consecutive positions and a
// private slice of four LDS buffers: A (a_ij, then a_ji), the old L and U
// runs (which the new rows overwrite, so diagonal and padding positions
// write their blocks back unchanged), and U_JJ (each lane stages its row),
// which then carries inv(U_JJ): lane r of a group computes column r of the
// inverse and reads the whole inverse back.  Every lane runs the same
// straight-line code (the 64 - BS PPW lanes past the last group shadow it);
// only the writes depend on whether a lane's position is real, so no load
// waits on a branch.  No workgroup barrier on the way; the arithmetic is
// that of the kernels above, operation by operation.
template <int BS, typename TC, typename TS, bool STEP, bool LOCAL>
__global__ __launch_bounds__(kBBS) void bilu_sym_sweep_wave_kernel(
    int np, const int4* __restrict__ meta, const int* __restrict__ pend, const int4* __restrict__ pairs,
    const TC* __restrict__ AL, const TC* __restrict__ AU, const TS* __restrict__ Lin,
    const TS* __restrict__ Uin, TS* __restrict__ Lout, TS* __restrict__ Uout,
    const float* __restrict__ rs, double* __restrict__ partial, const uint8_t* __restrict__ whalo) {
  constexpr int LB = BS * BS, PPW = 64 / BS, CE = PPW * LB, NWV = kBBS / 64;
  constexpr int HE = kHalo * LB;  // halo elements
  constexpr int HS = (HE * (int)sizeof(TS) + 15) / 16 * 16 / (int)sizeof(TS);  // halo slots (16-byte multiple)
  constexpr int SS = is16<TS>() ? (kHalo + PPW) * BS : 1;
  __shared__ __align__(16) TC shA[NWV * CE];
  __shared__ __align__(16) TC shJ[NWV * CE];
  __shared__ __align__(16) TS shL[NWV * (HS + CE)];
  __shared__ __align__(16) TS shU[NWV * (HS + CE)];
  __shared__ float shS[NWV * SS];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int pw = (xcd_block(blockIdx.x, gridDim.x) * NWV + wv) * PPW;  // the wave's first position
  const size_t off = (size_t)pw * LB;
  TC* sA = shA + wv * CE;
  TS* sL = shL + wv * (HS + CE) + HS;  // the run; the halo at [-HE, 0)
  TS* sU = shU + wv * (HS + CE) + HS;
  float* sS = shS + wv * SS + kHalo * BS;  // U row scales, indexed (q - pw) * BS + row
  const int slot = lane / BS, r = lane - slot * BS;
  const bool lane_on = slot < PPW;
  const int lp = min(slot, PPW - 1);
  const int4 m = meta[pw + lp];
  const int pe = pend[pw + lp];
  WaveRun<TC, CE> ra, rb;
  WaveRun<TS, CE> rl, ru;
  ra.load(AL + off, lane);
  rl.load(Lin + off, lane);
  ru.load(Uin + off, lane);
  const int I = m.x, J = m.y, qd = m.z;
  const bool active = lane_on && pw + lp < np;
  const bool strict = active && I != J;
  const float sI = is16<TS>() ? rs[I * BS + r] : 1.0f;
  const float sJs = is16<TS>() ? rs[J * BS + r] : 1.0f;
  TS ujr[BS];  // this lane's row of U_JJ
#pragma unroll
  for (int c = 0; c < BS; c++) ujr[c] = Uin[(size_t)qd * LB + r * BS + c];
  // the first three pair-table entries (a diagonal position of a 7-point
  // block pattern has three), fetched alongside U_JJ so the product loop
  // below does not wait on a dependent load per step (indices clamped into
  // the table; entries past pend are never used)
  const int4 pf0 = pairs[max(0, min(m.w, pe - 1))];
  const int4 pf1 = pairs[max(0, min(m.w + 1, pe - 1))];
  const int4 pf2 = pairs[max(0, min(m.w + 2, pe - 1))];
  rb.load(AU + off, lane);
  // the halo's old L and U blocks, element by element (the halo does not
  // start on a word boundary); the first run has none
  // (clamped indices, every lane loading; only the runs whose products
  // reach before the run -- a set-up flag per run, the same in every lane --
  // load one)
  constexpr int HR = (HE + 63) / 64;
  const bool halo = __builtin_amdgcn_readfirstlane(whalo[pw / PPW]) != 0;
  const size_t hoff = pw > 0 ? off - HE : off;
  TS hl[HR], hu[HR];
  float hs = 1.0f;  // the scales of U's rows at the halo positions (lanes h * BS + row)
  if (halo) {
#pragma unroll
    for (int k = 0; k < HR; k++) {
      const int e = min(lane + 64 * k, HE - 1);
      hl[k] = Lin[hoff + e];
      hu[k] = Uin[hoff + e];
    }
    if constexpr (is16<TS>()) {
      const int h = min(lane, kHalo * BS - 1);
      hs = rs[meta[max(pw - kHalo, 0) + h / BS].y * BS + h % BS];
    }
  }
  const int row0 = lp * LB + r * BS;
  TC* jb = shJ + wv * CE + lp * LB;  // this group's U_JJ, later inv(U_JJ)
  ra.to_lds(sA, lane);
  rl.to_lds(sL, lane);
  ru.to_lds(sU, lane);
  if (halo) {
#pragma unroll
    for (int k = 0; k < HR; k++) {
      const int e = lane + 64 * k;
      if (pw > 0 && e < HE) {
        sL[e - HE] = hl[k];
        sU[e - HE] = hu[k];
      }
    }
  }
  if (is16<TS>()) {
    if (lane_on) sS[lp * BS + r] = sJs;
    if (halo && lane < kHalo * BS) sS[lane - kHalo * BS] = hs;
  }
  if (lane_on) {
#pragma unroll
    for (int c = 0; c < BS; c++) jb[r * BS + c] = dequant<TC, TS>(ujr[c], sJs);
  }
  wave_lds_sync();
  TC aL[BS], aU[BS], lold[BS], uold[BS];
#pragma unroll
  for (int c = 0; c < BS; c++) {
    aL[c] = sA[row0 + c];
    lold[c] = dequant<TC, TS>(sL[row0 + c], sI);
    uold[c] = dequant<TC, TS>(sU[row0 + c], sJs);
  }
  wave_lds_sync();
  rb.to_lds(sA, lane);
  wave_lds_sync();
#pragma unroll
  for (int c = 0; c < BS; c++) aU[c] = sA[row0 + c];
  // products over K < J (on a 7-point block pattern only the diagonal
  // positions have them, three each, all from the position's own block row,
  // which the run and its halo hold)
  const StagedFactors<BS, PPW, TS> F{sL, sU, sS, Lin, Uin, rs, pw};
  if constexpr (LOCAL) {
    // every position has at most three products and all their operands are
    // staged (checked at set-up)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (j < pe - m.w) {
        const int4 t = j == 0 ? pf0 : (j == 1 ? pf1 : pf2);  // (q_IK, q_JK, K)
        // L_IJ row r -= L_IK(r,:) U_KJ (U_KJ at position (J, K)); U_JI row
        // r -= L_JK(r,:) U_KI (U_KI at position (I, K))
        if (I != J) staged_row_times_block_sub<BS, PPW, TC, TS, true>(F, t.x, r, sI, t.y, t.z, aL);
        staged_row_times_block_sub<BS, PPW, TC, TS, true>(F, t.y, r, sJs, t.x, t.z, aU);
      }
    }
  } else {
    for (int k = m.w; k < pe; k++) {
      const int j = k - m.w;
      int4 t;  // (q_IK, q_JK, K)
      if (j == 0) t = pf0;
      else if (j == 1) t = pf1;
      else if (j == 2) t = pf2;
      else t = pairs[k];
      if (I != J) staged_row_times_block_sub<BS, PPW, TC, TS, false>(F, t.x, r, sI, t.y, t.z, aL);
      staged_row_times_block_sub<BS, PPW, TC, TS, false>(F, t.y, r, sJs, t.x, t.z, aU);
    }
  }
  // U's row first (its registers are free for what follows); its squared
  // changes are added after L's, in the order of the other kernels
  double du2[STEP ? BS : 1];
  {
    const float sinv = 1.0f / sJs;
#pragma unroll
    for (int c = 0; c < BS; c++) {
      // L_JJ is the identity: the closing product is U_JI's own row
      TC v = aU[c] - uold[c];
      v = v + uold[c];
      const TS q = quant<TC, TS>(v, sinv);
      if (active) sU[row0 + c] = q;
      if constexpr (STEP) {
        const double d = (double)uold[c] - (double)dequant<TC, TS>(q, sJs);
        du2[c] = active ? d * d : 0.0;
      }
    }
  }
  // the closing product K = J, subtracted and added back
  TC ujj[LB];
#pragma unroll
  for (int e = 0; e < LB; e++) ujj[e] = jb[e];
#pragma unroll
  for (int c = 0; c < BS; c++) {
    TC acc = TC(0);
#pragma unroll
    for (int mm = 0; mm < BS; mm++) acc = acc + lold[mm] * ujj[mm * BS + c];
    aL[c] = aL[c] - acc;
    aL[c] = aL[c] + acc;
  }
  TC icol[BS];
  binv_col<BS, TC>(ujj, r, icol);
  if (lane_on) {
#pragma unroll
    for (int mm = 0; mm < BS; mm++) jb[mm * BS + r] = icol[mm];
  }
  wave_lds_sync();
  TC inv[LB];
#pragma unroll
  for (int e = 0; e < LB; e++) inv[e] = jb[e];
  double d2 = 0.0;
  {
    const float sinv = 1.0f / sI;
#pragma unroll
    for (int c = 0; c < BS; c++) {
      TC acc = TC(0);
#pragma unroll
      for (int mm = 0; mm < BS; mm++) acc = acc + aL[mm] * inv[mm * BS + c];
      const TS q = quant<TC, TS>(acc, sinv);
      if (strict) sL[row0 + c] = q;
      if (STEP && strict) {
        const double d = (double)lold[c] - (double)dequant<TC, TS>(q, sI);
        d2 += d * d;
      }
    }
  }
  if constexpr (STEP) {
#pragma unroll
    for (int c = 0; c < BS; c++) d2 += du2[c];
  }
  wave_lds_sync();
  WaveRun<TS, CE>::from_lds(sL, Lout + off, lane);
  WaveRun<TS, CE>::from_lds(sU, Uout + off, lane);
  if (STEP) {
    __shared__ double sm[kBBS / 64];
    const double tot = block_sum<kBBS>(d2, sm);
    if (threadIdx.x == 0) partial[xcd_block(blockIdx.x, gridDim.x)] = tot;
  }
}

// The wave-run form applies when its four buffers fit 64 KiB of LDS.
template <int BS, typename TC, typename TS>
constexpr bool sym_wave_ok() {
  constexpr int e = (kBBS / 64) * (64 / BS) * BS * BS;
  return e * (2 * sizeof(TC) + 2 * sizeof(TS)) + (kBBS / 64) * (2 * 16 * BS * BS * kHalo + 4 * (kHalo + 64) * BS) +
             64 <= 65536;
}

// General sweep: group (k, r) takes block k = (I, J) of A in BCSR order,
// merges L row I (BCSRL) with U column J (BCSCU) and writes row r of L_IJ
// (I > J) or U_IJ (I <= J), as src/parilu_v0_3_bcsr.cpp:130-247 does.  The
// merge always ends on the match K = min(I, J); that product is
// subtracted and added back.
template <int BS, typename TC, typename TS, bool STEP>
__global__ __launch_bounds__(kBBS) void bilu_gen_sweep_kernel(
    int nblk, const int* __restrict__ arowidx, const int* __restrict__ acol,
    const TC* __restrict__ AV, const int* __restrict__ lptr, const int* __restrict__ lcol,
    const int* __restrict__ ucp, const int* __restrict__ uri, const TS* __restrict__ Lin,
    const TS* __restrict__ Uin, TS* __restrict__ Lout, TS* __restrict__ Uout,
    const float* __restrict__ rs, double* __restrict__ partial) {
  constexpr int LB = BS * BS;
  const Group<BS> G(nblk);
  double d2 = 0.0;
  if (G.valid) {
    const int k = G.g, r = G.r;
    const int I = arowidx[k], J = acol[k];
    const float sI = is16<TS>() ? rs[I * BS + r] : 1.0f;
    TC s[BS], last[BS];
#pragma unroll
    for (int c = 0; c < BS; c++) {
      s[c] = AV[(size_t)k * LB + r * BS + c];
      last[c] = TC(0);
    }
    int il = lptr[I], iu = ucp[J];
    const int il_end = lptr[I + 1], iu_end = ucp[J + 1];
    while (il < il_end && iu < iu_end) {
      const int jl = lcol[il], ju = uri[iu];
      if (jl == ju) {
        if (jl == I) {
          // L_II is the identity: the product is U_IJ's row r
          load_row<BS, TC, TS>(Uin, iu, r, sI, last);
        } else {
          TC ub[LB], lr[BS];
          load_block<BS, TC, TS>(Uin, iu, rs, ju, ub);
          load_row<BS, TC, TS>(Lin, il, r, sI, lr);
#pragma unroll
          for (int c = 0; c < BS; c++) {
            TC acc = TC(0);
#pragma unroll
            for (int m = 0; m < BS; m++) acc = acc + lr[m] * ub[m * BS + c];
            last[c] = acc;
          }
        }
#pragma unroll
        for (int c = 0; c < BS; c++) s[c] = s[c] - last[c];
      }
      il += jl <= ju;
      iu += jl >= ju;
    }
#pragma unroll
    for (int c = 0; c < BS; c++) s[c] = s[c] + last[c];
    const float sinv = 1.0f / sI;
    if (I > J) {
      TC ujj[LB], inv[LB], lold[BS];
      load_block<BS, TC, TS>(Uin, iu_end - 1, rs, J, ujj);
      load_row<BS, TC, TS>(Lin, il - 1, r, sI, lold);
      binv<BS, TC>(ujj, inv);
#pragma unroll
      for (int c = 0; c < BS; c++) {
        TC acc = TC(0);
#pragma unroll
        for (int m = 0; m < BS; m++) acc = acc + s[m] * inv[m * BS + c];
        const TS q = quant<TC, TS>(acc, sinv);
        Lout[(size_t)(il - 1) * LB + r * BS + c] = q;
        if (STEP) {
          const double d = (double)lold[c] - (double)dequant<TC, TS>(q, sI);
          d2 += d * d;
        }
      }
    } else {
      TC uold[BS];
      load_row<BS, TC, TS>(Uin, iu - 1, r, sI, uold);
#pragma unroll
      for (int c = 0; c < BS; c++) {
        const TS q = quant<TC, TS>(s[c], sinv);
        Uout[(size_t)(iu - 1) * LB + r * BS + c] = q;
        if (STEP) {
          const double d = (double)uold[c] - (double)dequant<TC, TS>(q, sI);
          d2 += d * d;
        }
      }
    }
  }
  if (STEP) {
    __shared__ double sm[kBBS / 64];
    const double tot = block_sum<kBBS>(d2, sm);
    if (threadIdx.x == 0) partial[xcd_block(blockIdx.x, gridDim.x)] = tot;
  }
}

__global__ __launch_bounds__(256) void bilu_sum_kernel(int nb, const double* __restrict__ partial,
                                                       double* __restrict__ out) {
  __shared__ double sm[4];
  double v = 0.0;
  for (int t = threadIdx.x; t < nb; t += 256) v += partial[t];
  const double tot = block_sum<256>(v, sm);
  if (threadIdx.x == 0) *out = tot;
}

// Inverses of U's diagonal blocks (last block of each U column), one lane
// per block row.
template <int BS, typename TC, typename TS>
__global__ __launch_bounds__(kBBS) void bilu_dinv_kernel(int nbr, const int* __restrict__ ucp,
                                                         const TS* __restrict__ Uv,
                                                         const float* __restrict__ rs,
                                                         TC* __restrict__ Dinv) {
  const int I = blockIdx.x * kBBS + threadIdx.x;
  if (I >= nbr) return;
  TC u[BS * BS], inv[BS * BS];
  load_block<BS, TC, TS>(Uv, ucp[I + 1] - 1, rs, I, u);
  binv<BS, TC>(u, inv);
#pragma unroll
  for (int e = 0; e < BS * BS; e++) Dinv[(size_t)I * BS * BS + e] = inv[e];
}

// ---------------------------------------------------------------- solves
// Lower (identity diagonal blocks): y_I = r_I - sum_K L_IK yin_K over the
// strict entries of L row I.  LEVEL: the group list is rows[0..count).
template <int BS, typename TC, typename TS, bool LEVEL>
__global__ __launch_bounds__(kBBS) void btri_lower_kernel(int count, const int* __restrict__ rows,
                                                          const int* __restrict__ lptr,
                                                          const int* __restrict__ lcol,
                                                          const TS* __restrict__ Lv,
                                                          const float* __restrict__ rs,
                                                          const TC* __restrict__ rhs,
                                                          const TC* yin, TC* yout) {
  const Group<BS> G(count);
  if (!G.valid) return;
  const int I = LEVEL ? rows[G.g] : G.g, r = G.r;
  const float s = is16<TS>() ? rs[I * BS + r] : 1.0f;
  TC acc = TC(0);
  const int q1 = lptr[I + 1] - 1;
  for (int q = lptr[I]; q < q1; q++) {
    const int K = lcol[q];
#pragma unroll
    for (int c = 0; c < BS; c++)
      acc = acc + ldv<TC>(Lv, (size_t)q * BS * BS + r * BS + c, s) * yin[(size_t)K * BS + c];
  }
  yout[(size_t)I * BS + r] = rhs[(size_t)I * BS + r] - acc;
}

// Upper: y_I = inv(U_II) (r_I - sum_{K>I} U_IK yin_K); U row I lists
// (index into U's storage, block column K).  FIRST skips the sum (yin = 0).
template <int BS, typename TC, typename TS, bool LEVEL, bool FIRST>
__global__ __launch_bounds__(kBBS) void btri_upper_kernel(int count, const int* __restrict__ rows,
                                                          const int* __restrict__ urp,
                                                          const int* __restrict__ uidx,
                                                          const int* __restrict__ ucolk,
                                                          const TS* __restrict__ Uv,
                                                          const float* __restrict__ rs,
                                                          const TC* __restrict__ Dinv,
                                                          const TC* __restrict__ rhs, const TC* yin,
                                                          TC* yout) {
  const Group<BS> G(count);
  // every lane of a group takes part in the exchange below
  const bool valid = G.valid;
  const int I = valid ? (LEVEL ? rows[G.g] : G.g) : 0, r = G.r;
  TC t = TC(0);
  if (valid) {
    TC acc = TC(0);
    if (!FIRST) {
      const float s = is16<TS>() ? rs[I * BS + r] : 1.0f;
      for (int k = urp[I]; k < urp[I + 1]; k++) {
        const int q = uidx[k], K = ucolk[k];
#pragma unroll
        for (int c = 0; c < BS; c++)
          acc = acc + ldv<TC>(Uv, (size_t)q * BS * BS + r * BS + c, s) * yin[(size_t)K * BS + c];
      }
    }
    t = rhs[(size_t)I * BS + r] - acc;
  }
  const int base = (threadIdx.x & 63) - r;
  TC y = TC(0);
#pragma unroll
  for (int c = 0; c < BS; c++) {
    const TC tc = __shfl(t, base + c, 64);
    if (valid) y = y + Dinv[(size_t)I * BS * BS + r * BS + c] * tc;
  }
  if (valid) yout[(size_t)I * BS + r] = y;
}

// ---------------------------------------------------------------- host
static void host_scale_rows(const data_d_matrix* A, std::vector<float>& rs) {
  const int bs = A->blocksize, nbr = A->num_rows, lb = bs * bs;
  rs.assign((size_t)nbr * bs, 1.0f);
  for (int I = 0; I < nbr; I++)
    for (int a = 0; a < bs; a++) {
      double m = 0.0;
      for (int k = A->row[I]; k < A->row[I + 1]; k++)
        for (int c = 0; c < bs; c++) m = std::max(m, std::fabs(A->val[(size_t)k * lb + a * bs + c]));
      if (m > 0.0 && std::isfinite(m)) {
        int e = 0;
        (void)std::frexp(m, &e);  // m = f 2^e, f in [0.5, 1)
        e = std::max(-126, std::min(127, e - 1));
        rs[(size_t)I * bs + a] = std::ldexp(1.0f, e);
      }
    }
}

template <typename TC>
static int upload_as(DevBuf& dst, const std::vector<double>& v) {
  std::vector<TC> t(v.begin(), v.end());
  if (t.empty()) t.push_back(TC(0));
  return upload(dst, t.data(), sizeof(TC) * t.size());
}

static int upload_ints(DevBuf& dst, const std::vector<int>& v) {
  if (v.empty()) {
    const int z = 0;
    return upload(dst, &z, sizeof(int));
  }
  return upload(dst, v.data(), sizeof(int) * v.size());
}

int BILU::create(const data_d_matrix* A, lace_prec_t storage) {
  LACE_TRY(ensure_init());
  LACE_REQUIRE(A && A->row && A->col && A->val, DEV_ERR_INVALID_PTR, "block ParILU: null arrays");
  LACE_REQUIRE(A->storage_type == Magma_BCSR, DEV_ERR_NOT_SUPPORTED,
               "block ParILU: A must be BCSR (storage %d)", (int)A->storage_type);
  LACE_REQUIRE(A->blocksize >= 1 && A->blocksize <= kMaxBlock, DEV_ERR_NOT_SUPPORTED,
               "block ParILU: block size %d (1..%d)", A->blocksize, kMaxBlock);
  LACE_REQUIRE(storage == LACE_F64 || storage == LACE_F32 || storage == LACE_BF16 || storage == LACE_F16,
               DEV_ERR_NOT_SUPPORTED, "block ParILU: storage precision %d", (int)storage);
  LACE_REQUIRE(A->num_rows == A->num_cols, DEV_ERR_ILLEGAL_VALUE, "block ParILU: A not square");
  nbr = A->num_rows;
  bs = A->blocksize;
  store = storage;
  nblocksA = A->row[nbr];
  const int lb = bs * bs;
  // pattern checks: sorted block columns, diagonal block present
  std::vector<int> dpos(nbr, -1);
  for (int I = 0; I < nbr; I++)
    for (int k = A->row[I]; k < A->row[I + 1]; k++) {
      const int J = A->col[k];
      LACE_REQUIRE(J >= 0 && J < nbr, DEV_ERR_ILLEGAL_VALUE, "block ParILU: column %d out of range", J);
      LACE_REQUIRE(k == A->row[I] || A->col[k - 1] < J, DEV_ERR_ILLEGAL_VALUE,
                   "block ParILU: block row %d not sorted", I);
      if (J == I) dpos[I] = k;
    }
  for (int I = 0; I < nbr; I++)
    LACE_REQUIRE(dpos[I] >= 0, DEV_ERR_ILLEGAL_VALUE, "block ParILU: diagonal block %d missing", I);
  // L (BCSRL): strict-lower blocks, then the diagonal
  std::vector<int> lf;
  h_lptr.assign(nbr + 1, 0);
  h_lcol.clear();
  h_lrow.clear();
  for (int I = 0; I < nbr; I++) {
    h_lptr[I] = (int)h_lcol.size();
    for (int k = A->row[I]; k < A->row[I + 1] && A->col[k] <= I; k++) {
      h_lcol.push_back(A->col[k]);
      h_lrow.push_back(I);
      lf.push_back(A->col[k] == I ? -1 : k);
    }
  }
  h_lptr[nbr] = (int)h_lcol.size();
  nnzL = h_lptr[nbr];
  nstrict = nnzL - nbr;
  // U (BCSCU): column J holds rows K <= J ascending, diagonal last
  std::vector<int> uf;
  h_ucp.assign(nbr + 1, 0);
  for (int I = 0; I < nbr; I++)
    for (int k = A->row[I]; k < A->row[I + 1]; k++)
      if (A->col[k] >= I) h_ucp[A->col[k] + 1]++;
  for (int J = 0; J < nbr; J++) h_ucp[J + 1] += h_ucp[J];
  nnzU = h_ucp[nbr];
  h_uri.assign(nnzU, 0);
  uf.assign(nnzU, 0);
  {
    std::vector<int> fill(h_ucp.begin(), h_ucp.end() - 1);
    for (int I = 0; I < nbr; I++)
      for (int k = A->row[I]; k < A->row[I + 1]; k++)
        if (A->col[k] >= I) {
          const int q = fill[A->col[k]]++;
          h_uri[q] = I;
          uf[q] = k;
        }
  }
  sym = (h_ucp == h_lptr) && (h_uri == h_lcol);
  // U by rows for the solves: (U index, column K > I), columns ascending
  h_urp.assign(nbr + 1, 0);
  for (int J = 0; J < nbr; J++)
    for (int q = h_ucp[J]; q < h_ucp[J + 1]; q++)
      if (h_uri[q] != J) h_urp[h_uri[q] + 1]++;
  for (int I = 0; I < nbr; I++) h_urp[I + 1] += h_urp[I];
  h_uidx.assign(h_urp[nbr], 0);
  h_ucolk.assign(h_urp[nbr], 0);
  {
    std::vector<int> fill(h_urp.begin(), h_urp.end() - 1);
    for (int J = 0; J < nbr; J++)
      for (int q = h_ucp[J]; q < h_ucp[J + 1]; q++)
        if (h_uri[q] != J) {
          const int t = fill[h_uri[q]]++;
          h_uidx[t] = q;
          h_ucolk[t] = J;
        }
  }
  double fro = 0.0;
  for (size_t e = 0; e < (size_t)nblocksA * lb; e++) fro += A->val[e] * A->val[e];
  anorm = std::sqrt(fro);
  // values in the iteration precision: A at L positions for the symmetric
  // kernel (then L / U initial values come from AL / AU), A itself for the
  // general kernel
  std::vector<int> lfrom_h(nnzL), ufrom_h(nnzU);
  // the symmetric kernel's block arrays and tables are padded to whole
  // workgroups of wave runs (padding: zero blocks, positions (0, 0) with
  // no products, never written back as factors)
  const int npad = sym ? wave_pad(bs, nnzL) : nnzL;
  if (sym) {
    std::vector<double> al((size_t)npad * lb, 0.0), au((size_t)npad * lb, 0.0);
    for (int q = 0; q < nnzL; q++) {
      if (lf[q] >= 0)
        memcpy(&al[(size_t)q * lb], A->val + (size_t)lf[q] * lb, sizeof(double) * lb);
      memcpy(&au[(size_t)q * lb], A->val + (size_t)uf[q] * lb, sizeof(double) * lb);
      lfrom_h[q] = lf[q] >= 0 ? q : -1;
      ufrom_h[q] = q;
    }
    if (c64()) {
      LACE_TRY(upload_as<double>(AL, al));
      LACE_TRY(upload_as<double>(AU, au));
    } else {
      LACE_TRY(upload_as<float>(AL, al));
      LACE_TRY(upload_as<float>(AU, au));
    }
    // per-position tables of the symmetric kernel: (I, J, q_JJ, first
    // pair), end of pairs, and the pairs (q_IK, q_JK, K) for K < J in
    // increasing K, where rows I and J both hold block column K
    std::vector<int> mt((size_t)npad * 4, 0), pe(npad, 0), pr;
    for (int q = 0; q < nnzL; q++) {
      const int I = h_lrow[q], J = h_lcol[q];
      mt[(size_t)q * 4] = I;
      mt[(size_t)q * 4 + 1] = J;
      mt[(size_t)q * 4 + 2] = h_lptr[J + 1] - 1;
      mt[(size_t)q * 4 + 3] = (int)(pr.size() / 4);
      int qi = h_lptr[I], qj = h_lptr[J];
      const int qj_end = h_lptr[J + 1] - 1;
      while (qi < q && qj < qj_end) {
        const int ki = h_lcol[qi], kj = h_lcol[qj];
        if (ki == kj) pr.insert(pr.end(), {qi, qj, ki, 0});
        qi += ki <= kj;
        qj += ki >= kj;
      }
      LACE_REQUIRE(pr.size() / 4 <= (size_t)INT_MAX, DEV_ERR_ILLEGAL_VALUE,
                   "block ParILU: too many block products");
      pe[q] = (int)(pr.size() / 4);
    }
    // whether the wave-run sweep finds every product's operands in the
    // staged run of its position or the halo before it, with at most three
    // products per position
    // and, per run, whether any product reaches before it (the runs that
    // load the halo)
    wave_local = true;
    std::vector<uint8_t> wh(npad / (64 / bs), 0);
    {
      const int ppw = 64 / bs;
      for (int q = 0; q < nnzL; q++) {
        const int w0 = q / ppw * ppw;
        if (pe[q] - mt[(size_t)q * 4 + 3] > 3) wave_local = false;
        for (int k = mt[(size_t)q * 4 + 3]; k < pe[q]; k++)
          for (int e = 0; e < 2; e++) {
            const int x = pr[(size_t)k * 4 + e];
            if (x < w0 - kHalo || x >= w0 + ppw) wave_local = false;
            if (x < w0) wh[w0 / ppw] = 1;
          }
      }
    }
    LACE_TRY(upload(swhalo, wh.data(), wh.size()));
    if (pr.empty()) pr.assign(4, 0);
    LACE_TRY(upload_ints(smeta, mt));
    LACE_TRY(upload_ints(spend, pe));
    LACE_TRY(upload_ints(spairs, pr));
  }
  // host copy of A for the general kernel (now, or after a variant switch)
  h_arow.assign(A->row, A->row + nbr + 1);
  h_acol.assign(A->col, A->col + nblocksA);
  h_aval.assign(A->val, A->val + (size_t)nblocksA * lb);
  if (!sym) {
    LACE_TRY(upload_general());
    lfrom_h = lf;
    ufrom_h = uf;
  }
  LACE_TRY(upload_ints(lfrom, lfrom_h));
  LACE_TRY(upload_ints(ufrom, ufrom_h));
  std::vector<float> h_rs;
  host_scale_rows(A, h_rs);
  LACE_TRY(upload(rs, h_rs.data(), sizeof(float) * h_rs.size()));
  LACE_TRY(upload_ints(lptr, h_lptr));
  LACE_TRY(upload_ints(lcol, h_lcol));
  LACE_TRY(upload_ints(lrow, h_lrow));
  LACE_TRY(upload_ints(ucp, h_ucp));
  LACE_TRY(upload_ints(uri, h_uri));
  LACE_TRY(upload_ints(urp, h_urp));
  LACE_TRY(upload_ints(uidx, h_uidx));
  LACE_TRY(upload_ints(ucolk, h_ucolk));
  for (int s = 0; s < 2; s++) {
    LACE_TRY(Lv[s].alloc((size_t)std::max(nnzL, npad) * lb * selem()));
    LACE_TRY(Uv[s].alloc((size_t)std::max(nnzU, npad) * lb * selem()));
    LACE_HIP(hipMemsetAsync(Lv[s].p, 0, Lv[s].bytes, stream()));
    LACE_HIP(hipMemsetAsync(Uv[s].p, 0, Uv[s].bytes, stream()));
  }
  LACE_TRY(Dinv.alloc((size_t)nbr * lb * celem()));
  LACE_TRY(work.alloc((size_t)2 * n() * celem()));
  const int nbmax = group_blocks<kMaxBlock>(std::max(nnzL, nblocksA));
  LACE_TRY(partial.alloc(sizeof(double) * nbmax));
  LACE_TRY(scalar.alloc(sizeof(double) * 2));
  // level schedules of the block triangles
  {
    std::vector<int> lrs(nbr + 1, 0), lcs;
    for (int I = 0; I < nbr; I++) {
      lrs[I] = (int)lcs.size();
      for (int q = h_lptr[I]; q < h_lptr[I + 1] - 1; q++) lcs.push_back(h_lcol[q]);
    }
    lrs[nbr] = (int)lcs.size();
    if (lcs.empty()) lcs.push_back(0);
    h_lperm.assign(nbr, 0);
    h_llev.assign(nbr + 1, 0);
    const int nl = lace_level_sets(nbr, lrs.data(), lcs.data(), 0, h_lperm.data(), h_llev.data());
    LACE_REQUIRE(nl >= 0, DEV_ERR_ILLEGAL_VALUE, "block ParILU: level sets of L failed");
    h_llev.resize(nl + 1);
    std::vector<int> uc(h_ucolk);
    if (uc.empty()) uc.push_back(0);
    h_uperm.assign(nbr, 0);
    h_ulev.assign(nbr + 1, 0);
    const int nu = lace_level_sets(nbr, h_urp.data(), uc.data(), 1, h_uperm.data(), h_ulev.data());
    LACE_REQUIRE(nu >= 0, DEV_ERR_ILLEGAL_VALUE, "block ParILU: level sets of U failed");
    h_ulev.resize(nu + 1);
    LACE_TRY(upload_ints(lperm, h_lperm));
    LACE_TRY(upload_ints(uperm, h_uperm));
  }
  return reset();
}

// A's coordinates and values for the general kernel (uploaded on demand).
int BILU::upload_general() {
  if (have_general) return DEV_SUCCESS;
  const int lb = bs * bs;
  std::vector<int> ri(nblocksA);
  for (int I = 0; I < nbr; I++)
    for (int k = h_arow[I]; k < h_arow[I + 1]; k++) ri[k] = I;
  LACE_TRY(upload_ints(arowidx, ri));
  LACE_TRY(upload_ints(acol, h_acol));
  std::vector<double> v(h_aval.begin(), h_aval.begin() + (size_t)nblocksA * lb);
  if (c64()) LACE_TRY(upload_as<double>(AV, v));
  else LACE_TRY(upload_as<float>(AV, v));
  have_general = true;
  return DEV_SUCCESS;
}

int BILU::set_variant(int v) {
  LACE_REQUIRE(v >= 0 && v <= 6, DEV_ERR_ILLEGAL_VALUE, "block ParILU: variant %d", v);
  variant = v;
  if (v == 1) LACE_TRY(upload_general());
  return DEV_SUCCESS;
}

template <typename TC, typename TS>
static int reset_t(BILU* B) {
  const int lb = B->bs * B->bs;
  const long cl = (long)B->nnzL * lb, cu = (long)B->nnzU * lb;
  const long count = std::max(cl, cu);
  const TC* src_l = B->sym ? B->AL.as<TC>() : B->AV.as<TC>();
  const TC* src_u = B->sym ? B->AU.as<TC>() : B->AV.as<TC>();
  bilu_reset_kernel<TC, TS><<<(int)((count + kBBS - 1) / kBBS), kBBS, 0, stream()>>>(
      cl, cu, B->bs, B->lrow.as<int>(), B->uri.as<int>(), B->lfrom.as<int>(), B->ufrom.as<int>(),
      src_l, src_u, B->rs.as<float>(), B->Lv[0].as<TS>(), B->Lv[1].as<TS>(), B->Uv[0].as<TS>(),
      B->Uv[1].as<TS>());
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

int BILU::reset() {
  cur = 0;
  dirty = true;
  switch (store) {
    case LACE_F64: LACE_TRY((reset_t<double, double>(this))); break;
    case LACE_F32: LACE_TRY((reset_t<float, float>(this))); break;
    case LACE_BF16: LACE_TRY((reset_t<float, uint16_t>(this))); break;
    default: LACE_TRY((reset_t<float, _Float16>(this))); break;
  }
  return DEV_SUCCESS;
}

template <int BS, typename TC, typename TS>
static int sweep_bs(BILU* B, bool with_step) {
  const int c = B->cur, o = 1 - B->cur;
  int nb = 0;
  if (B->use_sym()) {
    nb = group_blocks<BS>(B->nnzL);
    auto go = [&](auto kern) {
      kern<<<nb, kBBS, 0, stream()>>>(B->nnzL, B->lrow.as<int>(), B->lcol.as<int>(), B->lptr.as<int>(),
                                       B->AL.as<TC>(), B->AU.as<TC>(), B->Lv[c].as<TS>(),
                                       B->Uv[c].as<TS>(), B->Lv[o].as<TS>(), B->Uv[o].as<TS>(),
                                       B->rs.as<float>(), B->partial.as<double>());
    };
    // variant 0 takes the wave-run form where its LDS fits, else the
    // tables with L, U and U_JJ staged where that staging applies, else the
    // device-side merge; 2 the merge, 3 the tables, 4 the tables with every
    // block set staged, 5 with L, U and U_JJ staged (wherever the staging
    // applies), 6 the wave-run form.  Measured on 5x5 blocks of a 64^3 grid
    // (FP32 / BF16 / FP64 storage): merge 235 / 237 / 434 us, tables 251 /
    // 257 / 464, all staged 219 / 206 / 493 (48 KB of LDS per workgroup in
    // FP64), L, U and U_JJ staged 204 / 191 / 399, wave runs 138 / 155 / 297.
    constexpr bool can_stage = sym_lds_ok<BS, TC, TS>();
    constexpr bool can_stage_lu = sym_lds_lu_ok<BS, TC, TS>();
    constexpr bool can_wave = sym_wave_ok<BS, TC, TS>();
    const bool wave = can_wave && (B->variant == 0 || B->variant == 6);
    const bool tables = B->variant >= 3 || (B->variant == 0 && can_stage_lu);
    if (wave) {
      auto run = [&](auto kern) {
        kern<<<nb, kBBS, 0, stream()>>>(B->nnzL, B->smeta.as<int4>(), B->spend.as<int>(),
                                         B->spairs.as<int4>(), B->AL.as<TC>(), B->AU.as<TC>(),
                                         B->Lv[c].as<TS>(), B->Uv[c].as<TS>(), B->Lv[o].as<TS>(),
                                         B->Uv[o].as<TS>(), B->rs.as<float>(), B->partial.as<double>(),
                                         B->swhalo.as<uint8_t>());
      };
      if constexpr (can_wave) {
        if (B->wave_local) {
          if (with_step) run(bilu_sym_sweep_wave_kernel<BS, TC, TS, true, true>);
          else run(bilu_sym_sweep_wave_kernel<BS, TC, TS, false, true>);
        } else {
          if (with_step) run(bilu_sym_sweep_wave_kernel<BS, TC, TS, true, false>);
          else run(bilu_sym_sweep_wave_kernel<BS, TC, TS, false, false>);
        }
      }
    } else if (tables) {
      const bool stage = can_stage && B->variant == 4;
      const bool stage_lu = can_stage_lu && (B->variant == 0 || B->variant == 5);
      auto run = [&](auto kern) {
        kern<<<nb, kBBS, 0, stream()>>>(B->nnzL, B->smeta.as<int4>(), B->spend.as<int>(),
                                         B->spairs.as<int4>(), B->AL.as<TC>(), B->AU.as<TC>(),
                                         B->Lv[c].as<TS>(), B->Uv[c].as<TS>(), B->Lv[o].as<TS>(),
                                         B->Uv[o].as<TS>(), B->rs.as<float>(), B->partial.as<double>());
      };
      constexpr int full = can_stage ? 2 : 0, lu = can_stage_lu ? 1 : 0;
      if (stage) {
        if (with_step) run(bilu_sym_sweep_meta_kernel<BS, TC, TS, true, full>);
        else run(bilu_sym_sweep_meta_kernel<BS, TC, TS, false, full>);
      } else if (stage_lu) {
        if (with_step) run(bilu_sym_sweep_meta_kernel<BS, TC, TS, true, lu>);
        else run(bilu_sym_sweep_meta_kernel<BS, TC, TS, false, lu>);
      } else {
        if (with_step) run(bilu_sym_sweep_meta_kernel<BS, TC, TS, true, 0>);
        else run(bilu_sym_sweep_meta_kernel<BS, TC, TS, false, 0>);
      }
    } else {
      if (with_step) go(bilu_sym_sweep_kernel<BS, TC, TS, true>);
      else go(bilu_sym_sweep_kernel<BS, TC, TS, false>);
    }
  } else {
    nb = group_blocks<BS>(B->nblocksA);
    auto go = [&](auto kern) {
      kern<<<nb, kBBS, 0, stream()>>>(B->nblocksA, B->arowidx.as<int>(), B->acol.as<int>(),
                                       B->AV.as<TC>(), B->lptr.as<int>(), B->lcol.as<int>(),
                                       B->ucp.as<int>(), B->uri.as<int>(), B->Lv[c].as<TS>(),
                                       B->Uv[c].as<TS>(), B->Lv[o].as<TS>(), B->Uv[o].as<TS>(),
                                       B->rs.as<float>(), B->partial.as<double>());
    };
    if (with_step) go(bilu_gen_sweep_kernel<BS, TC, TS, true>);
    else go(bilu_gen_sweep_kernel<BS, TC, TS, false>);
  }
  LACE_HIP(hipGetLastError());
  B->last_nb = nb;
  B->cur = o;
  B->dirty = true;
  return DEV_SUCCESS;
}

template <typename TC, typename TS>
static int sweep_t(BILU* B, bool with_step) {
  switch (B->bs) {
    case 1: return sweep_bs<1, TC, TS>(B, with_step);
    case 2: return sweep_bs<2, TC, TS>(B, with_step);
    case 3: return sweep_bs<3, TC, TS>(B, with_step);
    case 4: return sweep_bs<4, TC, TS>(B, with_step);
    case 5: return sweep_bs<5, TC, TS>(B, with_step);
    case 6: return sweep_bs<6, TC, TS>(B, with_step);
    case 7: return sweep_bs<7, TC, TS>(B, with_step);
    default: return sweep_bs<8, TC, TS>(B, with_step);
  }
}

int BILU::sweep(bool with_step) {
  TimedRegion tr("bilu_sweep");
  switch (store) {
    case LACE_F64: return sweep_t<double, double>(this, with_step);
    case LACE_F32: return sweep_t<float, float>(this, with_step);
    case LACE_BF16: return sweep_t<float, uint16_t>(this, with_step);
    default: return sweep_t<float, _Float16>(this, with_step);
  }
}

int BILU::read_step(double* step) {
  bilu_sum_kernel<<<1, 256, 0, stream()>>>(last_nb, partial.as<double>(), scalar.as<double>());
  LACE_HIP(hipGetLastError());
  LACE_HIP(hipMemcpyAsync(step, scalar.p, sizeof(double), hipMemcpyDeviceToHost, stream()));
  LACE_HIP(hipStreamSynchronize(stream()));
  return DEV_SUCCESS;
}

double BILU::sweep_bytes() const {
  // A's blocks once, old L (strict) and U read once, new ones written once,
  // the pattern (two ints per entry updated), and row scales for 16-bit
  // storage.
  const double lb = (double)bs * bs;
  const double a = (double)nblocksA * lb * celem();
  const double f = 2.0 * ((double)nstrict + nnzU) * lb * selem();
  return a + f + 8.0 * nblocksA + 4.0 * (nbr + 1) + (is16s() ? 4.0 * n() : 0.0);
}

template <int BS, typename TC, typename TS>
static int dinv_bs(BILU* B) {
  bilu_dinv_kernel<BS, TC, TS><<<ceil_div(B->nbr, kBBS), kBBS, 0, stream()>>>(
      B->nbr, B->ucp.as<int>(), B->Uv[B->cur].as<TS>(), B->rs.as<float>(), B->Dinv.as<TC>());
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

template <typename TC, typename TS>
static int dinv_t(BILU* B) {
  switch (B->bs) {
    case 1: return dinv_bs<1, TC, TS>(B);
    case 2: return dinv_bs<2, TC, TS>(B);
    case 3: return dinv_bs<3, TC, TS>(B);
    case 4: return dinv_bs<4, TC, TS>(B);
    case 5: return dinv_bs<5, TC, TS>(B);
    case 6: return dinv_bs<6, TC, TS>(B);
    case 7: return dinv_bs<7, TC, TS>(B);
    default: return dinv_bs<8, TC, TS>(B);
  }
}

int BILU::refresh() {
  if (!dirty) return DEV_SUCCESS;
  switch (store) {
    case LACE_F64: LACE_TRY((dinv_t<double, double>(this))); break;
    case LACE_F32: LACE_TRY((dinv_t<float, float>(this))); break;
    case LACE_BF16: LACE_TRY((dinv_t<float, uint16_t>(this))); break;
    default: LACE_TRY((dinv_t<float, _Float16>(this))); break;
  }
  dirty = false;
  return DEV_SUCCESS;
}

template <int BS, typename TC, typename TS>
static int apply_bs(BILU* B, lace_trsv_kind_t kind, int sweeps, const TC* r, TC* z) {
  const int nbr = B->nbr;
  const int* lptr = B->lptr.as<int>();
  const int* lcol = B->lcol.as<int>();
  const int* urp = B->urp.as<int>();
  const int* uidx = B->uidx.as<int>();
  const int* ucolk = B->ucolk.as<int>();
  const float* rs = B->rs.as<float>();
  const TS* Lv = B->Lv[B->cur].as<TS>();
  const TS* Uv = B->Uv[B->cur].as<TS>();
  const TC* Dinv = B->Dinv.as<TC>();
  TC* w0 = B->work.as<TC>();
  TC* w1 = w0 + (size_t)B->n();
  if (kind == LACE_TRSV_LEVEL) {
    {
      TimedRegion tr("btrsv_lower");
      for (size_t l = 0; l + 1 < B->h_llev.size(); l++) {
        const int cnt = B->h_llev[l + 1] - B->h_llev[l];
        btri_lower_kernel<BS, TC, TS, true><<<group_blocks<BS>(cnt), kBBS, 0, stream()>>>(
            cnt, B->lperm.as<int>() + B->h_llev[l], lptr, lcol, Lv, rs, r, w0, w0);
      }
      LACE_HIP(hipGetLastError());
    }
    TimedRegion tr("btrsv_upper");
    for (size_t l = 0; l + 1 < B->h_ulev.size(); l++) {
      const int cnt = B->h_ulev[l + 1] - B->h_ulev[l];
      btri_upper_kernel<BS, TC, TS, true, false><<<group_blocks<BS>(cnt), kBBS, 0, stream()>>>(
          cnt, B->uperm.as<int>() + B->h_ulev[l], urp, uidx, ucolk, Uv, rs, Dinv, w0, z, z);
    }
    LACE_HIP(hipGetLastError());
    return DEV_SUCCESS;
  }
  LACE_REQUIRE(kind == LACE_TRSV_JACOBI && sweeps >= 1, DEV_ERR_ILLEGAL_VALUE,
               "block trsv: kind %d with %d sweeps", (int)kind, sweeps);
  const int nb = group_blocks<BS>(nbr);
  // lower: y1 = r, then sweeps-1 Jacobi updates alternating w0/w1
  TC* y = w0;
  {
    TimedRegion tr("btrsv_jacobi_lower");
    LACE_HIP(hipMemcpyAsync(w0, r, sizeof(TC) * B->n(), hipMemcpyDeviceToDevice, stream()));
    for (int s = 1; s < sweeps; s++) {
      TC* nxt = (y == w0) ? w1 : w0;
      btri_lower_kernel<BS, TC, TS, false><<<nb, kBBS, 0, stream()>>>(nbr, nullptr, lptr, lcol, Lv, rs,
                                                                      r, y, nxt);
      y = nxt;
    }
    LACE_HIP(hipGetLastError());
  }
  // upper on rhs y: iterates u_1..u_k with u_k landing in z
  TimedRegion tr("btrsv_jacobi_upper");
  TC* other = (y == w0) ? w1 : w0;
  for (int s = 1; s <= sweeps; s++) {
    TC* dst = ((sweeps - s) % 2 == 0) ? z : other;
    TC* src = (dst == z) ? other : z;
    if (s == 1)
      btri_upper_kernel<BS, TC, TS, false, true><<<nb, kBBS, 0, stream()>>>(
          nbr, nullptr, urp, uidx, ucolk, Uv, rs, Dinv, y, nullptr, dst);
    else
      btri_upper_kernel<BS, TC, TS, false, false><<<nb, kBBS, 0, stream()>>>(
          nbr, nullptr, urp, uidx, ucolk, Uv, rs, Dinv, y, src, dst);
  }
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

template <typename TC, typename TS>
static int apply_t(BILU* B, lace_trsv_kind_t kind, int sweeps, const void* r, void* z) {
  const TC* rr = static_cast<const TC*>(r);
  TC* zz = static_cast<TC*>(z);
  switch (B->bs) {
    case 1: return apply_bs<1, TC, TS>(B, kind, sweeps, rr, zz);
    case 2: return apply_bs<2, TC, TS>(B, kind, sweeps, rr, zz);
    case 3: return apply_bs<3, TC, TS>(B, kind, sweeps, rr, zz);
    case 4: return apply_bs<4, TC, TS>(B, kind, sweeps, rr, zz);
    case 5: return apply_bs<5, TC, TS>(B, kind, sweeps, rr, zz);
    case 6: return apply_bs<6, TC, TS>(B, kind, sweeps, rr, zz);
    case 7: return apply_bs<7, TC, TS>(B, kind, sweeps, rr, zz);
    default: return apply_bs<8, TC, TS>(B, kind, sweeps, rr, zz);
  }
}

int BILU::apply(lace_trsv_kind_t kind, int sweeps, const void* r, void* z) {
  LACE_TRY(refresh());
  switch (store) {
    case LACE_F64: return apply_t<double, double>(this, kind, sweeps, r, z);
    case LACE_F32: return apply_t<float, float>(this, kind, sweeps, r, z);
    case LACE_BF16: return apply_t<float, uint16_t>(this, kind, sweeps, r, z);
    default: return apply_t<float, _Float16>(this, kind, sweeps, r, z);
  }
}

// Host copies of the factors in the reference's containers: L as BCSRL
// (identity diagonal blocks written out), U as BCSCU.  Values are the
// stored ones, widened to double.
int BILU::get(data_d_matrix* L, data_d_matrix* U) {
  const size_t lb = (size_t)bs * bs;
  std::vector<float> h_rs((size_t)n());
  LACE_TRY(download(h_rs.data(), rs, sizeof(float) * h_rs.size()));
  auto widen = [&](const DevBuf& src, int count, const std::vector<int>& rowblk,
                   std::vector<double>& dst) -> int {
    const size_t nv = (size_t)count * lb;
    dst.assign(nv, 0.0);
    if (store == LACE_F64) return download(dst.data(), src, sizeof(double) * nv);
    if (store == LACE_F32) {
      std::vector<float> f(nv);
      LACE_TRY(download(f.data(), src, sizeof(float) * nv));
      for (size_t e = 0; e < nv; e++) dst[e] = f[e];
      return DEV_SUCCESS;
    }
    std::vector<uint16_t> h(nv);
    LACE_TRY(download(h.data(), src, sizeof(uint16_t) * nv));
    for (int q = 0; q < count; q++)
      for (size_t e = 0; e < lb; e++) {
        const uint16_t bits = h[(size_t)q * lb + e];
        float v;
        if (store == LACE_BF16) {
          const uint32_t u = (uint32_t)bits << 16;
          memcpy(&v, &u, 4);
        } else {
          _Float16 hv;
          memcpy(&hv, &bits, 2);
          v = (float)hv;
        }
        dst[(size_t)q * lb + e] = (double)(v * h_rs[(size_t)rowblk[q] * bs + e / bs]);
      }
    return DEV_SUCCESS;
  };
  std::vector<double> lv, uv;
  LACE_TRY(widen(Lv[cur], nnzL, h_lrow, lv));
  LACE_TRY(widen(Uv[cur], nnzU, h_uri, uv));
  data_zmfree(L);
  data_zmfree(U);
  auto fill = [&](data_d_matrix* M, data_storage_t st, data_uplo_t fm, data_order_t mj,
                  data_diagorder_t dg, const std::vector<int>& ptr, const std::vector<int>& idx) -> int {
    const int cnt = ptr[nbr];
    M->storage_type = st;
    M->fill_mode = fm;
    M->major = mj;
    M->diagorder_type = dg;
    M->num_rows = M->num_cols = nbr;
    M->blocksize = bs;
    M->ldblock = (int)lb;
    M->numblocks = cnt;
    M->nnz = M->true_nnz = (int)(cnt * lb);
    LACE_TRY(host_calloc(&M->row, nbr + 1));
    LACE_TRY(host_calloc(&M->col, std::max(cnt, 1)));
    LACE_TRY(host_calloc(&M->val, (size_t)std::max(cnt, 1) * lb));
    memcpy(M->row, ptr.data(), sizeof(int) * (nbr + 1));
    memcpy(M->col, idx.data(), sizeof(int) * cnt);
    return DEV_SUCCESS;
  };
  // the reference tags its BCSRL result Magma_BCSR (magma_zmconverter.cpp:1251)
  LACE_TRY(fill(L, Magma_BCSR, MagmaLower, MagmaRowMajor, Magma_UNITY, h_lptr, h_lcol));
  LACE_TRY(fill(U, Magma_BCSCU, MagmaUpper, MagmaColMajor, Magma_VALUE, h_ucp, h_uri));
  for (int q = 0; q < nnzL; q++) {
    const bool diag = h_lrow[q] == h_lcol[q];
    for (size_t e = 0; e < lb; e++)
      L->val[(size_t)q * lb + e] = diag ? ((e / bs == e % bs) ? 1.0 : 0.0) : lv[(size_t)q * lb + e];
  }
  memcpy(U->val, uv.data(), sizeof(double) * uv.size());
  return DEV_SUCCESS;
}

}  // namespace lace

using namespace lace;

extern "C" {

int lace_bcsr_create(const data_d_matrix* A, lace_prec_t prec, lace_bcsr** out) {
  auto* M = new DevBCSR();
  const int rc = M->create(A, prec);
  if (rc != DEV_SUCCESS) {
    delete M;
    return rc;
  }
  *out = reinterpret_cast<lace_bcsr*>(M);
  return DEV_SUCCESS;
}

int lace_bcsr_destroy(lace_bcsr* A) {
  delete reinterpret_cast<DevBCSR*>(A);
  return DEV_SUCCESS;
}

int lace_bcsr_spmv(const lace_bcsr* Ah, const void* x, void* y) {
  LACE_TRY(ensure_init());
  const DevBCSR& A = *reinterpret_cast<const DevBCSR*>(Ah);
  if (A.prec == LACE_F64)
    return bcsr_spmv<double, double>(A, static_cast<const double*>(x), static_cast<double*>(y));
  return bcsr_spmv<float, float>(A, static_cast<const float*>(x), static_cast<float*>(y));
}

double lace_bcsr_spmv_bytes(const lace_bcsr* Ah) {
  const DevBCSR& A = *reinterpret_cast<const DevBCSR*>(Ah);
  const int s = A.prec == LACE_F64 ? 8 : 4;
  return A.bytes(s, s, false);
}

int lace_bilu_create(const data_d_matrix* A, lace_prec_t prec, lace_bilu** out) {
  auto* B = new BILU();
  const int rc = B->create(A, prec);
  if (rc != DEV_SUCCESS) {
    delete B;
    return rc;
  }
  *out = reinterpret_cast<lace_bilu*>(B);
  return DEV_SUCCESS;
}

int lace_bilu_destroy(lace_bilu* B) {
  delete reinterpret_cast<BILU*>(B);
  return DEV_SUCCESS;
}

int lace_bilu_reset(lace_bilu* B) {
  LACE_TRY(ensure_init());
  return reinterpret_cast<BILU*>(B)->reset();
}

int lace_bilu_sweeps(lace_bilu* Bh, int nsweeps, double* step) {
  LACE_TRY(ensure_init());
  auto* B = reinterpret_cast<BILU*>(Bh);
  for (int s = 0; s < nsweeps; s++) LACE_TRY(B->sweep(step != nullptr && s == nsweeps - 1));
  if (step) {
    if (nsweeps > 0) LACE_TRY(B->read_step(step));
    else *step = 0.0;
  }
  return DEV_SUCCESS;
}

int lace_bilu_get(lace_bilu* B, data_d_matrix* L, data_d_matrix* U) {
  LACE_TRY(ensure_init());
  return reinterpret_cast<BILU*>(B)->get(L, U);
}

double lace_bilu_sweep_bytes(const lace_bilu* B) {
  return reinterpret_cast<const BILU*>(B)->sweep_bytes();
}

int lace_bilu_apply(lace_bilu* B, lace_trsv_kind_t kind, int jacobi_sweeps, const void* r, void* z) {
  LACE_TRY(ensure_init());
  return reinterpret_cast<BILU*>(B)->apply(kind, jacobi_sweeps, r, z);
}

int lace_bilu_levels(const lace_bilu* Bh, int* nlev_L, int* nlev_U) {
  const auto* B = reinterpret_cast<const BILU*>(Bh);
  if (nlev_L) *nlev_L = (int)B->h_llev.size() - 1;
  if (nlev_U) *nlev_U = (int)B->h_ulev.size() - 1;
  return DEV_SUCCESS;
}

int lace_bilu_set_variant(lace_bilu* B, int variant) {
  LACE_TRY(ensure_init());
  return reinterpret_cast<BILU*>(B)->set_variant(variant);
}

int lace_bilu_is_symmetric(const lace_bilu* B) { return reinterpret_cast<const BILU*>(B)->sym ? 1 : 0; }

// src/parilu_v0_3_bcsr.cpp:16-276 (which defines the symbol
// data_PariLU_v0_3 a second time; here it has its own name).  Sweeps run
// on the device in FP64 until the unscaled step sum_(new-old)^2 is at most
// `reduction` (:61, :118); the reference has no sweep cap, this one stops
// at 1000 sweeps and returns DEV_NOTCONVERGED.  L is returned as BCSRL
// (identity diagonal blocks), U as BCSCU.
int data_PariLU_v0_3_bcsr(data_d_matrix* A, data_d_matrix* L, data_d_matrix* U, dataType reduction,
                          data_d_preconditioner_log* log) {
  LACE_TRY(ensure_init());
  auto t0 = std::chrono::steady_clock::now();
  BILU B;
  LACE_TRY(B.create(A, LACE_F64));
  double step = 1e300;
  int iter = 0;
  const int cap = 1000;
  while (step > reduction && iter < cap) {
    LACE_TRY(B.sweep(true));
    LACE_TRY(B.read_step(&step));
    iter++;
  }
  LACE_TRY(B.get(L, U));
  auto t1 = std::chrono::steady_clock::now();
  log->sweeps = iter;
  log->tol = reduction;
  log->finalStep = step;
  log->A_Frobenius = B.anorm;
  log->precond_generation_time = std::chrono::duration<double>(t1 - t0).count();
  log->initial_residual = 0.0;
  log->initial_nonlinear_residual = 0.0;
  return step <= reduction ? DEV_SUCCESS : DEV_NOTCONVERGED;
}

}  // extern "C"
