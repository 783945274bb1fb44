// tsgemm.hip -- tall-skinny FP64 products on MFMA for the multi-RHS
// (block) Krylov path.
//
// Operands are n x 32 blocks stored row-major (row i's 32 values
// contiguous), the layout of the multi-RHS vectors.  Two products:
//   tn: C_b = X_b^T W over rows [rs_b, n) for a list of blocks X_b
//       (32 x 32 result per block): per-slab partials, then a fixed-order
//       reduction -- deterministic.
//   nn: out = base - sum_b X_b D_b over rows [rlo, n) (D_b 32 x 32); rows
//       below rlo keep base.  base is a block, zero, or an identity block
//       column (1 at row e0 + c, column c).
// Both run on v_mfma_f64_16x16x4f64: a wave owns 16 x 16 output tiles and
// steps the contraction 4 at a time, with operands read from LDS tiles that
// the workgroup stages by coalesced 16-byte loads (whole 256-byte rows).  Operand layout (CDNA3/4): A
// (16 x 4): lane l holds A[l % 16][l / 16]; B (4 x 16): lane l holds
// B[l / 16][l % 16]; C (16 x 16): lane l holds rows 4*v + l/16, v < 4,
// of column l % 16.
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "tsgemm.hpp"

namespace lace {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma4(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int kPitch = 34;   // LDS row pitch in doubles: 16-byte aligned rows, banks spread
constexpr int kTnG = 8;      // X blocks per workgroup
constexpr int kTnRows = 16;  // rows per LDS chunk

// 16-byte word e of a 32-column FP64 row chunk: row e / 16, columns
// 2 (e % 16) and 2 (e % 16) + 1.
__device__ __forceinline__ double2 ld2(const double* p) { return *reinterpret_cast<const double2*>(p); }
__device__ __forceinline__ void st2(double* p, double2 v) { *reinterpret_cast<double2*>(p) = v; }

// grid (nslab, ngroups); workgroup (slab s, group g) handles rows
// [s*rpb, (s+1)*rpb) and blocks [8g, 8g+8).  Each 16-row chunk of W and of
// the group's X blocks is staged in LDS with coalesced 16-byte loads (a
// half-wave reads one whole 256-byte row); wave w then owns tile (w/2, w%2)
// of every block: A = X_b^T (lane l: X_b[row k0 + l/16][16 ah + l%16]),
// B = W (lane l: W[row k0 + l/16][16 bh + l%16]).
__global__ __launch_bounds__(256) void tn_partial_kernel(int n, int rpb, BlockList X,
                                                         const double* __restrict__ W, int rs_main,
                                                         int rs_last, double* __restrict__ partial) {
  __shared__ double Ws[kTnRows][kPitch];
  __shared__ double Xs[kTnG][kTnRows][kPitch];
  const int nb = X.nb;
  const int slab = blockIdx.x, g = blockIdx.y;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int ah = w >> 1, bh = w & 1;
  const int r_begin = slab * rpb, r_end = min(r_begin + rpb, n);
  const int b0 = g * kTnG, gb = min(kTnG, nb - b0);
  d4 acc[kTnG];
  const double* xb[kTnG];
  int rsb[kTnG];
#pragma unroll
  for (int q = 0; q < kTnG; q++) {
    acc[q] = d4{0.0, 0.0, 0.0, 0.0};
    const int b = b0 + q;
    xb[q] = q < gb ? static_cast<const double*>(X.p[b]) : nullptr;
    rsb[q] = b == nb - 1 ? rs_last : rs_main;
  }
  // staging: 16 rows x 16 words = 256 words per operand, one per thread;
  // the next chunk's words are loaded into registers while the MFMAs run on
  // the staged one
  const int srow = t >> 4, scol = (t & 15) * 2;
  const double2 z2 = make_double2(0.0, 0.0);
  double2 wv = z2, xv[kTnG];
  for (int r0 = r_begin - kTnRows; r0 < r_end; r0 += kTnRows) {
    if (r0 >= r_begin) {
      st2(&Ws[srow][scol], wv);
#pragma unroll
      for (int q = 0; q < kTnG; q++) st2(&Xs[q][srow][scol], xv[q]);
      __syncthreads();
    }
    {
      const int row = r0 + kTnRows + srow;  // the next chunk
      const bool in = row < r_end;
      wv = in ? ld2(W + (size_t)row * 32 + scol) : z2;
#pragma unroll
      for (int q = 0; q < kTnG; q++)
        xv[q] = (q < gb && in && row >= rsb[q]) ? ld2(xb[q] + (size_t)row * 32 + scol) : z2;
    }
    if (r0 < r_begin) continue;
#pragma unroll
    for (int k0 = 0; k0 < kTnRows; k0 += 4) {
      const int kr = k0 + (lane >> 4);
      const double bv = Ws[kr][bh * 16 + (lane & 15)];
#pragma unroll
      for (int q = 0; q < kTnG; q++)
        if (q < gb) acc[q] = mfma4(Xs[q][kr][ah * 16 + (lane & 15)], bv, acc[q]);
    }
    __syncthreads();
  }
  // partial[slab][b][32][32], tile (ah, bh): rows ah*16 + 4*v + l/16, col bh*16 + l%16
#pragma unroll
  for (int q = 0; q < kTnG; q++) {
    if (q >= gb) continue;
    double* P = partial + ((size_t)slab * nb + b0 + q) * 1024;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int rr = ah * 16 + 4 * v + (lane >> 4), cc = bh * 16 + (lane & 15);
      P[rr * 32 + cc] = acc[q][v];
    }
  }
}

// Two-stage fixed-order reduction over the slabs: stage 1 sums slab group
// g (blockIdx.y) with four independent accumulators, stage 2 adds the group
// sums in group order.
__global__ __launch_bounds__(256) void tn_reduce1_kernel(int nslab, int count, int per,
                                                         const double* __restrict__ partial,
                                                         double* __restrict__ mid) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= count) return;
  const int g = blockIdx.y;
  const int s0i = g * per, s1i = min(nslab, s0i + per);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int s = s0i;
  for (; s + 3 < s1i; s += 4) {
    s0 += partial[(size_t)s * count + e];
    s1 += partial[(size_t)(s + 1) * count + e];
    s2 += partial[(size_t)(s + 2) * count + e];
    s3 += partial[(size_t)(s + 3) * count + e];
  }
  for (; s < s1i; s++) s0 += partial[(size_t)s * count + e];
  mid[(size_t)g * count + e] = (s0 + s1) + (s2 + s3);
}

__global__ __launch_bounds__(256) void tn_reduce2_kernel(int ng, int count, const double* __restrict__ mid,
                                                         double* __restrict__ C) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= count) return;
  double s = 0.0;
  for (int g = 0; g < ng; g++) s += mid[(size_t)g * count + e];
  C[e] = s;
}

constexpr int kRedGroups = 16;

constexpr int kNnRows = 64;  // rows per workgroup

// Element-wise widening of a staged 16-byte word.
template <typename TX>
struct Stage;
template <>
struct Stage<double> {
  static constexpr int WORDS = 16;  // 16-byte words per 32-column row
  // a 16-byte word of the row, held in registers, into LDS as doubles
  __device__ __forceinline__ static void store(const uint4& u, double* dst) {
    st2(dst, __builtin_bit_cast(double2, u));
  }
};
template <>
struct Stage<float> {
  static constexpr int WORDS = 8;
  __device__ __forceinline__ static void store(const uint4& u, double* dst) {
    const float4 v = __builtin_bit_cast(float4, u);
    st2(dst, make_double2(v.x, v.y));
    st2(dst + 2, make_double2(v.z, v.w));
  }
};

// A workgroup owns a band of 64 rows; for each block the band of X_b and
// D_b are staged in LDS (coalesced 16-byte loads), then wave w computes
// rows band + 16 w .. +15 in both 16-column halves: A = X_b (lane l:
// X_b[16 w + l%16][k0 + l/16]), B = D_b (lane l: D_b[k0 + l/16][16 ch + l%16]).
template <typename TX, typename TO>
__global__ __launch_bounds__(256) void nn_kernel(int n, BlockList X,
                                                 const double* __restrict__ D, const double* base,
                                                 int base_mode, int e0, int rlo, TO* out) {
  __shared__ double Xs[kNnRows][kPitch];
  __shared__ double Ds[32][kPitch];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int band = xcd_block(blockIdx.x, gridDim.x) * kNnRows;
  constexpr int WPR = Stage<TX>::WORDS, EPW = 32 / WPR;  // words per row, elements per word
  d4 acc0 = d4{0.0, 0.0, 0.0, 0.0}, acc1 = d4{0.0, 0.0, 0.0, 0.0};
  // block b + 1's words of the band and of D are loaded into registers
  // while the MFMAs run on block b's staged ones
  constexpr int NXW = kNnRows * WPR / 256, NDW = 512 / 256;
  uint4 xw[NXW], dw[NDW];
  for (int b = -1; b < X.nb; b++) {
    if (b >= 0) {
#pragma unroll
      for (int k = 0; k < NXW; k++) {
        const int e = t + 256 * k;
        Stage<TX>::store(xw[k], &Xs[e / WPR][(e % WPR) * EPW]);
      }
#pragma unroll
      for (int k = 0; k < NDW; k++) {
        const int e = t + 256 * k;
        st2(&Ds[e >> 4][(e & 15) * 2], __builtin_bit_cast(double2, dw[k]));
      }
      __syncthreads();
    }
    if (b + 1 < X.nb) {
      const TX* Xb = static_cast<const TX*>(X.p[b + 1]);
      const double* Db = D + (size_t)(b + 1) * 1024;
#pragma unroll
      for (int k = 0; k < NXW; k++) {
        const int e = t + 256 * k;
        const int row = band + e / WPR, cw = (e % WPR) * EPW;
        xw[k] = make_uint4(0u, 0u, 0u, 0u);
        if (row < n && row >= rlo) xw[k] = *reinterpret_cast<const uint4*>(Xb + (size_t)row * 32 + cw);
      }
#pragma unroll
      for (int k = 0; k < NDW; k++) {
        const int e = t + 256 * k;
        dw[k] = *reinterpret_cast<const uint4*>(Db + (e >> 4) * 32 + (e & 15) * 2);
      }
    }
    if (b < 0) continue;
#pragma unroll
    for (int k0 = 0; k0 < 32; k0 += 4) {
      const int kk = k0 + (lane >> 4);
      const double av = Xs[16 * w + (lane & 15)][kk];
      acc0 = mfma4(av, Ds[kk][lane & 15], acc0);
      acc1 = mfma4(av, Ds[kk][16 + (lane & 15)], acc1);
    }
    __syncthreads();
  }
  // C layout: rows band + 16 w + 4 v + l/16, columns 16 ch + l%16
#pragma unroll
  for (int ch = 0; ch < 2; ch++) {
    const int col = ch * 16 + (lane & 15);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int row = band + 16 * w + 4 * v + (lane >> 4);
      if (row >= n) continue;
      double bval = 0.0;
      if (base_mode == 1) bval = base[(size_t)row * 32 + col];
      else if (base_mode == 2) bval = (row - e0 == col) ? 1.0 : 0.0;
      const double a = ch == 0 ? acc0[v] : acc1[v];
      out[(size_t)row * 32 + col] = (TO)(row >= rlo ? bval - a : bval);
    }
  }
}

int TsGemm::init(int n_) {
  n = n_;
  // about 1024 row slabs, a multiple of the 16-row staging chunk each
  rpb = std::max(256, ((n + 1023) / 1024 + 15) / 16 * 16);
  nslab = ceil_div(n, rpb);
  LACE_TRY(partial.alloc(sizeof(double) * 1024 * (size_t)nslab * kMaxBlocks));
  LACE_TRY(mid.alloc(sizeof(double) * 1024 * (size_t)kRedGroups * kMaxBlocks));
  return DEV_SUCCESS;
}

int TsGemm::tn(const BlockList& L, const double* W, int rs_main, int rs_last, double* C) {
  LACE_REQUIRE(L.nb >= 1 && L.nb <= kMaxBlocks, DEV_ERR_ILLEGAL_VALUE, "tsgemm: %d blocks", L.nb);
  // algorithmic work: 2*32*32 flops per block row; each X_b row and each W
  // row read once, the 32 x 32 results written once
  const double rows_x = (double)(L.nb - 1) * (n - rs_main) + (double)(n - rs_last);
  TimedRegion tr("tsgemm_tn", 2048.0 * rows_x,
                 256.0 * (rows_x + (n - std::min(rs_main, rs_last))) + 8192.0 * L.nb);
  dim3 grid(nslab, ceil_div(L.nb, kTnG));
  tn_partial_kernel<<<grid, 256, 0, stream()>>>(n, rpb, L, W, rs_main, rs_last, partial.as<double>());
  const int count = L.nb * 1024;
  const int per = ceil_div(nslab, kRedGroups), ng = ceil_div(nslab, per);
  tn_reduce1_kernel<<<dim3(ceil_div(count, 256), ng), 256, 0, stream()>>>(nslab, count, per,
                                                                         partial.as<double>(), mid.as<double>());
  tn_reduce2_kernel<<<ceil_div(count, 256), 256, 0, stream()>>>(ng, count, mid.as<double>(), C);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

template <typename TX, typename TO>
int TsGemm::nn(const BlockList& L, const double* D, const double* base, int base_mode, int e0, int rlo,
               TO* out) {
  LACE_REQUIRE(L.nb >= 0 && L.nb <= kMaxBlocks, DEV_ERR_ILLEGAL_VALUE, "tsgemm: %d blocks", L.nb);
  // algorithmic work: 2*32*32 flops per block row below rlo; X_b rows from
  // rlo read once, D once, out written once, base read once (mode 1)
  const double rows = (double)std::max(0, n - rlo);
  TimedRegion tr("tsgemm_nn", 2048.0 * L.nb * rows,
                 32.0 * sizeof(TX) * L.nb * rows + 8192.0 * L.nb + 32.0 * sizeof(TO) * n +
                     (base_mode == 1 ? 256.0 * n : 0.0));
  nn_kernel<TX, TO><<<ceil_div(n, kNnRows), 256, 0, stream()>>>(n, L, D, base, base_mode, e0, rlo, out);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

template int TsGemm::nn<double, double>(const BlockList&, const double*, const double*, int, int, int,
                                        double*);
template int TsGemm::nn<double, float>(const BlockList&, const double*, const double*, int, int, int,
                                       float*);
template int TsGemm::nn<float, double>(const BlockList&, const double*, const double*, int, int, int,
                                       double*);

}  // namespace lace

using namespace lace;

extern "C" {

// Tall-skinny products on device arrays (blocks: host array of nb device
// pointers to n x 32 row-major blocks).
int lace_tsgemm_tn(int n, int nb, const double* const* blocks, const double* W, int rs_main,
                   int rs_last, double* C) {
  LACE_TRY(ensure_init());
  LACE_REQUIRE(nb >= 1 && nb <= kMaxBlocks, DEV_ERR_ILLEGAL_VALUE, "tsgemm: %d blocks", nb);
  TsGemm T;
  LACE_TRY(T.init(n));
  BlockList L;
  L.nb = nb;
  for (int b = 0; b < nb; b++) L.p[b] = blocks[b];
  LACE_TRY(T.tn(L, W, rs_main, rs_last, C));
  LACE_HIP(hipStreamSynchronize(stream()));
  return DEV_SUCCESS;
}

int lace_tsgemm_nn(int n, int nb, const double* const* blocks, const double* D, const double* base,
                   int base_mode, int e0, int rlo, double* out) {
  LACE_TRY(ensure_init());
  LACE_REQUIRE(nb >= 0 && nb <= kMaxBlocks, DEV_ERR_ILLEGAL_VALUE, "tsgemm: %d blocks", nb);
  TsGemm T;
  LACE_TRY(T.init(n));
  BlockList L;
  L.nb = nb;
  for (int b = 0; b < nb; b++) L.p[b] = blocks[b];
  LACE_TRY((T.nn<double, double>(L, D, base, base_mode, e0, rlo, out)));
  LACE_HIP(hipStreamSynchronize(stream()));
  return DEV_SUCCESS;
}

}  // extern "C"
