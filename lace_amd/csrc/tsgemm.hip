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
// Both run on v_mfma_f64_16x16x4f64: a wave owns one 16 x 16 output tile
// and steps the contraction 4 at a time.  Operand layout (CDNA3/4): A
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

constexpr int kTnBlocksPerGroup = 4;  // X blocks per workgroup (4 waves = 4 tiles each)

// grid (nslab, ngroups); workgroup (slab s, group g) handles rows
// [s*rpb, (s+1)*rpb) and blocks [4g, 4g+4).  Wave w owns tile (w/2, w%2)
// of every block of its group.
__global__ __launch_bounds__(256) void tn_partial_kernel(int n, int rpb, BlockList X,
                                                         const double* __restrict__ W, int rs_main,
                                                         int rs_last, double* __restrict__ partial) {
  const int nb = X.nb;
  const int slab = blockIdx.x, g = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ah = w >> 1, bh = w & 1;
  const int r_begin = slab * rpb, r_end = min(r_begin + rpb, n);
  d4 acc[kTnBlocksPerGroup];
  const double* xb[kTnBlocksPerGroup];
  int rsb[kTnBlocksPerGroup];
#pragma unroll
  for (int t = 0; t < kTnBlocksPerGroup; t++) {
    acc[t] = d4{0.0, 0.0, 0.0, 0.0};
    const int b = g * kTnBlocksPerGroup + t;
    xb[t] = b < nb ? static_cast<const double*>(X.p[b]) : nullptr;
    rsb[t] = b == nb - 1 ? rs_last : rs_main;
  }
  const int col_a = ah * 16 + (lane & 15), col_b = bh * 16 + (lane & 15);
  for (int r = r_begin; r < r_end; r += 4) {
    const int row = r + (lane >> 4);
    const bool in = row < r_end;
    const double bv = in ? W[(size_t)row * 32 + col_b] : 0.0;
#pragma unroll
    for (int t = 0; t < kTnBlocksPerGroup; t++) {
      if (!xb[t]) continue;
      const double av = (in && row >= rsb[t]) ? xb[t][(size_t)row * 32 + col_a] : 0.0;
      acc[t] = mfma4(av, bv, acc[t]);
    }
  }
  // partial[slab][b][32][32], tile (ah, bh): rows ah*16 + 4*v + l/16, col bh*16 + l%16
#pragma unroll
  for (int t = 0; t < kTnBlocksPerGroup; t++) {
    const int b = g * kTnBlocksPerGroup + t;
    if (b >= nb) continue;
    double* P = partial + ((size_t)slab * nb + b) * 1024;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int rr = ah * 16 + 4 * v + (lane >> 4), cc = bh * 16 + (lane & 15);
      P[rr * 32 + cc] = acc[t][v];
    }
  }
}

// C[b][e] = sum over slabs in order (four independent accumulators).
__global__ __launch_bounds__(256) void tn_reduce_kernel(int nslab, int count, const double* __restrict__ partial,
                                                        double* __restrict__ C) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= count) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int s = 0;
  for (; s + 3 < nslab; s += 4) {
    s0 += partial[(size_t)s * count + e];
    s1 += partial[(size_t)(s + 1) * count + e];
    s2 += partial[(size_t)(s + 2) * count + e];
    s3 += partial[(size_t)(s + 3) * count + e];
  }
  for (; s < nslab; s++) s0 += partial[(size_t)s * count + e];
  C[e] = (s0 + s1) + (s2 + s3);
}

// A workgroup owns a band of 32 rows; wave w computes the 16 x 16 tile of
// rows band + 16*(w/2) .. +15 and columns 16*(w%2) .. +15.
template <typename TX, typename TO>
__global__ __launch_bounds__(256) void nn_kernel(int n, BlockList X,
                                                 const double* __restrict__ D, const double* base,
                                                 int base_mode, int e0, int rlo, TO* out) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int band = xcd_block(blockIdx.x, gridDim.x) * 32;  // 32 rows per workgroup
  const int rt = band + 16 * (w >> 1);                      // tile's first row
  const int ch = w & 1;
  // A operand: X[rt + l%16][k0 + l/16]; B operand: D[k0 + l/16][ch*16 + l%16]
  const int arow = rt + (lane & 15);
  d4 acc = d4{0.0, 0.0, 0.0, 0.0};
  const bool arow_in = arow < n && arow >= rlo;
  for (int b = 0; b < X.nb; b++) {
    const TX* Xb = static_cast<const TX*>(X.p[b]);
    const double* Db = D + (size_t)b * 1024;
#pragma unroll
    for (int k0 = 0; k0 < 32; k0 += 4) {
      const int kk = k0 + (lane >> 4);
      const double av = arow_in ? (double)Xb[(size_t)arow * 32 + kk] : 0.0;
      const double bv = Db[kk * 32 + ch * 16 + (lane & 15)];
      acc = mfma4(av, bv, acc);
    }
  }
  // C layout: rows rt + 4*v + l/16, column ch*16 + l%16
  const int col = ch * 16 + (lane & 15);
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int row = rt + 4 * v + (lane >> 4);
    if (row >= n) continue;
    double bval = 0.0;
    if (base_mode == 1) bval = base[(size_t)row * 32 + col];
    else if (base_mode == 2) bval = (row - e0 == col) ? 1.0 : 0.0;
    out[(size_t)row * 32 + col] = (TO)(row >= rlo ? bval - acc[v] : bval);
  }
}

int TsGemm::init(int n_) {
  n = n_;
  // about 1024 row slabs, a multiple of 4 rows each
  rpb = std::max(256, ((n + 1023) / 1024 + 3) / 4 * 4);
  nslab = ceil_div(n, rpb);
  LACE_TRY(partial.alloc(sizeof(double) * 1024 * (size_t)nslab * kMaxBlocks));
  return DEV_SUCCESS;
}

int TsGemm::tn(const BlockList& L, const double* W, int rs_main, int rs_last, double* C) {
  LACE_REQUIRE(L.nb >= 1 && L.nb <= kMaxBlocks, DEV_ERR_ILLEGAL_VALUE, "tsgemm: %d blocks", L.nb);
  TimedRegion tr("tsgemm_tn");
  dim3 grid(nslab, ceil_div(L.nb, kTnBlocksPerGroup));
  tn_partial_kernel<<<grid, 256, 0, stream()>>>(n, rpb, L, W, rs_main, rs_last, partial.as<double>());
  const int count = L.nb * 1024;
  tn_reduce_kernel<<<ceil_div(count, 256), 256, 0, stream()>>>(nslab, count, partial.as<double>(), C);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

template <typename TX, typename TO>
int TsGemm::nn(const BlockList& L, const double* D, const double* base, int base_mode, int e0, int rlo,
               TO* out) {
  LACE_REQUIRE(L.nb >= 0 && L.nb <= kMaxBlocks, DEV_ERR_ILLEGAL_VALUE, "tsgemm: %d blocks", L.nb);
  TimedRegion tr("tsgemm_nn");
  nn_kernel<TX, TO><<<ceil_div(n, 32), 256, 0, stream()>>>(n, L, D, base, base_mode, e0, rlo, out);
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
