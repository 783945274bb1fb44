// binner.hip -- batched inner FP32 GMRES: 32 independent right-preconditioned
// GMRES(m) iterations in lockstep on n x 32 row-major FP32 blocks.
//
// This is the inner solver of the nested mixed-precision scheme (an FP64
// flexible outer iteration preconditioned by a fixed number of FP32 Krylov
// steps), here for the 32 columns of the block solver at once: the ILU
// sweeps and the FP32 product read the factors and A once for all columns,
// and every vector access is a whole 128-byte row.  Per step j (column c):
//   p_j = M^{-1} W_j, w = A p_j, h_ij = s_i (W_i . w) (i <= j, classical
//   Gram-Schmidt, FP64 sums), W_{j+1} = w - sum_i h_ij s_i W_i,
//   h_{j+1,j} = ||W_{j+1}||, Givens rotations on the Hessenberg column;
// after m steps z = sum_j p_j y_j with y from the triangular least squares.
// The basis is kept unnormalised past W_0 = v / ||v||: column k of the
// basis block holds W_k, and the scales s_k = 1 / ||W_k|| (s_0 = 1) enter
// the coefficients instead, so no pass rescales a block.  With p_j =
// M^{-1} W_j as stored, A p_j = sum_{i <= j+1} h_ij (s_i W_i) holds for the
// stored vectors, and z = P y is a GMRES iterate of the preconditioned
// operator.
// Dot products and norms are reduced per 1024-row slab and then over the
// slabs in a fixed order, so results do not depend on scheduling.
#include <algorithm>
#include <cmath>

#include "binner.hpp"
#include "mrhs.hpp"
#include "smalldense.hpp"

namespace lace {

// FP64 state per column c at [x * 32 + c]: g (m+1), cs (m), sn (m),
// R (m x m, R[i][j] at i*m + j), y (m), h (m+1), nrm2 (1), inv (1),
// scales s (m+1).
struct InnerLayout {
  int m;
  __host__ __device__ int g() const { return 0; }
  __host__ __device__ int cs() const { return (m + 1) * 32; }
  __host__ __device__ int sn() const { return cs() + m * 32; }
  __host__ __device__ int R() const { return sn() + m * 32; }
  __host__ __device__ int y() const { return R() + m * m * 32; }
  __host__ __device__ int h() const { return y() + m * 32; }
  __host__ __device__ int nrm2() const { return h() + (m + 1) * 32; }
  __host__ __device__ int inv() const { return nrm2() + 32; }  // 1 / ||w_c|| (0 for a zero norm)
  __host__ __device__ int s() const { return inv() + 32; }     // s_k = 1 / ||W_k|| of the basis
  __host__ __device__ int total() const { return s() + (m + 1) * 32; }
};

__device__ __forceinline__ void load4f(const float* p, double (&v)[4]) {
  const float4 a = *reinterpret_cast<const float4*>(p);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}

// partial[slab][i][c] = sum over the slab's rows of q_i[row][c] w[row][c].
// Thread t: column group 4 (t % 8), rows t / 8 + 32 k of the slab, one row
// per iteration (the loads of one row in flight per thread keep the
// register count, and so the waves per SIMD, up).  NQ = nq.
template <int NQ>
__global__ __launch_bounds__(256) void bi_dots_kernel(int n, int rpb, const float* __restrict__ Q,
                                                      size_t bstride, const float* __restrict__ w,
                                                      double* __restrict__ partial) {
  __shared__ double red[4][NQ][32];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, cg = t & 7, rg = t >> 3;
  const int slab = blockIdx.x;
  const int r0 = slab * rpb, r1 = min(n, r0 + rpb);
  double acc[NQ][4];
#pragma unroll
  for (int i = 0; i < NQ; i++)
#pragma unroll
    for (int e = 0; e < 4; e++) acc[i][e] = 0.0;
  for (int row = r0 + rg; row < r1; row += 32) {
    const float4 wa = *reinterpret_cast<const float4*>(w + (size_t)row * 32 + 4 * cg);
    float4 qa[NQ];
#pragma unroll
    for (int i = 0; i < NQ; i++)
      qa[i] = *reinterpret_cast<const float4*>(Q + i * bstride + (size_t)row * 32 + 4 * cg);
#pragma unroll
    for (int i = 0; i < NQ; i++) {
      acc[i][0] += (double)qa[i].x * (double)wa.x;
      acc[i][1] += (double)qa[i].y * (double)wa.y;
      acc[i][2] += (double)qa[i].z * (double)wa.z;
      acc[i][3] += (double)qa[i].w * (double)wa.w;
    }
  }
  constexpr int nq = NQ;
#pragma unroll
  for (int i = 0; i < NQ; i++) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
      double v = acc[i][e];
      v += __shfl_xor(v, 8, 64);
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      if (lane < 8) red[wv][i][4 * lane + e] = v;
    }
  }
  __syncthreads();
  for (int idx = t; idx < nq * 32; idx += 256) {
    const int i = idx >> 5, c = idx & 31;
    partial[((size_t)slab * nq + i) * 32 + c] = (red[0][i][c] + red[1][i][c]) + (red[2][i][c] + red[3][i][c]);
  }
}

// w <- w - sum_{i < nq} q_i h_i (per column); partial[slab][c] = sum of the
// squares of the new (FP32-rounded) w.  One row per iteration, as above.
template <int NQ>
__global__ __launch_bounds__(256) void bi_update_kernel(int n, int rpb, const float* __restrict__ Q,
                                                        size_t bstride, const double* __restrict__ h,
                                                        const double* __restrict__ sc,
                                                        const float* __restrict__ w, float* __restrict__ wn,
                                                        double* __restrict__ partial) {
  __shared__ double red[4][32];
  __shared__ __align__(16) double coef[NQ * 32];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, cg = t & 7, rg = t >> 3;
  const int slab = blockIdx.x;
  const int r0 = slab * rpb, r1 = min(n, r0 + rpb);
  // the multiplier of W_i in column c: h_ic s_i with h_ic = s_i (W_i . w),
  // formed once per workgroup
  for (int k = t; k < NQ * 32; k += 256) {
    const double si = sc[k];
    coef[k] = (si * h[k]) * si;
  }
  __syncthreads();
  double s2[4] = {0.0, 0.0, 0.0, 0.0};
  for (int row = r0 + rg; row < r1; row += 32) {
    const float4 wx = *reinterpret_cast<const float4*>(w + (size_t)row * 32 + 4 * cg);
    float4 qx[NQ];
#pragma unroll
    for (int i = 0; i < NQ; i++)
      qx[i] = *reinterpret_cast<const float4*>(Q + i * bstride + (size_t)row * 32 + 4 * cg);
    double x[4] = {wx.x, wx.y, wx.z, wx.w};
    // the multipliers are read from LDS in every row (a compiler-only
    // barrier keeps them from being hoisted into 8 NQ registers), which
    // leaves room for more waves per SIMD
    asm volatile("" ::: "memory");
#pragma unroll
    for (int i = 0; i < NQ; i++) {
      const double2 h01 = *reinterpret_cast<const double2*>(&coef[i * 32 + 4 * cg]);
      const double2 h23 = *reinterpret_cast<const double2*>(&coef[i * 32 + 4 * cg + 2]);
      x[0] -= (double)qx[i].x * h01.x;
      x[1] -= (double)qx[i].y * h01.y;
      x[2] -= (double)qx[i].z * h23.x;
      x[3] -= (double)qx[i].w * h23.y;
    }
    const float4 o = make_float4((float)x[0], (float)x[1], (float)x[2], (float)x[3]);
    *reinterpret_cast<float4*>(wn + (size_t)row * 32 + 4 * cg) = o;
    s2[0] += (double)o.x * o.x;
    s2[1] += (double)o.y * o.y;
    s2[2] += (double)o.z * o.z;
    s2[3] += (double)o.w * o.w;
  }
#pragma unroll
  for (int e = 0; e < 4; e++) {
    double v = s2[e];
    v += __shfl_xor(v, 8, 64);
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if (lane < 8) red[wv][4 * lane + e] = v;
  }
  __syncthreads();
  if (t < 32) partial[(size_t)slab * 32 + t] = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
}

// Launch the NQ instance for a runtime nq (1 .. kMaxInner + 1).
#define LACE_NQ_SWITCH(nq, CALL) \
  switch (nq) {                  \
    case 1: CALL(1); break;      \
    case 2: CALL(2); break;      \
    case 3: CALL(3); break;      \
    case 4: CALL(4); break;      \
    case 5: CALL(5); break;      \
    case 6: CALL(6); break;      \
    case 7: CALL(7); break;      \
    case 8: CALL(8); break;      \
    case 9: CALL(9); break;      \
    case 10: CALL(10); break;    \
    case 11: CALL(11); break;    \
    case 12: CALL(12); break;    \
    default: CALL(13); break;    \
  }

static void launch_dots(int nq, int nslab, int n, int rpb, const float* Q, size_t bs, const float* w,
                        double* part) {
#define LACE_DOTS(K) bi_dots_kernel<K><<<nslab, 256, 0, stream()>>>(n, rpb, Q, bs, w, part)
  LACE_NQ_SWITCH(nq, LACE_DOTS)
#undef LACE_DOTS
}

static void launch_update(int nq, int nslab, int n, int rpb, const float* Q, size_t bs, const double* h,
                          const double* sc, const float* w, float* wn, double* part) {
#define LACE_UPD(K) bi_update_kernel<K><<<nslab, 256, 0, stream()>>>(n, rpb, Q, bs, h, sc, w, wn, part)
  LACE_NQ_SWITCH(nq, LACE_UPD)
#undef LACE_UPD
}

// q = w * inv[c] per column (inv = 1 / ||w_c||, 0 where the norm is zero),
// four consecutive columns per thread: a streaming pass (the reciprocal is
// formed once per column by the small kernels).  m is a multiple of 4.
__global__ __launch_bounds__(256) void bi_scale_kernel(size_t m, const float* __restrict__ w,
                                                       const double* __restrict__ inv, float* __restrict__ q) {
  const size_t e4 = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e4 * 4 >= m) return;
  const int c0 = (int)((e4 * 4) & 31);
  const float4 a = *reinterpret_cast<const float4*>(w + e4 * 4);
  const float4 o = make_float4((float)((double)a.x * inv[c0]), (float)((double)a.y * inv[c0 + 1]),
                               (float)((double)a.z * inv[c0 + 2]), (float)((double)a.w * inv[c0 + 3]));
  *reinterpret_cast<float4*>(q + e4 * 4) = o;
}

__device__ __forceinline__ double inv_norm(double nrm) { return nrm > 0.0 ? 1.0 / nrm : 0.0; }

__global__ void bi_init_kernel(InnerLayout L, double* __restrict__ st) {
  const int c = threadIdx.x;
  if (c >= 32) return;
  const double nr = sqrt(st[L.nrm2() + c]);
  st[L.g() + c] = nr;
  st[L.inv() + c] = inv_norm(nr);
  st[L.s() + c] = 1.0;  // W_0 = v / ||v|| is stored normalised
  for (int i = 1; i <= L.m; i++) st[L.g() + i * 32 + c] = 0.0;
}

// Givens step j on each column's Hessenberg column h[0..j] (h_ij =
// s_i (W_i . w)) and h_{j+1,j} = ||W_{j+1}||, whose reciprocal is
// s_{j+1}.
__global__ void bi_givens_kernel(InnerLayout L, int j, double* __restrict__ st) {
  const int c = threadIdx.x;
  if (c >= 32) return;
  double h[kMaxInner + 1];
  for (int i = 0; i <= j; i++) h[i] = st[L.s() + i * 32 + c] * st[L.h() + i * 32 + c];
  h[j + 1] = sqrt(st[L.nrm2() + c]);
  st[L.inv() + c] = inv_norm(h[j + 1]);
  st[L.s() + (j + 1) * 32 + c] = inv_norm(h[j + 1]);
  for (int i = 0; i < j; i++) {
    const double cs = st[L.cs() + i * 32 + c], sn = st[L.sn() + i * 32 + c];
    const double a = cs * h[i] + sn * h[i + 1];
    h[i + 1] = -sn * h[i] + cs * h[i + 1];
    h[i] = a;
  }
  const double r = hypot(h[j], h[j + 1]);
  double cs = 1.0, sn = 0.0;
  if (r > 0.0) {
    cs = h[j] / r;
    sn = h[j + 1] / r;
  }
  h[j] = r;
  st[L.cs() + j * 32 + c] = cs;
  st[L.sn() + j * 32 + c] = sn;
  const double gj = st[L.g() + j * 32 + c];
  st[L.g() + (j + 1) * 32 + c] = -sn * gj;
  st[L.g() + j * 32 + c] = cs * gj;
  for (int i = 0; i <= j; i++) st[L.R() + (i * L.m + j) * 32 + c] = h[i];
}

// y = R^{-1} g on the first mu steps (a zero pivot gives y_j = 0).
__global__ void bi_solve_kernel(InnerLayout L, int mu, double* __restrict__ st) {
  const int c = threadIdx.x;
  if (c >= 32) return;
  for (int j = mu - 1; j >= 0; j--) {
    double s = st[L.g() + j * 32 + c];
    for (int i = j + 1; i < mu; i++) s -= st[L.R() + (j * L.m + i) * 32 + c] * st[L.y() + i * 32 + c];
    const double d = st[L.R() + (j * L.m + j) * 32 + c];
    st[L.y() + j * 32 + c] = d != 0.0 ? s / d : 0.0;
  }
}

// Z[row][c] = sum_{j < mu} P_j[row][c] y_j[c].
__global__ __launch_bounds__(256) void bi_combine_kernel(int n, int mu, const float* __restrict__ P, size_t bstride,
                                                         const double* __restrict__ y, float* __restrict__ Z) {
  const size_t e4 = (size_t)blockIdx.x * 256 + threadIdx.x;  // index of a 4-column group
  if (e4 >= (size_t)n * 8) return;
  const int cg = (int)(e4 & 7);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = 0; j < mu; j++) {
    double p[4];
    load4f(P + j * bstride + e4 * 4, p);
#pragma unroll
    for (int e = 0; e < 4; e++) acc[e] += p[e] * y[j * 32 + 4 * cg + e];
  }
  *reinterpret_cast<float4*>(Z + e4 * 4) = make_float4((float)acc[0], (float)acc[1], (float)acc[2], (float)acc[3]);
}

int InnerBatch::init(int n_, int m_) {
  LACE_REQUIRE(m_ >= 1 && m_ <= kMaxInner, DEV_ERR_ILLEGAL_VALUE, "inner steps %d (1..%d)", m_, kMaxInner);
  n = n_;
  m = m_;
  nslab = ceil_div(n, rpb);
  const size_t blk = (size_t)n * 32 * sizeof(float);
  LACE_TRY(Q.alloc(blk * (m + 1)));
  LACE_TRY(P.alloc(blk * m));
  LACE_TRY(w.alloc(blk));
  LACE_TRY(t1.alloc(blk));
  LACE_TRY(t2.alloc(blk));
  InnerLayout L{m};
  LACE_TRY(state.alloc(sizeof(double) * L.total()));
  LACE_TRY(partial.alloc(sizeof(double) * (size_t)nslab * (kMaxInner + 1) * 32));
  return DEV_SUCCESS;
}

int InnerBatch::solve(const DevCSR& A32, ParILU* F, lace_trsv_kind_t kind, int sweeps, int trans,
                      const float* V, float* Z) {
  TimedRegion tr("block_inner");
  const size_t bs = (size_t)n * 32;
  float* Qb = Q.as<float>();
  float* Pb = P.as<float>();
  float* wb = w.as<float>();
  double* st = state.as<double>();
  double* part = partial.as<double>();
  InnerLayout L{m};
  const int eb = ceil_div((long)bs / 4, 256);
  // beta_c = ||v_c||, q_0 = v / beta
  launch_dots(1, nslab, n, rpb, V, bs, V, part);
  LACE_TRY(sd_slab_reduce(nslab, 32, part, st + L.nrm2(), false));
  bi_init_kernel<<<1, 32, 0, stream()>>>(L, st);
  bi_scale_kernel<<<eb, 256, 0, stream()>>>(bs, V, st + L.inv(), Qb);
  for (int j = 0; j < m; j++) {
    float* pj = Pb + j * bs;
    LACE_TRY(block_precond<float>(F, kind, sweeps, trans, Qb + j * bs, pj, t1.as<float>(), t2.as<float>()));
    LACE_TRY(spmm32_f32(A32, pj, wb));
    {
      // algorithmic bytes: q_0..q_j and w read once
      TimedRegion td("inner_dots", 2.0 * (j + 1) * bs, 4.0 * (j + 2) * bs);
      launch_dots(j + 1, nslab, n, rpb, Qb, bs, wb, part);
    }
    LACE_TRY(sd_slab_reduce(nslab, (j + 1) * 32, part, st + L.h(), false));
    {
      // W_0..W_j and w read once, W_{j+1} written once into the basis
      // block (which has room for W_m, whose norm closes the last step)
      TimedRegion tu("inner_update", 2.0 * (j + 1) * bs, 4.0 * (j + 3) * bs);
      launch_update(j + 1, nslab, n, rpb, Qb, bs, st + L.h(), st + L.s(), wb, Qb + (j + 1) * bs, part);
    }
    LACE_TRY(sd_slab_reduce(nslab, 32, part, st + L.nrm2(), false));
    bi_givens_kernel<<<1, 32, 0, stream()>>>(L, j, st);
  }
  bi_solve_kernel<<<1, 32, 0, stream()>>>(L, m, st);
  bi_combine_kernel<<<ceil_div((long)n * 8, 256), 256, 0, stream()>>>(n, m, Pb, bs, st + L.y(), Z);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

}  // namespace lace
