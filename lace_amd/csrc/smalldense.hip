// smalldense.hip -- small FP64 dense kernels of the block Krylov path:
// Cholesky of 32 x 32 Gram matrices, Householder reconstruction from a
// CholeskyQR panel, the block Hessenberg window QR with its least-squares
// right-hand side, block back substitution, and column norms of n x 32
// blocks.  The 32 x 32 and 64 x 32 kernels run in one workgroup of 1024
// threads, thread t owning element (t / 32, t % 32), with the operands in
// LDS (rows padded to 33 doubles against bank conflicts).
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "device.hpp"
#include "smalldense.hpp"

namespace lace {

constexpr int kSd = 32;
constexpr double kUnitRoundoff = 1.1102230246251565e-16;  // 2^-53

// ---------------------------------------------------------------- gemm
__global__ __launch_bounds__(256) void sd_gemm_kernel(int M, int N, int K, double alpha,
                                                      const double* __restrict__ A, int lda, int ta,
                                                      const double* __restrict__ B, int ldb, int tb,
                                                      double beta, double* __restrict__ C, int ldc) {
  __shared__ double As[32][33], Bs[32][33];
  const int t = threadIdx.x;
  const int tr = t >> 3, tc = t & 7;
  const int bi = blockIdx.y * 32, bj = blockIdx.x * 32;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < K; k0 += 32) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int idx = t + 256 * e;
      const int rr = idx >> 5, cc = idx & 31;
      {
        const int i = bi + rr, k = k0 + cc;
        As[rr][cc] = (i < M && k < K) ? (ta ? A[(size_t)k * lda + i] : A[(size_t)i * lda + k]) : 0.0;
      }
      {
        const int k = k0 + rr, j = bj + cc;
        Bs[rr][cc] = (k < K && j < N) ? (tb ? B[(size_t)j * ldb + k] : B[(size_t)k * ldb + j]) : 0.0;
      }
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < 32; kk++) {
      const double a = As[tr][kk];
#pragma unroll
      for (int q = 0; q < 4; q++) acc[q] += a * Bs[kk][tc + 8 * q];
    }
    __syncthreads();
  }
  const int i = bi + tr;
  if (i >= M) return;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int j = bj + tc + 8 * q;
    if (j >= N) continue;
    double v = alpha * acc[q];
    if (beta != 0.0) v += beta * C[(size_t)i * ldc + j];
    C[(size_t)i * ldc + j] = v;
  }
}

int sd_gemm(int M, int N, int K, double alpha, const double* A, int lda, bool ta, const double* B,
            int ldb, bool tb, double beta, double* C, int ldc) {
  if (M <= 0 || N <= 0) return DEV_SUCCESS;
  dim3 grid(ceil_div(N, 32), ceil_div(M, 32));
  sd_gemm_kernel<<<grid, 256, 0, stream()>>>(M, N, K, alpha, A, lda, ta ? 1 : 0, B, ldb, tb ? 1 : 0, beta,
                                             C, ldc);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

// Inverse of an upper-triangular 32 x 32 matrix held in LDS, by the 32
// threads t < 32 (thread j solves column j by back substitution).
__device__ __forceinline__ void upper_inverse_cols(const double (*U)[33], double (*X)[33]) {
  const int j = threadIdx.x;
  if (j >= kSd) return;
  for (int i = kSd - 1; i >= 0; i--) {
    if (i > j) {
      X[i][j] = 0.0;
      continue;
    }
    double s = (i == j) ? 1.0 : 0.0;
    for (int q = i + 1; q <= j; q++) s -= U[i][q] * X[q][j];
    X[i][j] = s / U[i][i];
  }
}

// Inverse of a unit lower-triangular 32 x 32 matrix, threads t < 32.
__device__ __forceinline__ void unit_lower_inverse_cols(const double (*L)[33], double (*X)[33]) {
  const int j = threadIdx.x;
  if (j >= kSd) return;
  for (int i = 0; i < kSd; i++) {
    if (i < j) {
      X[i][j] = 0.0;
      continue;
    }
    double s = (i == j) ? 1.0 : 0.0;
    for (int q = j; q < i; q++) s -= L[i][q] * X[q][j];
    X[i][j] = s;
  }
}

// ---------------------------------------------------------------- Cholesky
__global__ __launch_bounds__(1024) void chol32_kernel(const double* __restrict__ G, double sigma_coef,
                                                      double* __restrict__ R, double* __restrict__ Rinv,
                                                      int* __restrict__ status) {
  __shared__ double A[32][33], X[32][33];
  __shared__ double diag[32];
  __shared__ int fail[2];
  __shared__ double shift;
  const int t = threadIdx.x, r = t >> 5, c = t & 31;
  const double g = G[r * 32 + c];
  if (r == c) diag[r] = g;
  if (t == 0) {
    shift = 0.0;
    fail[0] = fail[1] = 0;
  }
  __syncthreads();
  int st = 2;
  for (int attempt = 0; attempt < 2; attempt++) {
    A[r][c] = g + ((r == c) ? shift : 0.0);
    __syncthreads();
    for (int j = 0; j < kSd; j++) {
      if (t == 0) {
        const double d = A[j][j];
        const double ref = diag[j] + shift;
        if (!(d > 1e-13 * ref) || !isfinite(d)) fail[attempt] = 1;
        else A[j][j] = sqrt(d);
      }
      __syncthreads();
      if (fail[attempt]) break;  // uniform: read after the barrier, never reset
      if (r == j && c > j) A[j][c] = A[j][c] / A[j][j];
      __syncthreads();
      if (r > j && c >= r) A[r][c] -= A[j][r] * A[j][c];
      __syncthreads();
    }
    if (!fail[attempt]) {
      st = attempt;
      break;
    }
    if (attempt == 0) {
      if (t == 0) {
        double tr = 0.0;
        for (int q = 0; q < kSd; q++) tr += diag[q];
        shift = sigma_coef * tr;
      }
      __syncthreads();
    }
  }
  if (t == 0) {
    unsigned mask = 0;
    for (int q = 0; q < kSd; q++)
      if (diag[q] == 0.0) mask |= 1u << q;
    status[1] = (int)mask;
  }
  if (st == 2) {
    R[r * 32 + c] = 0.0;
    Rinv[r * 32 + c] = 0.0;
    if (t == 0) status[0] = 2;
    return;
  }
  if (c < r) A[r][c] = 0.0;
  __syncthreads();
  upper_inverse_cols(A, X);
  __syncthreads();
  R[r * 32 + c] = A[r][c];
  Rinv[r * 32 + c] = X[r][c];
  if (t == 0) status[0] = st;
}

int sd_chol32(const double* G, long mrows, double* R, double* Rinv, int* status) {
  const double coef = 11.0 * ((double)mrows * kSd + kSd * (kSd + 1.0)) * kUnitRoundoff;
  chol32_kernel<<<1, 1024, 0, stream()>>>(G, coef, R, Rinv, status);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

// ---------------------------------------------------------------- reconstruction
__global__ __launch_bounds__(1024) void recon32_kernel(const double* __restrict__ Qlast, int r0,
                                                       const double* __restrict__ Rl,
                                                       const double* __restrict__ Rlinv,
                                                       const double* __restrict__ Rprev,
                                                       double* __restrict__ Ltop, double* __restrict__ Tb,
                                                       double* __restrict__ Rt, double* __restrict__ Dy) {
  __shared__ double M[32][33], L[32][33], U[32][33], W1[32][33];
  __shared__ double S[32], piv[32];
  const int t = threadIdx.x, r = t >> 5, c = t & 31;
  // M = Qtop * Rl^{-1}: the top rows of the orthonormal panel
  W1[r][c] = Qlast[(size_t)(r0 + r) * 32 + c];
  U[r][c] = Rlinv[r * 32 + c];
  L[r][c] = 0.0;
  __syncthreads();
  {
    double s = 0.0;
    for (int q = 0; q <= c; q++) s += W1[r][q] * U[q][c];
    M[r][c] = s;
  }
  __syncthreads();
  // LU of I - M S with S_p = -sign(m~_pp) chosen at step p: pivot 1 + |m~_pp|
  for (int p = 0; p < kSd; p++) {
    if (t == 0) {
      const double mpp = M[p][p];
      S[p] = mpp >= 0.0 ? -1.0 : 1.0;
      piv[p] = 1.0 - S[p] * mpp;
    }
    __syncthreads();
    if (r > p) {
      const double l = -S[p] * M[r][p] / piv[p];
      if (c > p) M[r][c] -= l * M[p][c];
      else if (c == p) L[r][p] = l;
    }
    __syncthreads();
  }
  {
    const double lv = r > c ? L[r][c] : (r == c ? 1.0 : 0.0);
    const double uv = c > r ? -S[c] * M[r][c] : (c == r ? piv[r] : 0.0);
    __syncthreads();
    L[r][c] = lv;
    U[r][c] = uv;
  }
  __syncthreads();
  // W1 = U^{-1} (threads 0..31); M = L^{-1} (threads 32..63, shifted ids)
  if (t < 32) {
    upper_inverse_cols(U, W1);
  } else if (t < 64) {
    const int j = t - 32;
    for (int i = 0; i < kSd; i++) {
      if (i < j) {
        M[i][j] = 0.0;
        continue;
      }
      double s = (i == j) ? 1.0 : 0.0;
      for (int q = j; q < i; q++) s -= L[i][q] * M[q][j];
      M[i][j] = s;
    }
  }
  __syncthreads();
  // Tb = U L^{-T}
  {
    double s = 0.0;
    for (int q = r; q <= c; q++) s += U[r][q] * M[c][q];
    Tb[r * 32 + c] = s;
  }
  // Rt = S (Rl Rprev)
  {
    double s = 0.0;
    for (int q = r; q <= c; q++) s += Rl[r * 32 + q] * Rprev[q * 32 + c];
    Rt[r * 32 + c] = c >= r ? S[r] * s : 0.0;
  }
  // Dy = Rl^{-1} S U^{-1}
  {
    double s = 0.0;
    for (int q = r; q <= c; q++) s += Rlinv[r * 32 + q] * S[q] * W1[q][c];
    Dy[r * 32 + c] = s;
  }
  Ltop[r * 32 + c] = L[r][c];
}

int sd_recon32(const double* Qlast, int r0, const double* Rl, const double* Rlinv, const double* Rprev,
               double* Ltop, double* Tb, double* Rt, double* Dy) {
  recon32_kernel<<<1, 1024, 0, stream()>>>(Qlast, r0, Rl, Rlinv, Rprev, Ltop, Tb, Rt, Dy);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

__global__ __launch_bounds__(1024) void write_top_kernel(const double* __restrict__ Ltop, int r0,
                                                         double* __restrict__ Y) {
  const int t = threadIdx.x;
  Y[(size_t)r0 * 32 + t] = Ltop[t];
}

int sd_write_top(const double* Ltop, int r0, double* Y) {
  write_top_kernel<<<1, 1024, 0, stream()>>>(Ltop, r0, Y);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

// out (32 x 32 at leading dimension ldo) = alpha * in (32 x 32, ld 32);
// in == nullptr gives alpha * I.
__global__ __launch_bounds__(1024) void copy32_kernel(double alpha, const double* __restrict__ in,
                                                      double* __restrict__ out, int ldo) {
  const int t = threadIdx.x, r = t >> 5, c = t & 31;
  out[(size_t)r * ldo + c] = in ? alpha * in[t] : (r == c ? alpha : 0.0);
}

int sd_copy32(double alpha, const double* in, double* out, int ldo) {
  copy32_kernel<<<1, 1024, 0, stream()>>>(alpha, in, out, ldo);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

// Columns in `mask` of rows [r0, n) of W <- deterministic values in
// [-1, 1) (a 64-bit integer hash of the position).
__global__ __launch_bounds__(256) void fill_cols_kernel(int n, int r0, unsigned mask, double* __restrict__ W) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  const int i = r0 + (int)(e >> 5), c = (int)(e & 31);
  if (i >= n || !((mask >> c) & 1u)) return;
  unsigned long long h = ((unsigned long long)i << 5 | (unsigned)c) * 0x9E3779B97F4A7C15ull;
  h ^= h >> 31;
  h *= 0xBF58476D1CE4E5B9ull;
  h ^= h >> 29;
  W[(size_t)i * 32 + c] = (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

int sd_fill_cols(int n, int r0, unsigned mask, double* W) {
  if (!mask || n <= r0) return DEV_SUCCESS;
  fill_cols_kernel<<<ceil_div((long)(n - r0) * 32, 256), 256, 0, stream()>>>(n, r0, mask, W);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

__global__ __launch_bounds__(1024) void zero_cols32_kernel(unsigned mask, double* __restrict__ M) {
  const int t = threadIdx.x, c = t & 31;
  if ((mask >> c) & 1u) M[t] = 0.0;
}

int sd_zero_cols32(unsigned mask, double* M) {
  if (!mask) return DEV_SUCCESS;
  zero_cols32_kernel<<<1, 1024, 0, stream()>>>(mask, M);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

__global__ __launch_bounds__(1024) void gather_rows_kernel(BlockList Y, int e0, double* __restrict__ a) {
  const int b = blockIdx.x, t = threadIdx.x, r = t >> 5, c = t & 31;
  const double* Yb = static_cast<const double*>(Y.p[b]);
  a[(size_t)(b * 32 + c) * 32 + r] = Yb[(size_t)(e0 + r) * 32 + c];
}

int sd_gather_rows(const BlockList& Y, int e0, double* a) {
  if (Y.nb == 0) return DEV_SUCCESS;
  gather_rows_kernel<<<Y.nb, 1024, 0, stream()>>>(Y, e0, a);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

// ---------------------------------------------------------------- Hessenberg
// X (64 x 32) <- (I - V T V^T)^T X = X - V T^T (V^T X); P, Q scratch.
__device__ __forceinline__ void apply_window_t(double (*X)[33], const double (*V)[33], const double (*T)[33],
                                               double (*P)[33], double (*Q)[33]) {
  const int t = threadIdx.x, r = t >> 5, c = t & 31;
  {
    double s = 0.0;
    for (int q = 0; q < 64; q++) s += V[q][r] * X[q][c];
    P[r][c] = s;
  }
  __syncthreads();
  {
    double s = 0.0;
    for (int b = 0; b <= r; b++) s += T[b][r] * P[b][c];
    Q[r][c] = s;
  }
  __syncthreads();
  {
    double s0 = 0.0, s1 = 0.0;
    for (int a = 0; a < 32; a++) {
      s0 += V[r][a] * Q[a][c];
      s1 += V[r + 32][a] * Q[a][c];
    }
    X[r][c] -= s0;
    X[r + 32][c] -= s1;
  }
  __syncthreads();
}

__global__ __launch_bounds__(1024) void hess_step_kernel(int k, int ldh, double* __restrict__ Hm,
                                                         double* __restrict__ Vw, double* __restrict__ Tw,
                                                         double* __restrict__ Rdinv, double* __restrict__ Gls,
                                                         const double* __restrict__ W,
                                                         const double* __restrict__ Rt,
                                                         double* __restrict__ res) {
  __shared__ double X[64][33], V[64][33], T[32][33], P[32][33], Q[32][33];
  __shared__ double tau[32];
  __shared__ double sh_alpha, sh_scale;
  const int t = threadIdx.x, r = t >> 5, c = t & 31;
  const int col0 = k * 32;
  X[r][c] = W[(size_t)r * 32 + c];
  // windows 0..k-1 on the new column block
  for (int j = 0; j < k; j++) {
    X[32 + r][c] = W[(size_t)((j + 1) * 32 + r) * 32 + c];
    V[r][c] = Vw[(size_t)j * 2048 + r * 32 + c];
    V[r + 32][c] = Vw[(size_t)j * 2048 + (r + 32) * 32 + c];
    T[r][c] = Tw[(size_t)j * 1024 + r * 32 + c];
    __syncthreads();
    apply_window_t(X, V, T, P, Q);
    Hm[(size_t)(j * 32 + r) * ldh + col0 + c] = X[r][c];
    X[r][c] = X[32 + r][c];
    __syncthreads();
  }
  // window k: the carried rows on top, the subdiagonal block below
  X[32 + r][c] = Rt[r * 32 + c];
  __syncthreads();
  for (int i = 0; i < 32; i++) {
    if (t < 64) {
      const double v = t > i ? X[t][i] : 0.0;
      const double ss = wave_sum(v * v);
      if (t == 0) {
        const double x0 = X[i][i];
        if (ss == 0.0) {
          tau[i] = 0.0;
          sh_alpha = x0;
          sh_scale = 0.0;
        } else {
          const double nrm = sqrt(x0 * x0 + ss);
          const double alpha = x0 >= 0.0 ? -nrm : nrm;
          tau[i] = (alpha - x0) / alpha;
          sh_alpha = alpha;
          sh_scale = 1.0 / (x0 - alpha);
        }
      }
    }
    __syncthreads();
    if (t < 64) V[t][i] = t < i ? 0.0 : (t == i ? 1.0 : X[t][i] * sh_scale);
    __syncthreads();
    {
      const int j = t >> 5, p = t & 31;
      double w = V[p][i] * X[p][j] + V[p + 32][i] * X[p + 32][j];
      w = group_sum<32>(w);
      const double ti = tau[i];
      if (j > i) {
        X[p][j] -= ti * V[p][i] * w;
        X[p + 32][j] -= ti * V[p + 32][i] * w;
      } else if (j == i) {
        if (p == i) X[p][i] = sh_alpha;
        else if (p > i) X[p][i] = 0.0;
        X[p + 32][i] = 0.0;
      }
    }
    __syncthreads();
  }
  // compact WY factor of the window: T upper triangular, P = V^T V
  {
    double s = 0.0;
    for (int q = 0; q < 64; q++) s += V[q][r] * V[q][c];
    P[r][c] = s;
    T[r][c] = 0.0;
  }
  __syncthreads();
  for (int i = 0; i < 32; i++) {
    double acc = 0.0;
    if (t < i)
      for (int b = t; b < i; b++) acc += T[t][b] * P[b][i];
    __syncthreads();
    if (t < i) T[t][i] = -tau[i] * acc;
    if (t == 0) T[i][i] = tau[i];
    __syncthreads();
  }
  Vw[(size_t)k * 2048 + r * 32 + c] = V[r][c];
  Vw[(size_t)k * 2048 + (r + 32) * 32 + c] = V[r + 32][c];
  Tw[(size_t)k * 1024 + r * 32 + c] = T[r][c];
  Hm[(size_t)(k * 32 + r) * ldh + col0 + c] = X[r][c];
  Hm[(size_t)((k + 1) * 32 + r) * ldh + col0 + c] = 0.0;
  __syncthreads();
  upper_inverse_cols(X, Q);
  __syncthreads();
  Rdinv[(size_t)k * 1024 + r * 32 + c] = Q[r][c];
  __syncthreads();
  // right-hand side rows [32 k, 32 k + 64)
  X[r][c] = Gls[(size_t)(k * 32 + r) * 32 + c];
  X[r + 32][c] = Gls[(size_t)(k * 32 + 32 + r) * 32 + c];
  __syncthreads();
  apply_window_t(X, V, T, P, Q);
  Gls[(size_t)(k * 32 + r) * 32 + c] = X[r][c];
  Gls[(size_t)(k * 32 + 32 + r) * 32 + c] = X[r + 32][c];
  if (t < 32) {
    double s = 0.0;
    for (int q = 32; q < 64; q++) s += X[q][t] * X[q][t];
    res[t] = sqrt(s);
  }
}

int sd_hess_step(const HessState& H, int k, const double* W, const double* Rt, double* res) {
  LACE_REQUIRE(k >= 0 && k < H.m, DEV_ERR_ILLEGAL_VALUE, "hess step %d of %d", k, H.m);
  hess_step_kernel<<<1, 1024, 0, stream()>>>(k, H.ldh, H.Hm, H.Vw, H.Tw, H.Rdinv, H.Gls, W, Rt, res);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

__global__ __launch_bounds__(1024) void backsolve_kernel(int nblk, int ldh, const double* __restrict__ Hm,
                                                         const double* __restrict__ Rdinv,
                                                         const double* __restrict__ Gls,
                                                         double* __restrict__ Dneg) {
  __shared__ double Rh[32][33];
  const int t = threadIdx.x, r = t >> 5, c = t & 31;
  for (int jb = nblk - 1; jb >= 0; jb--) {
    double rhs = Gls[(size_t)(jb * 32 + r) * 32 + c];
    const double* hrow = Hm + (size_t)(jb * 32 + r) * ldh;
    for (int l = jb + 1; l < nblk; l++)
      for (int q = 0; q < 32; q++) rhs += hrow[l * 32 + q] * Dneg[(size_t)(l * 32 + q) * 32 + c];
    Rh[r][c] = rhs;
    __syncthreads();
    double y = 0.0;
    for (int q = r; q < 32; q++) y += Rdinv[(size_t)jb * 1024 + r * 32 + q] * Rh[q][c];
    Dneg[(size_t)(jb * 32 + r) * 32 + c] = -y;
    __threadfence();
    __syncthreads();
  }
}

int sd_backsolve(const HessState& H, int nblk, double* Dneg) {
  if (nblk <= 0) return DEV_SUCCESS;
  backsolve_kernel<<<1, 1024, 0, stream()>>>(nblk, H.ldh, H.Hm, H.Rdinv, H.Gls, Dneg);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

// ---------------------------------------------------------------- column norms
constexpr int kNormRows = 1024;

__global__ __launch_bounds__(256) void colnorm_partial_kernel(int n, const double* __restrict__ X,
                                                              double* __restrict__ partial) {
  __shared__ double red[8][32];
  const int t = threadIdx.x, c = t & 31, g = t >> 5;
  const int i0 = blockIdx.x * kNormRows, i1 = min(n, i0 + kNormRows);
  double s = 0.0;
  for (int i = i0 + g; i < i1; i += 8) {
    const double v = X[(size_t)i * 32 + c];
    s += v * v;
  }
  red[g][c] = s;
  __syncthreads();
  if (t < 32) {
    double tot = 0.0;
    for (int q = 0; q < 8; q++) tot += red[q][t];
    partial[(size_t)blockIdx.x * 32 + t] = tot;
  }
}

// out[e] = sum_s partial[s * count + e] (sqrt of it with SQRT): one
// workgroup per output, thread t summing slabs t, t + 256, ... in order, then
// the fixed-order block tree -- deterministic.
template <bool SQRT>
__global__ __launch_bounds__(256) void slab_reduce_kernel(int nslab, int count, const double* __restrict__ partial,
                                                          double* __restrict__ out) {
  __shared__ double sm[4];
  const int e = blockIdx.x;
  double s = 0.0;
  for (int q = threadIdx.x; q < nslab; q += 256) s += partial[(size_t)q * count + e];
  s = block_sum<256>(s, sm);
  if (threadIdx.x == 0) out[e] = SQRT ? sqrt(s) : s;
}

int sd_slab_reduce(int nslab, int count, const double* partial, double* out, bool take_sqrt) {
  if (count <= 0) return DEV_SUCCESS;
  if (take_sqrt) slab_reduce_kernel<true><<<count, 256, 0, stream()>>>(nslab, count, partial, out);
  else slab_reduce_kernel<false><<<count, 256, 0, stream()>>>(nslab, count, partial, out);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

size_t sd_colnorms_partial_elems(int n) { return (size_t)std::max(1, ceil_div(n, kNormRows)) * 32; }

int sd_colnorms(int n, const double* X, double* partial, double* out) {
  const int nslab = std::max(1, ceil_div(n, kNormRows));
  colnorm_partial_kernel<<<nslab, 256, 0, stream()>>>(n, X, partial);
  LACE_HIP(hipGetLastError());
  return sd_slab_reduce(nslab, 32, partial, out, true);
}

}  // namespace lace
