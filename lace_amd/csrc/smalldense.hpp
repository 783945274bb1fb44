// smalldense.hpp -- small FP64 dense kernels of the block Krylov path
// (row-major matrices; 32 x 32 panels, block Hessenberg windows).
#pragma once

#include "common.hpp"
#include "tsgemm.hpp"

namespace lace {

// C = alpha * op(A) op(B) + beta * C (M x N, inner K), row-major with
// leading dimensions; op(X) = X^T when the flag is set.  beta == 0 never
// reads C.  Fixed summation order (deterministic).
int sd_gemm(int M, int N, int K, double alpha, const double* A, int lda, bool ta, const double* B,
            int ldb, bool tb, double beta, double* C, int ldc);

// Cholesky G = R^T R of a 32 x 32 Gram matrix of an mrows x 32 panel, with
// R^{-1}.  A pivot below 1e-13 of its column's Gram diagonal (or not
// finite) restarts the factorization with the shift
// 11 (mrows*32 + 32*33) u trace(G) of shifted CholeskyQR3 (Fukaya et al.,
// SIAM J. Sci. Comput. 42 (2020)); status[0] = 0 (unshifted), 1 (shifted),
// 2 (failed: G is zero or not finite); status[1] = bit mask of the columns
// whose Gram diagonal is exactly zero.
int sd_chol32(const double* G, long mrows, double* R, double* Rinv, int* status);

// Householder reconstruction from a CholeskyQR panel (Ballard et al.,
// "Reconstructing Householder vectors from tall-skinny QR", JPDC 2015):
// Q = Qlast * Rl^{-1} has orthonormal columns, its top 32 rows start at
// row r0 of the block Qlast; the panel's R is Rl * Rprev.  With S =
// diag(-sign) chosen during the elimination, I - Q_1 S = L U (L unit lower),
// the block reflector I - Y Tb Y^T has Y = [0; L; -Q_2 S U^{-1}] and maps the
// panel to [S R; 0].  Outputs: Ltop (L, unit diagonal and zeros stored),
// Tb = U L^{-T}, Rt = S R, Dy = Rl^{-1} S U^{-1} (so Y_2 = -Qlast_2 Dy).
int sd_recon32(const double* Qlast, int r0, const double* Rl, const double* Rlinv, const double* Rprev,
               double* Ltop, double* Tb, double* Rt, double* Dy);

// Rows [r0, r0 + 32) of block Y <- Ltop.
int sd_write_top(const double* Ltop, int r0, double* Y);

// out (32 x 32 at leading dimension ldo) = alpha * in (ld 32); in ==
// nullptr gives alpha * I.
int sd_copy32(double alpha, const double* in, double* out, int ldo);

// Exactly zero panel columns: fill rows [r0, n) of the columns in mask with
// deterministic values in [-1, 1) before the QR, and zero those columns of
// the triangular factor after it.  With Q from the filled panel, Q^T of the
// original panel is the filled R with those columns zeroed, still upper
// triangular, so the factorization of the original panel stays exact.
int sd_fill_cols(int n, int r0, unsigned mask, double* W);
int sd_zero_cols32(unsigned mask, double* M);

// a[(b*32 + c)][r] = Y_b[(e0 + r)][c]: the rows e0.. of every block, i.e.
// Y_all^T E for the identity block column E starting at row e0.
int sd_gather_rows(const BlockList& Y, int e0, double* a);

// Block Hessenberg least-squares state.  Hm: (m+1)*32 x m*32 row-major
// (ld = m*32); column block k holds the triangularised Hessenberg column.
// Window j is the Householder QR (compact WY: Vw[j] 64 x 32, Tw[j]) of rows
// [32 j, 32 j + 64) of column block j after windows 0..j-1; Gls is the
// transformed right-hand side ((m+1)*32 x 32).
struct HessState {
  int m = 0, ldh = 0;
  double *Hm = nullptr, *Vw = nullptr, *Tw = nullptr, *Rdinv = nullptr, *Gls = nullptr;
};

// Column block k of the block Hessenberg matrix: top (k+1)*32 rows from
// rows [0, (k+1)*32) of block W, the subdiagonal block Rt (32 x 32).
// Applies windows 0..k-1, forms window k, applies it to Gls, writes res[c] =
// || Gls[(k+1)*32 .. (k+2)*32, c] || (the residual norm of right-hand side c)
// and the inverse of the new diagonal block.
int sd_hess_step(const HessState& H, int k, const double* W, const double* Rt, double* res);

// Dneg = -Y for the least-squares solution of the first nblk column blocks:
// R_H Y = Gls[0 : nblk*32] (block back substitution with the stored
// diagonal-block inverses).
int sd_backsolve(const HessState& H, int nblk, double* Dneg);

// out[e] = sum over slabs s of partial[s * count + e] (its square root with
// take_sqrt), e < count; fixed summation order.
int sd_slab_reduce(int nslab, int count, const double* partial, double* out, bool take_sqrt);

// out[c] = ||X[:, c]||_2 for an n x 32 block (fixed-order reduction);
// partial holds ceil(n / 4096) * 32 doubles.
int sd_colnorms(int n, const double* X, double* partial, double* out);
size_t sd_colnorms_partial_elems(int n);

}  // namespace lace
