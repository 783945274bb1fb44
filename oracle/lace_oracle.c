/*
 * lace_oracle.c -- CPU restatement of the reference's preconditioned-solve
 * path.  TEST INFRASTRUCTURE ONLY: it is the checker for the HIP library
 * and the CPU baseline timed beside it.  Nothing in the shipped library
 * links, loads or calls this file.
 *
 * Every function cites the reference file:line it restates.  Build:
 * oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  Contraction is off
 * so that products and sums round exactly like the GPU kernels, which are
 * compiled under `#pragma clang fp contract(off)` for the parity paths.
 *
 * Pinning (what each function is checked against, tests/test_oracle_*.py):
 *   orc_csr_split / orc_transpose : golden arrays of
 *       testing/test_matrix_io.cpp:474-733 (io_test.mtx) and
 *       testing/test_iLU_operations.cpp:172-330 (Trefethen_20).
 *   orc_spgemm                    : SpGEMM goldens of test_iLU_operations.cpp
 *       :186-232,261-318; dense numpy product.
 *   orc_parilu_sweep (Jacobi)     : sweep-1 closed form from the
 *       A-initialised guess (dense numpy), finite termination
 *       (k >= DAG depth sweeps == one Gauss-Seidel sweep, bitwise),
 *       orc_ilu0 to 1e-13, reference bounds Ares <= 10x exact.
 *   orc_ilu0                      : dense no-pivot LU (numpy) on no-fill
 *       patterns; invariant (LU)_ij = a_ij on pattern(A).
 *   orc_trsv_* / orc_parcsrtrsv   : scipy.linalg.solve_triangular; the
 *       Neumann closed form of Jacobi-k.
 *   orc_spmv                      : scipy.sparse product.
 *   orc_fgmres_householder / orc_fgmres_mgs : per-step residual history
 *       equals the minimal residual over the preconditioned Krylov space
 *       (dense least squares in numpy); true residual <= tol; A = I happy
 *       breakdown; diagonal A with d distinct values converges in d steps.
 *   orc_givens                    : closed form c*a + s*b = +-hypot(a,b),
 *       -s*a + c*b = 0.
 *   orc_levels                    : closed form lev = i+j (2-D) and
 *       i+j+k (3-D) on natural-order stencils; brute-force longest path.
 *   orc_housegen / orc_hqrd       : testing/matrices/magic7_Householder_{U,R}.mtx.
 *   orc_bilu_sweep_d/f (blocks)   : block size 1 equals orc_parilu_sweep to
 *       1e-14 (s*(1/u) vs s/u); sweep-1 dense closed form; the fixed point
 *       and one in-place sweep equal a dense block ILU(0) (IKJ, numpy) on
 *       testing/matrices/bcsr_test.mtx and on 7-point block grids.
 *   orc_binv (block inverse)      : numpy.linalg.inv, block sizes 1..8.
 *   orc_bcsr_spmv                 : dense numpy product.
 *   orc_btrsv                     : numpy.linalg.solve (exact); the Neumann
 *       closed form of block Jacobi-k.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* Pattern splits: control/magma_zmconverter.cpp:303-460.
 * upper = 0: CSRL -- entries with col < i, then the diagonal (if kept).
 * upper = 1: CSRU -- the diagonal (if kept), then entries with col > i.
 * Rows of A are assumed sorted by column, as the MatrixMarket reader
 * leaves them (control/magma_zmio.cpp:323-336), so CSRL has the diagonal
 * last and CSRU first.  diag_mode: 0 omit, 1 unit value, 2 A's value,
 * 3 zero value.  Returns nnz of the result; with Trow == NULL only counts. */
ORC_API int orc_csr_split(int n, const int *row, const int *col,
                          const double *val, int upper, int diag_mode,
                          int *Trow, int *Tcol, double *Tval) {
  int z = 0;
  for (int i = 0; i < n; i++) {
    if (Trow) Trow[i] = z;
    for (int k = row[i]; k < row[i + 1]; k++) {
      int j = col[k];
      int take = upper ? (j > i) : (j < i);
      double v = val ? val[k] : 0.0;
      if (j == i) {
        if (diag_mode == 0) continue;
        take = 1;
        v = diag_mode == 1 ? 1.0 : (diag_mode == 2 ? v : 0.0);
      }
      if (!take) continue;
      if (Trow) {
        Tcol[z] = j;
        if (Tval) Tval[z] = v;
      }
      z++;
    }
  }
  if (Trow) Trow[n] = z;
  return z;
}

/* CSR -> CSC by counting sort (control/magma_zmtranspose.cpp:85-235):
 * entries of each column appear in increasing row order. */
ORC_API void orc_transpose(int nrows, int ncols, const int *row,
                           const int *col, const double *val, int *trow,
                           int *tcol, double *tval) {
  memset(trow, 0, sizeof(int) * (ncols + 1));
  for (int i = 0; i < nrows; i++)
    for (int k = row[i]; k < row[i + 1]; k++) trow[col[k] + 1]++;
  for (int j = 0; j < ncols; j++) trow[j + 1] += trow[j];
  int *fill = (int *)malloc(sizeof(int) * (ncols + 1));
  memcpy(fill, trow, sizeof(int) * (ncols + 1));
  for (int i = 0; i < nrows; i++)
    for (int k = row[i]; k < row[i + 1]; k++) {
      int p = fill[col[k]]++;
      tcol[p] = i;
      if (val) tval[p] = val[k];
    }
  free(fill);
}

/* ------------------------------------------------------------------ */
/* Sparse x sparse C = A*B with sorted rows (blas/zspmm.cpp:47-167 uses MKL;
 * the reference tests fix the result arrays).  Two passes, dense marker. */
ORC_API int orc_spgemm(int m, int n, const int *arow, const int *acol,
                       const double *aval, const int *brow, const int *bcol,
                       const double *bval, int *crow, int *ccol, double *cval) {
  int *mark = (int *)malloc(sizeof(int) * n);
  double *acc = (double *)calloc(n, sizeof(double));
  int *list = (int *)malloc(sizeof(int) * n);
  for (int j = 0; j < n; j++) mark[j] = -1;
  int z = 0;
  for (int i = 0; i < m; i++) {
    int cnt = 0;
    if (crow) crow[i] = z;
    for (int p = arow[i]; p < arow[i + 1]; p++) {
      int k = acol[p];
      for (int q = brow[k]; q < brow[k + 1]; q++) {
        int j = bcol[q];
        if (mark[j] != i) {
          mark[j] = i;
          list[cnt++] = j;
          acc[j] = 0.0;
        }
        acc[j] += aval[p] * bval[q];
      }
    }
    /* insertion sort of the row's columns */
    for (int a = 1; a < cnt; a++) {
      int v = list[a], b = a - 1;
      while (b >= 0 && list[b] > v) {
        list[b + 1] = list[b];
        b--;
      }
      list[b + 1] = v;
    }
    if (ccol)
      for (int a = 0; a < cnt; a++) {
        ccol[z + a] = list[a];
        cval[z + a] = acc[list[a]];
      }
    z += cnt;
  }
  if (crow) crow[m] = z;
  free(mark);
  free(acc);
  free(list);
  return z;
}

/* ||A - LU||_F over LU's pattern (res) and over A's pattern (nonlinres):
 * control/norms.cpp:301-405.  L: CSRL with unit diagonal, U: CSR upper
 * (diagonal first).  Dense row accumulator instead of MKL SpGEMM. */
ORC_API void orc_ilures(int n, const int *arow, const int *acol,
                        const double *aval, const int *lrow, const int *lcol,
                        const double *lval, const int *urow, const int *ucol,
                        const double *uval, double *res, double *nonlinres) {
  double *acc = (double *)calloc(n, sizeof(double));
  int *mark = (int *)malloc(sizeof(int) * n);
  int *list = (int *)malloc(sizeof(int) * n);
  for (int j = 0; j < n; j++) mark[j] = -1;
  double r = 0.0, nl = 0.0;
  for (int i = 0; i < n; i++) {
    int cnt = 0;
    for (int p = lrow[i]; p < lrow[i + 1]; p++) {
      int k = lcol[p];
      for (int q = urow[k]; q < urow[k + 1]; q++) {
        int j = ucol[q];
        if (mark[j] != i) {
          mark[j] = i;
          list[cnt++] = j;
          acc[j] = 0.0;
        }
        acc[j] += lval[p] * uval[q];
      }
    }
    for (int p = arow[i]; p < arow[i + 1]; p++) {
      int j = acol[p];
      if (mark[j] == i) {
        double t = acc[j] - aval[p];
        acc[j] = t;
        nl += t * t;
      }
    }
    for (int a = 0; a < cnt; a++) r += acc[list[a]] * acc[list[a]];
  }
  *res = sqrt(r);
  *nonlinres = sqrt(nl);
  free(acc);
  free(mark);
  free(list);
}

/* ------------------------------------------------------------------ */
/* One ParILU(0) sweep: src/parilu_v0_3.cpp:70-102.
 * For each nnz (i,j) of A in CSR order: s = a_ij; merge L row i (CSRL,
 * cols ascending, unit diagonal last) with U column j (CSCU, rows
 * ascending, diagonal last); every match subtracts the rounded product;
 * the last product is added back (:90).  i > j: l_ij = s / u_jj
 * (u_jj = last entry of U column j, :93); else u_ij = s.
 * Reads come from (Lin, Uin), writes go to (Lout, Uout).  Jacobi mode:
 * distinct buffers, Lout/Uout pre-filled with Lin/Uin.  Gauss-Seidel
 * mode (the reference with one OpenMP thread): Lout == Lin, Uout == Uin.
 * Returns sum over updates of (old - new)^2 (before the 1/||A||_F scale,
 * :104). */
#define ORC_SWEEP_BODY(T)                                                    \
  double step = 0.0;                                                         \
  for (int i = 0; i < n; i++) {                                              \
    for (int k = arow[i]; k < arow[i + 1]; k++) {                            \
      int j = acol[k];                                                       \
      T s = (T)aval[k];                                                      \
      T sp = (T)0;                                                           \
      int il = lrow[i], iu = ucolptr[j];                                     \
      while (il < lrow[i + 1] && iu < ucolptr[j + 1]) {                      \
        sp = (T)0;                                                           \
        int jl = lcol[il], ju = urowind[iu];                                 \
        if (jl == ju) {                                                      \
          sp = Lin[il] * Uin[iu];                                            \
          s = s - sp;                                                        \
        }                                                                    \
        il = (jl <= ju) ? il + 1 : il;                                       \
        iu = (jl >= ju) ? iu + 1 : iu;                                       \
      }                                                                      \
      s = s + sp;                                                            \
      if (i > j) {                                                           \
        T t = s / Uin[ucolptr[j + 1] - 1];                                   \
        double d = (double)Lin[il - 1] - (double)t;                          \
        step += d * d;                                                       \
        Lout[il - 1] = t;                                                    \
      } else {                                                               \
        double d = (double)Uin[iu - 1] - (double)s;                          \
        step += d * d;                                                       \
        Uout[iu - 1] = s;                                                    \
      }                                                                      \
    }                                                                        \
  }                                                                          \
  return step;

ORC_API double orc_parilu_sweep_d(int n, const int *arow, const int *acol,
                                  const double *aval, const int *lrow,
                                  const int *lcol, const double *Lin,
                                  double *Lout, const int *ucolptr,
                                  const int *urowind, const double *Uin,
                                  double *Uout) {
  ORC_SWEEP_BODY(double)
}

ORC_API double orc_parilu_sweep_f(int n, const int *arow, const int *acol,
                                  const float *aval, const int *lrow,
                                  const int *lcol, const float *Lin,
                                  float *Lout, const int *ucolptr,
                                  const int *urowind, const float *Uin,
                                  float *Uout) {
  ORC_SWEEP_BODY(float)
}

/* The reference's OpenMP ParILU v0.3 loop as the CPU baseline:
 * src/parilu_v0_3_prescribedSweeps.cpp:62-107 -- asynchronous in-place
 * updates over nnz (rowidx from data_rowindex), `#pragma omp for` static
 * schedule, step reduction; runs up to max_sweeps, stopping early when
 * step <= 4 eps (:62).  Returns the number of sweeps run. */
ORC_API int orc_parilu_v03_omp(int n, int nnz, const int *arow,
                               const int *rowidx, const int *acol,
                               const double *aval, const int *lrow,
                               const int *lcol, double *L, const int *ucolptr,
                               const int *urowind, double *U, int max_sweeps,
                               double inv_anorm, double *final_step) {
  int iter = 0;
  double step = FLT_MAX;
  while (iter < max_sweeps && step > DBL_EPSILON * 4.0) {
    step = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : step)
    for (int k = 0; k < nnz; k++) {
      int i = rowidx[k], j = acol[k];
      double s = aval[k], sp = 0.0;
      int il = lrow[i], iu = ucolptr[j];
      while (il < lrow[i + 1] && iu < ucolptr[j + 1]) {
        sp = 0.0;
        int jl = lcol[il], ju = urowind[iu];
        sp = (jl == ju) ? L[il] * U[iu] : sp;
        s = (jl == ju) ? s - sp : s;
        il = (jl <= ju) ? il + 1 : il;
        iu = (jl >= ju) ? iu + 1 : iu;
      }
      s += sp;
      if (i > j) {
        double t = s / U[ucolptr[j + 1] - 1];
        step += (L[il - 1] - t) * (L[il - 1] - t);
        L[il - 1] = t;
      } else {
        step += (U[iu - 1] - s) * (U[iu - 1] - s);
        U[iu - 1] = s;
      }
    }
    step *= inv_anorm;
    iter++;
  }
  (void)arow;
  if (final_step) *final_step = step;
  return iter;
}

/* Textbook ILU(0), IKJ variant (Saad, Iterative Methods, Alg. 10.4), in
 * place on CSR with sorted rows -- the exact factorization the reference
 * obtains from MKL dcsrilu0 (blas/zcsrilu0.cpp:15-75).  The result holds
 * strict-L and U in A's pattern (data_zmlumerge layout).  Returns the row
 * of the first pivot with |u_ii| < 1e-16 or -1. */
ORC_API int orc_ilu0(int n, const int *row, const int *col, double *val) {
  int *diag = (int *)malloc(sizeof(int) * n);
  int *pos = (int *)malloc(sizeof(int) * n);
  int bad = -1;
  for (int j = 0; j < n; j++) pos[j] = -1;
  for (int i = 0; i < n; i++) {
    diag[i] = -1;
    for (int k = row[i]; k < row[i + 1]; k++)
      if (col[k] == i) diag[i] = k;
  }
  for (int i = 0; i < n; i++) {
    for (int k = row[i]; k < row[i + 1]; k++) pos[col[k]] = k;
    for (int p = row[i]; p < row[i + 1] && col[p] < i; p++) {
      int k = col[p];
      val[p] = val[p] / val[diag[k]];
      for (int q = diag[k] + 1; q < row[k + 1]; q++) {
        int j = col[q];
        if (pos[j] >= 0) val[pos[j]] = val[pos[j]] - val[p] * val[q];
      }
    }
    if (diag[i] < 0 || fabs(val[diag[i]]) < 1e-16) {
      if (bad < 0) bad = i;
    }
    for (int k = row[i]; k < row[i + 1]; k++) pos[col[k]] = -1;
  }
  free(diag);
  free(pos);
  return bad;
}

/* ------------------------------------------------------------------ */
/* Exact triangular solves in natural order.  Lower: CSRL, diagonal last
 * (src/trisolve.cpp:283-288); upper: CSRU, diagonal first (:326-331),
 * rows from n-1 down to 0.  The inner sum accumulates left to right, as
 * the reference loop does. */
#define ORC_TRSV_L(T)                                                        \
  for (int i = 0; i < n; i++) {                                              \
    T t = (T)0;                                                              \
    for (int k = row[i]; k < row[i + 1] - 1; k++) t = t + val[k] * y[col[k]]; \
    y[i] = (r[i] - t) / val[row[i + 1] - 1];                                 \
  }
#define ORC_TRSV_U(T)                                                        \
  for (int i = n - 1; i >= 0; i--) {                                         \
    T t = (T)0;                                                              \
    for (int k = row[i] + 1; k < row[i + 1]; k++) t = t + val[k] * y[col[k]]; \
    y[i] = (r[i] - t) / val[row[i]];                                         \
  }

ORC_API void orc_trsv_lower_d(int n, const int *row, const int *col,
                              const double *val, const double *r, double *y) {
  ORC_TRSV_L(double)
}
ORC_API void orc_trsv_upper_d(int n, const int *row, const int *col,
                              const double *val, const double *r, double *y) {
  ORC_TRSV_U(double)
}
ORC_API void orc_trsv_lower_f(int n, const int *row, const int *col,
                              const float *val, const float *r, float *y) {
  ORC_TRSV_L(float)
}
ORC_API void orc_trsv_upper_f(int n, const int *row, const int *col,
                              const float *val, const float *r, float *y) {
  ORC_TRSV_U(float)
}

/* Synchronous Jacobi sweeps for a triangular system from y0 = 0:
 * y_{m+1,i} = (r_i - sum_{k != i} T_ik y_{m,k}) / T_ii, the update formula
 * of src/trisolve.cpp:281-290 with every read taken from sweep m.
 * upper selects the diagonal position (CSRU first / CSRL last). */
#define ORC_JAC(T)                                                           \
  T *prev = (T *)calloc(n, sizeof(T));                                       \
  for (int m = 0; m < sweeps; m++) {                                         \
    for (int i = 0; i < n; i++) {                                            \
      int kd = upper ? row[i] : row[i + 1] - 1;                              \
      int k0 = upper ? row[i] + 1 : row[i];                                  \
      int k1 = upper ? row[i + 1] : row[i + 1] - 1;                          \
      T t = (T)0;                                                            \
      for (int k = k0; k < k1; k++) t = t + val[k] * prev[col[k]];           \
      y[i] = (r[i] - t) / val[kd];                                           \
    }                                                                        \
    memcpy(prev, y, sizeof(T) * n);                                          \
  }                                                                          \
  if (sweeps <= 0) memset(y, 0, sizeof(T) * n);                              \
  free(prev);

ORC_API void orc_trsv_jacobi_d(int n, const int *row, const int *col,
                               const double *val, int upper, int sweeps,
                               const double *r, double *y) {
  ORC_JAC(double)
}
ORC_API void orc_trsv_jacobi_f(int n, const int *row, const int *col,
                               const float *val, int upper, int sweeps,
                               const float *r, float *y) {
  ORC_JAC(float)
}

/* ParCSRTRSV: src/trisolve.cpp:256-351.  In-place sweeps (Gauss-Seidel in
 * the given order with one thread; the reference's OpenMP static chunks
 * with more) until sum (y_old - y_new)^2 <= tol or step is not finite.
 * y holds the initial guess on entry.  Returns the sweep count. */
ORC_API int orc_parcsrtrsv(int n, const int *row, const int *col,
                           const double *val, int upper, const double *r,
                           double *y, double tol) {
  int iter = 0;
  double step = 1.e8;
  while (step > tol && isfinite(step)) {
    step = 0.0;
    if (!upper) {
#pragma omp parallel for schedule(static) reduction(+ : step)
      for (int i = 0; i < n; i++) {
        double t = 0.0;
        for (int k = row[i]; k < row[i + 1] - 1; k++) t += val[k] * y[col[k]];
        t = (r[i] - t) / val[row[i + 1] - 1];
        step += (y[i] - t) * (y[i] - t);
        y[i] = t;
      }
    } else {
#pragma omp parallel for schedule(static) reduction(+ : step)
      for (int ii = 0; ii < n; ii++) {
        int i = n - 1 - ii;
        double t = 0.0;
        for (int k = row[i] + 1; k < row[i + 1]; k++) t += val[k] * y[col[k]];
        t = (r[i] - t) / val[row[i]];
        step += (y[i] - t) * (y[i] - t);
        y[i] = t;
      }
    }
    iter++;
  }
  return iter;
}

/* Level sets of a triangular pattern (absent from the reference; defined
 * here): lower -- lev(i) = 1 + max lev(k), k < i in row i; upper --
 * lev(i) = 1 + max lev(k), k > i in row i; 0 when none.  Writes lev[]
 * and returns the number of levels. */
ORC_API int orc_levels(int n, const int *row, const int *col, int upper,
                       int *lev) {
  int nl = 0;
  for (int ii = 0; ii < n; ii++) {
    int i = upper ? n - 1 - ii : ii;
    int l = 0;
    for (int k = row[i]; k < row[i + 1]; k++) {
      int j = col[k];
      if ((upper && j > i) || (!upper && j < i))
        if (lev[j] + 1 > l) l = lev[j] + 1;
    }
    lev[i] = l;
    if (l + 1 > nl) nl = l + 1;
  }
  return nl;
}

/* y = alpha*A*x + beta*y (blas/zspmv.cpp:48-78 via mkl_dcsrmv). */
ORC_API void orc_spmv(int n, const int *row, const int *col, const double *val,
                      double alpha, const double *x, double beta, double *y) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    double t = 0.0;
    for (int k = row[i]; k < row[i + 1]; k++) t += val[k] * x[col[k]];
    y[i] = alpha * t + (beta == 0.0 ? 0.0 : beta * y[i]);
  }
}

/* ------------------------------------------------------------------ */
/* src/givens.cpp:46-67 */
ORC_API void orc_givens(double a, double b, double tol, double *c, double *s) {
  if (fabs(b) <= tol) {
    *c = 1.0;
    *s = 0.0;
  } else if (fabs(b) > fabs(a)) {
    double t = a / b;
    *s = 1.0 / sqrt(1.0 + t * t);
    *c = t * (*s);
  } else {
    double t = b / a;
    *c = 1.0 / sqrt(1.0 + t * t);
    *s = t * (*c);
  }
}

static double orc_nrm2(int n, const double *x) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += x[i] * x[i];
  return sqrt(s);
}
static double orc_sgn(double v) { return (double)(v >= 0.0) - (double)(v < 0.0); }

/* Preconditioner of the FGMRES restatements (src/fgmres.cpp:216-246):
 * choice 0 -- exact L (unit) then U solves; 1 -- ParCSRTRSV with tol 1.0
 * from y = 0; 2 -- identity; 3 -- jac synchronous Jacobi sweeps each. */
typedef struct {
  int n, choice, jac;
  const int *lrow, *lcol, *urow, *ucol;
  const double *lval, *uval;
  double *tmp;
  long trsv_sweeps;
} orc_prec;

static void orc_prec_apply(orc_prec *P, const double *q, double *z) {
  int n = P->n;
  if (P->choice == 2) {
    memcpy(z, q, sizeof(double) * n);
  } else if (P->choice == 0) {
    orc_trsv_lower_d(n, P->lrow, P->lcol, P->lval, q, P->tmp);
    orc_trsv_upper_d(n, P->urow, P->ucol, P->uval, P->tmp, z);
  } else if (P->choice == 1) {
    memset(P->tmp, 0, sizeof(double) * n);
    memset(z, 0, sizeof(double) * n);
    P->trsv_sweeps += orc_parcsrtrsv(n, P->lrow, P->lcol, P->lval, 0, q, P->tmp, 1.0);
    P->trsv_sweeps += orc_parcsrtrsv(n, P->urow, P->ucol, P->uval, 1, P->tmp, z, 1.0);
  } else {
    orc_trsv_jacobi_d(n, P->lrow, P->lcol, P->lval, 0, P->jac, q, P->tmp);
    orc_trsv_jacobi_d(n, P->urow, P->ucol, P->uval, 1, P->jac, P->tmp, z);
  }
}

/* Restarted Householder FGMRES: src/fgmres_householder_restart.cpp:63-788,
 * sequential.  W (n x (m+1), column-major) holds, as in the reference, the
 * Hessenberg column in rows 0..j of column j+1 and the reflector u_{j+1}
 * (unit 2-norm... scaled so ||u|| = 1) in rows j+1..n-1.  The reflector
 * norm runs over rows j+1..n-1; the reference's length n-j reads one
 * element past the column (:477), which is the zero first entry of the
 * next column except in the last search direction.
 * hist (may be NULL) receives |g_{j+1}| per search direction (max
 * hist_len).  Returns info 0 when |g| < rtol was reached. */
ORC_API int orc_fgmres_householder(
    int n, const int *arow, const int *acol, const double *aval,
    const double *b, double *x, int m, int restart_max, double rtol_in,
    int tol_type, int choice, int jac, const int *lrow, const int *lcol,
    const double *lval, const int *urow, const int *ucol, const double *uval,
    int *searches_out, int *restarts_out, double *hist, int hist_len,
    double *final_res, int nthreads) {
  const int nt = nthreads > 1 ? nthreads : 1;
  double *W = (double *)calloc((size_t)n * (m + 1), sizeof(double));
  double *Z = (double *)calloc((size_t)n * m, sizeof(double));
  double *r = (double *)calloc(n, sizeof(double));
  double *q = (double *)calloc(n, sizeof(double));
  double *g = (double *)calloc(m + 1, sizeof(double));
  double *cs = (double *)calloc(m, sizeof(double));
  double *sn = (double *)calloc(m, sizeof(double));
  double *y = (double *)calloc(m + 1, sizeof(double));
  orc_prec P = {n, choice, jac, lrow, lcol, urow, ucol, lval, uval,
                (double *)calloc(n, sizeof(double)), 0};
  const double eps = nextafter(0.0, 1.0);
  double rtol = rtol_in;
  int info = -1, total = 0, nh = 0, restarts = 0;
  double fres = 0.0;
#define Wc(i, j) W[(size_t)(j) * n + (i)]
#define Zc(i, j) Z[(size_t)(j) * n + (i)]
  for (int restart = 0; restart < restart_max; ++restart) {
    memset(W, 0, sizeof(double) * (size_t)n * (m + 1));
    memset(Z, 0, sizeof(double) * (size_t)n * m);
    memset(g, 0, sizeof(double) * (m + 1));
    memset(cs, 0, sizeof(double) * m);
    memset(sn, 0, sizeof(double) * m);
    /* r = b - A x   (:209-221) */
    for (int i = 0; i < n; i++) {
      double t = 0.0;
      for (int k = arow[i]; k < arow[i + 1]; k++) t = t - aval[k] * x[acol[k]];
      r[i] = t + b[i];
    }
    double rnorm2 = orc_nrm2(n, r);
    if (restart == 0 && tol_type == 1) rtol = rtol * rnorm2;
    if (rnorm2 < rtol) {
      info = 0;
      fres = rnorm2;
      break;
    }
    /* first reflector (:239-308) */
    for (int i = 0; i < n; i++) Wc(i, 0) = r[i];
    double dd = orc_sgn(Wc(0, 0)) * rnorm2;
    Wc(0, 0) = Wc(0, 0) + dd;
    double k1 = 1.0 / orc_nrm2(n, W);
    for (int i = 0; i < n; i++) Wc(i, 0) = Wc(i, 0) * k1;
    for (int i = 0; i < n; i++) q[i] = -r[i] / dd;
    g[0] = -dd;
    int done = 0;
    for (int s = 0; s < m; ++s) {
      int s1 = s + 1;
      orc_prec_apply(&P, q, &Zc(0, s));                       /* :320-349 */
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
      for (int i = 0; i < n; i++) {                           /* :356-367 */
        double t = Wc(i, s1);
        for (int k = arow[i]; k < arow[i + 1]; k++)
          t = t + aval[k] * Zc(acol[k], s);
        Wc(i, s1) = t;
      }
      for (int j = 0; j <= s; ++j) {                          /* :375-469 */
        double sum = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : sum) num_threads(nt) if (nt > 1)
        for (int i = j; i < n; i++) sum += Wc(i, j) * Wc(i, s1);
        sum = 2.0 * sum;
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
        for (int i = j; i < n; i++) Wc(i, s1) = Wc(i, s1) - sum * Wc(i, j);
      }
      if (s < n) {                                            /* :476-517 */
        double snrm2 = s1 < n ? orc_nrm2(n - s1, &Wc(s1, s1)) : 0.0;
        dd = orc_sgn(Wc(s1 < n ? s1 : n - 1, s1)) * snrm2;
        if (s1 < n) {
          Wc(s1, s1) = Wc(s1, s1) + dd;
          double nv = orc_nrm2(n - s1, &Wc(s1, s1));
          double sc = 1.0 / nv;
          for (int i = s1; i < n; i++) Wc(i, s1) = Wc(i, s1) * sc;
        }
        memset(q, 0, sizeof(double) * n);                     /* :519-616 */
        if (s1 < n) q[s1] = 1.0;
        for (int j = (s1 < n ? s1 : s); j >= 0; --j) {
          double sum = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : sum) num_threads(nt) if (nt > 1)
          for (int i = j; i < n; i++) sum += Wc(i, j) * q[i];
          sum = 2.0 * sum;
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
          for (int i = j; i < n; i++) q[i] = q[i] - sum * Wc(i, j);
        }
      }
      for (int j = 0; j < s; ++j) {                           /* :658-679 */
        double t0 = Wc(j, s1), t1 = Wc(j + 1, s1);
        Wc(j + 1, s1) = -sn[j] * t0 + cs[j] * t1;
        Wc(j, s1) = cs[j] * t0 + sn[j] * t1;
      }
      if (s < n) {                                            /* :686-701 */
        orc_givens(Wc(s, s1), -dd, eps, &cs[s], &sn[s]);
        Wc(s, s1) = cs[s] * Wc(s, s1) - sn[s] * dd;
        g[s1] = -sn[s] * g[s];
        g[s] = cs[s] * g[s];
      }
      total++;
      if (hist && nh < hist_len) hist[nh++] = fabs(g[s1]);
      if (fabs(g[s1]) < rtol || s == m - 1) {                 /* :714-757 */
        for (int i = 0; i <= s; ++i) y[i] = g[i] / Wc(i, i + 1);
        for (int j = s; j > 0; --j)
          for (int i = j - 1; i >= 0; --i)
            y[i] = y[i] - Wc(i, j + 1) * y[j] / Wc(i, i + 1);
        for (int j = 0; j <= s; j++)
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
          for (int i = 0; i < n; i++) x[i] = x[i] + Zc(i, j) * y[j];
        fres = fabs(g[s1]);
        restarts = restart;
        if (fabs(g[s1]) < rtol) {
          done = 1;
          info = 0;
        }
        break;
      }
    }
    if (done) break;
  }
#undef Wc
#undef Zc
  if (searches_out) *searches_out = total;
  if (restarts_out) *restarts_out = restarts;
  if (final_res) *final_res = fres;
  free(W); free(Z); free(r); free(q); free(g); free(cs); free(sn); free(y);
  free(P.tmp);
  return info;
}

/* Restarted MGS FGMRES: src/fgmres.cpp:64-517 (one cycle) repeated as in
 * src/fgmres_restart.cpp.  reorth: 0 Brown/Hindmarsh test (:314-315),
 * 1 never, 2 always.  Same outputs as orc_fgmres_householder. */
ORC_API int orc_fgmres_mgs(
    int n, const int *arow, const int *acol, const double *aval,
    const double *b, double *x, int m, int restart_max, double rtol_in,
    int tol_type, int reorth, int choice, int jac, const int *lrow,
    const int *lcol, const double *lval, const int *urow, const int *ucol,
    const double *uval, int *searches_out, int *restarts_out, double *hist,
    int hist_len, double *final_res) {
  double *V = (double *)calloc((size_t)n * (m + 1), sizeof(double));
  double *Z = (double *)calloc((size_t)n * m, sizeof(double));
  double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
  double *r = (double *)calloc(n, sizeof(double));
  double *u = (double *)calloc(n, sizeof(double));
  double *g = (double *)calloc(m + 1, sizeof(double));
  double *cs = (double *)calloc(m, sizeof(double));
  double *sn = (double *)calloc(m, sizeof(double));
  double *y = (double *)calloc(m + 1, sizeof(double));
  orc_prec P = {n, choice, jac, lrow, lcol, urow, ucol, lval, uval,
                (double *)calloc(n, sizeof(double)), 0};
  const double eps = nextafter(0.0, 1.0);
  double rtol = rtol_in, fres = 0.0;
  int info = -1, total = 0, nh = 0, restarts = 0;
#define Vc(i, j) V[(size_t)(j) * n + (i)]
#define Zc(i, j) Z[(size_t)(j) * n + (i)]
#define Hc(i, j) H[(size_t)(j) * (m + 1) + (i)]
  for (int restart = 0; restart < restart_max; ++restart) {
    memset(H, 0, sizeof(double) * (size_t)(m + 1) * m);
    memset(g, 0, sizeof(double) * (m + 1));
    for (int i = 0; i < n; i++) {                             /* :175-177 */
      double t = 0.0;
      for (int k = arow[i]; k < arow[i + 1]; k++) t = t + (-1.0) * aval[k] * x[acol[k]];
      r[i] = t + b[i];
    }
    double rnorm2 = orc_nrm2(n, r);
    if (restart == 0 && tol_type == 1) rtol = rtol * rnorm2;
    if (rnorm2 < rtol) {
      info = 0;
      fres = rnorm2;
      break;
    }
    for (int i = 0; i < n; i++) Vc(i, 0) = r[i] / rnorm2;    /* :193-196 */
    g[0] = rnorm2;
    int done = 0;
    for (int s = 0; s < m; s++) {
      int s1 = s + 1;
      orc_prec_apply(&P, &Vc(0, s), &Zc(0, s));
      for (int i = 0; i < n; i++) {                           /* :259-268 */
        double t = 0.0;
        for (int k = arow[i]; k < arow[i + 1]; k++) t = t + aval[k] * Zc(acol[k], s);
        u[i] = t;
      }
      double normav = orc_nrm2(n, u);
      for (int j = 0; j <= s; j++) {                          /* :281-306 */
        double t = 0.0;
        for (int i = 0; i < n; i++) t = t + Vc(i, j) * u[i];
        Hc(j, s) = t;
        for (int i = 0; i < n; i++) u[i] = u[i] - Hc(j, s) * Vc(i, j);
      }
      Hc(s1, s) = orc_nrm2(n, u);
      double normav2 = Hc(s1, s);
      double hr = (normav + 0.001 * normav2) - normav;        /* :314-338 */
      if ((reorth == 0 && hr <= eps) || reorth == 2) {
        for (int j = 0; j <= s; j++) {
          hr = 0.0;
          for (int i = 0; i < n; i++) hr = hr + Vc(i, j) * u[i];
          Hc(j, s) = Hc(j, s) + hr;
          for (int i = 0; i < n; i++) u[i] = u[i] - hr * Vc(i, j);
        }
        Hc(s1, s) = orc_nrm2(n, u);
      }
      if (fabs(Hc(s1, s)) > DBL_EPSILON)                      /* :341-356 */
        for (int i = 0; i < n; i++) Vc(i, s1) = u[i] / Hc(s1, s);
      for (int j = 0; j < s; j++) {                           /* :385-393 */
        double delta = Hc(j, s);
        Hc(j, s) = cs[j] * delta - sn[j] * Hc(j + 1, s);
        Hc(j + 1, s) = sn[j] * delta + cs[j] * Hc(j + 1, s);
      }
      double gamma = sqrt(Hc(s, s) * Hc(s, s) + Hc(s1, s) * Hc(s1, s));
      if (gamma > 0.0) {                                      /* :399-411 */
        cs[s] = Hc(s, s) / gamma;
        sn[s] = -Hc(s1, s) / gamma;
        Hc(s, s) = cs[s] * Hc(s, s) - sn[s] * Hc(s1, s);
        Hc(s1, s) = 0.0;
        double delta = g[s];
        g[s] = cs[s] * delta - sn[s] * g[s1];
        g[s1] = sn[s] * delta + cs[s] * g[s1];
      }
      total++;
      if (hist && nh < hist_len) hist[nh++] = fabs(g[s1]);
      if (fabs(g[s1]) < rtol || s == m - 1 || !isfinite(g[s1])) { /* :440-481 */
        for (int i = 0; i <= s; i++) y[i] = g[i] / Hc(i, i);
        for (int j = s; j > 0; j--)
          for (int i = j - 1; i > -1; i--)
            y[i] = y[i] - Hc(i, j) * y[j] / Hc(i, i);
        for (int i = 0; i < n; i++) {
          double t = 0.0;
          for (int j = 0; j <= s; j++) t = t + Zc(i, j) * y[j];
          x[i] = x[i] + t;
        }
        fres = fabs(g[s1]);
        restarts = restart;
        if (fabs(g[s1]) < rtol) {
          done = 1;
          info = 0;
        }
        break;
      }
    }
    if (done) break;
  }
#undef Vc
#undef Zc
#undef Hc
  if (searches_out) *searches_out = total;
  if (restarts_out) *restarts_out = restarts;
  if (final_res) *final_res = fres;
  free(V); free(Z); free(H); free(r); free(u); free(g); free(cs); free(sn);
  free(y); free(P.tmp);
  return info;
}

/* ------------------------------------------------------------------ */
/* Stewart housegen (src/householder.cpp:49-83): u with ||u|| = sqrt(2),
 * H = I - u u^T maps x to nu e_1. */
ORC_API void orc_housegen(int n, const double *x, double *u, double *nu) {
  const double eps = nextafter(0.0, 1.0);
  for (int i = 0; i < n; i++) u[i] = x[i];
  *nu = orc_nrm2(n, x);
  if (*nu < eps) {
    u[0] = sqrt(2.0);
    return;
  }
  for (int i = 0; i < n; i++) u[i] = x[i] / (*nu);
  if (u[0] >= 0) {
    u[0] = u[0] + 1.0;
    *nu = -(*nu);
  } else {
    u[0] = u[0] - 1.0;
  }
  double t = 1.0 / sqrt(fabs(u[0]));
  for (int i = 0; i < n; i++) u[i] = u[i] * t;
}

/* Stewart hqrd (src/householder.cpp:120-160): X (m x n, column-major) is
 * overwritten; U (m x n) receives the generators, R (n x n) the triangle. */
ORC_API void orc_hqrd(int m, int n, double *X, double *U, double *R) {
  int p = m < n ? m : n;
  memset(U, 0, sizeof(double) * (size_t)m * n);
  memset(R, 0, sizeof(double) * (size_t)n * n);
  for (int k = 0; k < p; k++) {
    double nu;
    orc_housegen(m - k, &X[(size_t)k * m + k], &U[(size_t)k * m + k], &nu);
    R[(size_t)k * n + k] = nu;
    for (int j = k + 1; j < n; j++) {
      double v = 0.0;
      for (int i = k; i < m; i++) v += U[(size_t)k * m + i] * X[(size_t)j * m + i];
      for (int i = k; i < m; i++) X[(size_t)j * m + i] -= v * U[(size_t)k * m + i];
      R[(size_t)j * n + k] = X[(size_t)j * m + k];
    }
  }
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11; Random123): the counter-based RNG
 * the synthetic generators use.  Checked against the Random123 known-answer
 * vectors in the tests. */
ORC_API void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                               uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Thread count of the OpenMP loops above (n <= 0 restores the default). */
ORC_API void orc_set_threads(int n) {
#ifdef _OPENMP
  static int dflt = 0;
  if (dflt == 0) dflt = omp_get_max_threads();
  omp_set_num_threads(n > 0 ? n : dflt);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* Block (BCSR) routines.  Blocks are bs x bs, row-major, ldblock = bs*bs
 * (the layout the reference's block ParILU multiplies with row-major
 * dgemm, src/parilu_v0_3_bcsr.cpp:167-171).  Block inverse: Gauss-Jordan
 * without pivoting on [M | I] in a fixed order (the reference iterates
 * dense ParLU + triangular sweeps to 1e-15, src/inverse.cpp:17-160; the
 * fixed elimination is the exact inverse that iteration approximates). */
#define ORC_MAXBS 8

#define ORC_BLOCK_FUNCS(T, SUF)                                               \
  static void orc_binv_##SUF(int bs, const T *M, T *Inv) {                    \
    T a[ORC_MAXBS * ORC_MAXBS];                                               \
    for (int e = 0; e < bs * bs; e++) {                                       \
      a[e] = M[e];                                                            \
      Inv[e] = (T)0;                                                          \
    }                                                                         \
    for (int r = 0; r < bs; r++) Inv[r * bs + r] = (T)1;                      \
    for (int p = 0; p < bs; p++) {                                            \
      const T inv = (T)1 / a[p * bs + p];                                     \
      for (int c = 0; c < bs; c++) {                                          \
        a[p * bs + c] = a[p * bs + c] * inv;                                  \
        Inv[p * bs + c] = Inv[p * bs + c] * inv;                              \
      }                                                                       \
      for (int r = 0; r < bs; r++) {                                          \
        if (r == p) continue;                                                 \
        const T f = a[r * bs + p];                                            \
        for (int c = 0; c < bs; c++) {                                        \
          a[r * bs + c] = a[r * bs + c] - f * a[p * bs + c];                  \
          Inv[r * bs + c] = Inv[r * bs + c] - f * Inv[p * bs + c];            \
        }                                                                     \
      }                                                                       \
    }                                                                         \
  }                                                                           \
  /* C = A * B, each entry summed over m in ascending order from zero */      \
  static void orc_bmul_##SUF(int bs, const T *A, const T *B, T *Cm) {         \
    for (int r = 0; r < bs; r++)                                              \
      for (int c = 0; c < bs; c++) {                                          \
        T acc = (T)0;                                                         \
        for (int m = 0; m < bs; m++) acc = acc + A[r * bs + m] * B[m * bs + c]; \
        Cm[r * bs + c] = acc;                                                 \
      }                                                                       \
  }

ORC_BLOCK_FUNCS(double, d)
ORC_BLOCK_FUNCS(float, f)

/* One synchronous block ParILU sweep in the reference's form
 * (src/parilu_v0_3_bcsr.cpp:118-250): for each block (i,j) of A in BCSR
 * order, S = A_ij; merge L block row i (BCSRL, identity diagonal block
 * last) with U block column j (BCSCU, diagonal block last); every match
 * subtracts SP = L_ik U_kj; the last SP is added back; i > j: L_ij =
 * S inv(U_jj), else U_ij = S.  Reads (Lin, Uin), writes (Lout, Uout).
 * Returns sum of squared changes (no ||A|| scaling, :61). */
#define ORC_BSWEEP_BODY(T, SUF)                                               \
  const int lb = bs * bs;                                                     \
  T s[ORC_MAXBS * ORC_MAXBS], sp[ORC_MAXBS * ORC_MAXBS];                      \
  T inv[ORC_MAXBS * ORC_MAXBS], t[ORC_MAXBS * ORC_MAXBS];                     \
  double step = 0.0;                                                          \
  for (int i = 0; i < nbr; i++) {                                             \
    for (int k = arow[i]; k < arow[i + 1]; k++) {                             \
      const int j = acol[k];                                                  \
      for (int e = 0; e < lb; e++) s[e] = (T)aval[(size_t)k * lb + e];        \
      for (int e = 0; e < lb; e++) sp[e] = (T)0;                              \
      int il = lrow[i], iu = ucolptr[j];                                      \
      while (il < lrow[i + 1] && iu < ucolptr[j + 1]) {                       \
        for (int e = 0; e < lb; e++) sp[e] = (T)0;                            \
        const int jl = lcol[il], ju = urowind[iu];                            \
        if (jl == ju) {                                                       \
          orc_bmul_##SUF(bs, &Lin[(size_t)il * lb], &Uin[(size_t)iu * lb], sp); \
          for (int e = 0; e < lb; e++) s[e] = s[e] - sp[e];                   \
        }                                                                     \
        il = (jl <= ju) ? il + 1 : il;                                        \
        iu = (jl >= ju) ? iu + 1 : iu;                                        \
      }                                                                       \
      for (int e = 0; e < lb; e++) s[e] = s[e] + sp[e];                       \
      if (i > j) {                                                            \
        orc_binv_##SUF(bs, &Uin[(size_t)(ucolptr[j + 1] - 1) * lb], inv);     \
        orc_bmul_##SUF(bs, s, inv, t);                                        \
        for (int e = 0; e < lb; e++) {                                        \
          const double d = (double)Lin[(size_t)(il - 1) * lb + e] - (double)t[e]; \
          step += d * d;                                                      \
          Lout[(size_t)(il - 1) * lb + e] = t[e];                             \
        }                                                                     \
      } else {                                                                \
        for (int e = 0; e < lb; e++) {                                        \
          const double d = (double)Uin[(size_t)(iu - 1) * lb + e] - (double)s[e]; \
          step += d * d;                                                      \
          Uout[(size_t)(iu - 1) * lb + e] = s[e];                             \
        }                                                                     \
      }                                                                       \
    }                                                                         \
  }                                                                           \
  return step;

ORC_API double orc_bilu_sweep_d(int nbr, int bs, const int *arow, const int *acol,
                                const double *aval, const int *lrow, const int *lcol,
                                const double *Lin, double *Lout, const int *ucolptr,
                                const int *urowind, const double *Uin, double *Uout) {
  ORC_BSWEEP_BODY(double, d)
}

ORC_API double orc_bilu_sweep_f(int nbr, int bs, const int *arow, const int *acol,
                                const double *aval, const int *lrow, const int *lcol,
                                const float *Lin, float *Lout, const int *ucolptr,
                                const int *urowind, const float *Uin, float *Uout) {
  ORC_BSWEEP_BODY(float, f)
}

ORC_API void orc_binv_double(int bs, const double *M, double *Inv) { orc_binv_d(bs, M, Inv); }

/* y = A x for BCSR A (bs x bs row-major blocks); y_{I,a} sums blocks in
 * order, within a block over c ascending. */
ORC_API void orc_bcsr_spmv(int nbr, int bs, const int *row, const int *col,
                           const double *val, const double *x, double *y) {
  const int lb = bs * bs;
  for (int I = 0; I < nbr; I++)
    for (int a = 0; a < bs; a++) {
      double t = 0.0;
      for (int k = row[I]; k < row[I + 1]; k++)
        for (int c = 0; c < bs; c++) t += val[(size_t)k * lb + a * bs + c] * x[(size_t)col[k] * bs + c];
      y[(size_t)I * bs + a] = t;
    }
}

/* Block triangular solves on L (BCSRL: diagonal block last in each block
 * row) or U in block-CSR form with the diagonal block first.  Exact:
 * t = r_I - sum_K T_IK y_K (blocks in order, entries c ascending), then
 * y_I = inv(T_II) t (Gauss-Jordan inverse above); jac > 0: that many
 * synchronous Jacobi sweeps from y = 0 instead. */
ORC_API void orc_btrsv(int nbr, int bs, const int *row, const int *col, const float *val,
                       int upper, int jac, const float *r, float *y) {
  const int lb = bs * bs;
  float inv[ORC_MAXBS * ORC_MAXBS], tt[ORC_MAXBS];
  float *prev = (float *)calloc((size_t)nbr * bs, sizeof(float));
  const int sweeps = jac > 0 ? jac : 1;
  for (int sw = 0; sw < sweeps; sw++) {
    for (int ii = 0; ii < nbr; ii++) {
      const int I = (upper && jac <= 0) ? nbr - 1 - ii : ii;
      const int kd = upper ? row[I] : row[I + 1] - 1;
      const int k0 = upper ? row[I] + 1 : row[I];
      const int k1 = upper ? row[I + 1] : row[I + 1] - 1;
      const float *yy = jac > 0 ? prev : y;
      for (int a = 0; a < bs; a++) {
        float t = 0.0f;
        if (!(jac > 0 && sw == 0))
          for (int k = k0; k < k1; k++)
            for (int c = 0; c < bs; c++)
              t = t + val[(size_t)k * lb + a * bs + c] * yy[(size_t)col[k] * bs + c];
        tt[a] = r[(size_t)I * bs + a] - t;
      }
      orc_binv_f(bs, &val[(size_t)kd * lb], inv);
      for (int a = 0; a < bs; a++) {
        float acc = 0.0f;
        for (int c = 0; c < bs; c++) acc = acc + inv[a * bs + c] * tt[c];
        y[(size_t)I * bs + a] = acc;
      }
    }
    if (jac > 0) memcpy(prev, y, sizeof(float) * (size_t)nbr * bs);
  }
  free(prev);
}
