/*
 * lace.h -- pure-C ABI of the MI355X-native LACE preconditioned-solve path.
 *
 * Two groups of entry points live here.
 *
 *  1. Drop-in entry points.  Same symbol names, argument lists and container
 *     layout as the reference's include/sparse.h, with host-array semantics:
 *     the caller's data_d_matrix arrays are authoritative, the library uploads
 *     them, runs the hot path on the GPU and writes results back into host
 *     arrays allocated with the reference allocator conventions (64-byte
 *     aligned, zero-filled, released with free()/data_zmfree).
 *     Differences from the reference are limited to:
 *       - errors are returned (DEV_* codes), never exit();
 *       - FGMRES returns info = 0 when converged (the reference inverts it,
 *         src/fgmres.cpp:495-497);
 *       - C++ default arguments are gone (sparse.h:152,665); pass every one.
 *
 *  2. lace_* device-resident entry points.  Opaque handles that keep A, the
 *     factors and the Krylov workspace in HBM across calls; no counterpart in
 *     the reference, which copies host<->device every sweep
 *     (src/parilu_v0_3_gpu.cpp:211).
 *
 * Struct layouts reproduce include/sparse_types.h:14-41,131-176 exactly,
 * including the 64-byte alignment of the pointer members, so a binary built
 * against the reference headers can pass its containers here unchanged.
 */
#ifndef LACE_MI355X_H
#define LACE_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* every declaration below is exported from the shared library */
#pragma GCC visibility push(default)

#define DEV_ALIGN 64

typedef int    data_int_t;   /* include/dense_types.h:95-99 */
typedef int    data_index_t;
typedef double dataType;     /* include/dense_types.h:105 */
typedef double real_Double_t;

/* Return codes: include/dense_types.h:134-160 */
#define DEV_SUCCESS               0
#define DEV_ERR                  -100
#define DEV_ERR_NOT_INITIALIZED  -101
#define DEV_ERR_NOT_SUPPORTED    -103
#define DEV_ERR_ILLEGAL_VALUE    -104
#define DEV_ERR_NOT_FOUND        -105
#define DEV_ERR_ALLOCATION       -106
#define DEV_ERR_HOST_ALLOC       -112
#define DEV_ERR_DEVICE_ALLOC     -113
#define DEV_ERR_INVALID_PTR      -115
#define DEV_ERR_UNKNOWN          -116
#define DEV_ERR_NOT_IMPLEMENTED  -117
#define DEV_ERR_NAN              -118
#define DEV_SLOW_CONVERGENCE     -201
#define DEV_DIVERGENCE           -202
#define DEV_ERR_BADPRECOND       -204
#define DEV_NOTCONVERGED         -205
/* library-specific: a HIP runtime call failed */
#define DEV_ERR_HIP              -301

/* Enumerations: include/dense_types.h:175-388 (MAGMA numbering kept). */
typedef enum { MagmaRowMajor = 101, MagmaColMajor = 102 } data_order_t;
typedef enum {
  MagmaUpper = 121, MagmaLower = 122, MagmaFull = 123, MagmaHessenberg = 124
} data_uplo_t;
typedef enum {
  Magma_CSR = 611, Magma_ELLPACKT = 612, Magma_ELL = 613, Magma_DENSE = 614,
  Magma_BCSR = 615, Magma_CSC = 616, Magma_HYB = 617, Magma_COO = 618,
  Magma_ELLRT = 619, Magma_SPMVFUNCTION = 620, Magma_SELLP = 621,
  Magma_ELLD = 622, Magma_CSRLIST = 623, Magma_CSRD = 624, Magma_CSRL = 627,
  Magma_CSRU = 628, Magma_CSRCOO = 629, Magma_CUCSR = 630, Magma_COOLIST = 631,
  Magma_CSR5 = 632, Magma_CSCD = 633, Magma_CSCL = 634, Magma_CSCU = 635,
  Magma_CSCCOO = 636, Magma_DENSEL = 637, Magma_DENSEU = 638,
  Magma_DENSED = 639, Magma_BCSRD = 640, Magma_BCSRL = 641, Magma_BCSRU = 642,
  Magma_BCSC = 643, Magma_BCSCD = 644, Magma_BCSCL = 645, Magma_BCSCU = 646
} data_storage_t;
typedef enum { Magma_GENERAL = 581, Magma_SYMMETRIC = 582 } data_symmetry_t;
typedef enum {
  Magma_ORDERED = 591, Magma_DIAGFIRST = 592, Magma_UNITY = 593,
  Magma_VALUE = 594, Magma_ZERO = 595, Magma_NODIAG = 596
} data_diagorder_t;

/* include/sparse_types.h:14-41 -- field order and alignment unchanged. */
typedef struct data_d_matrix {
  data_storage_t   storage_type;
  data_symmetry_t  sym;
  data_diagorder_t diagorder_type;
  data_uplo_t      fill_mode;
  data_int_t       num_rows;
  data_int_t       num_cols;
  data_int_t       nnz;
  data_int_t       max_nnz_row;
  data_int_t       diameter;
  data_int_t       true_nnz;
  dataType        *val    __attribute__((aligned(DEV_ALIGN)));
  dataType        *diag   __attribute__((aligned(DEV_ALIGN)));
  data_int_t      *row    __attribute__((aligned(DEV_ALIGN)));
  data_int_t      *rowidx __attribute__((aligned(DEV_ALIGN)));
  data_int_t      *col    __attribute__((aligned(DEV_ALIGN)));
  data_int_t      *list;
  data_int_t      *blockinfo;
  data_int_t       blocksize;
  data_int_t       ldblock;
  data_int_t       numblocks;
  data_int_t       alignment;
  data_order_t     major;
  data_int_t       ld;
  data_int_t       pad_rows;
  data_int_t       pad_cols;
} __attribute__((aligned(DEV_ALIGN))) data_d_matrix;

/* include/sparse_types.h:131-144 */
typedef struct data_z_preconditioner_log {
  data_int_t sweeps;
  data_int_t maxSweeps;
  dataType   tol;
  dataType   finalStep;
  dataType   A_Frobenius;
  dataType   precond_generation_time;
  dataType   initial_residual;
  dataType   initial_nonlinear_residual;
  dataType   residual;
  dataType   nonlinear_residual;
  data_int_t omp_num_threads;
} __attribute__((aligned(DEV_ALIGN))) data_d_preconditioner_log;

/* include/sparse_types.h:147-156 */
typedef struct data_z_gmres_log {
  data_int_t restarts;
  data_int_t search_directions;
  dataType   solve_time;
  dataType   initial_residual;
  dataType   final_residual;
  dataType   scaled_residual;
  dataType   original_residual;
} __attribute__((aligned(DEV_ALIGN))) data_d_gmres_log;

/* include/sparse_types.h:158-176.  user_csrtrsv_choice: 0 exact triangular
 * solves (the reference's mkl_dcsrtrsv; here level-scheduled on the GPU),
 * 1 ParCSRTRSV iterated sweeps to tol 1.0 (src/fgmres.cpp:129; here
 * synchronous Jacobi sweeps), 2 no preconditioner.  This library adds
 * 3 = fixed Jacobi-k sweeps with k = LACE_JACOBI_SWEEPS_DEFAULT. */
typedef struct data_z_gmres_param {
  data_int_t search_max;
  data_int_t tol_type;
  dataType   rtol;
  data_int_t reorth;
  data_int_t user_csrtrsv_choice;
  data_int_t monitorOrthog;
  data_int_t restart_max;
  data_int_t precondition;
  dataType   parilu_reduction;
  data_int_t reorder;
} __attribute__((aligned(DEV_ALIGN))) data_d_gmres_param;

#define LACE_JACOBI_SWEEPS_DEFAULT 5

/* ======================================================================
 * Drop-in entry points (reference include/sparse.h line cited per entry)
 * ====================================================================== */

/* Containers and formats (host).  sparse.h:147,165,170,177,181,380 */
/* major: the reference's C++ default argument is MagmaColMajor (sparse.h:152). */
int data_zvinit(data_d_matrix *x, int num_rows, int num_cols, dataType value,
                data_order_t major);
int data_rowindex(data_d_matrix *A, int **rowidx);
int data_zmconvert(data_d_matrix A, data_d_matrix *B,
                   data_storage_t old_format, data_storage_t new_format);
int data_zmtranspose(data_d_matrix A, data_d_matrix *B);
int data_zmfree(data_d_matrix *A);
int data_zmlumerge(data_d_matrix L, data_d_matrix U, data_d_matrix *A);

/* MatrixMarket I/O.  sparse.h:56 (read), control/magma_zmio.cpp:587 (write) */
int data_z_csr_mtx(data_d_matrix *A, const char *filename);
int data_zwrite_csr_mtx(data_d_matrix A, data_order_t MajorType,
                        const char *filename);
/* Dense MatrixMarket ("array") I/O: control/magma_zmio.cpp:1221-1318. */
int data_z_dense_mtx(data_d_matrix *A, data_order_t major, const char *filename);
int data_zwrite_dense_mtx(int n_row, int n_col, int nnz, data_order_t major,
                          dataType **val, const char *filename);
int data_zwrite_dense(data_d_matrix A, const char *filename);
/* dy <- alpha dx + dy on host arrays: blas/zaxpy.cpp:52-59. */
void data_zaxpy(int n, dataType alpha, const dataType *dx, int incx, dataType *dy, int incy);

/* Dot products on host arrays with strides.  sparse.h:507-517;
 * blas/zdot.cpp:50-78 (data_zdot: OpenMP reduction; data_zdot_mkl:
 * cblas_ddot).  Both sum in index order here. */
dataType data_zdot(int n, dataType *dx, int incx, dataType *dy, int incy);
dataType data_zdot_mkl(int n, dataType *dx, int incx, dataType *dy, int incy);

/* Norms and the ILU residual check.  sparse.h:410,462-468;
 * control/norms.cpp:69-92,301-405.  data_zilures evaluates ||A-LU||_F over
 * LU's pattern (res) and over A's pattern (nonlinres) on the host (a
 * once-per-factorization report); lace_parilu_nonlinres is the device form. */
dataType data_dnrm2(int n, dataType *x, int incx);
int data_zfrobenius(data_d_matrix A, dataType *res);
int data_zilures(data_d_matrix A, data_d_matrix L, data_d_matrix U,
                 data_d_matrix *LU, dataType *res, dataType *nonlinres);

/* Sparse products.  sparse.h:560-572.  data_z_spmv runs on the GPU;
 * data_z_spmm is the sparse x sparse product used by the residual check. */
int data_z_spmv(dataType alpha, data_d_matrix *A, data_d_matrix *x,
                dataType beta, data_d_matrix *y);
int data_z_spmm(dataType alpha, data_d_matrix A, data_d_matrix B,
                data_d_matrix *C);

/* Exact ILU(0) of A in place (the reference wraps MKL dcsrilu0,
 * blas/zcsrilu0.cpp:15-75).  Returns DEV_ERR_NAN on a pivot below 1e-16. */
int data_dcsrilu0_mkl(data_d_matrix *A);

/* ParILU(0).  sparse.h:774-804.  L is returned as CSRL (unit diagonal last
 * in each row), U as CSCU (U by columns, diagonal last).  All four run the
 * sweeps on the GPU in synchronous (Jacobi) mode: every update of sweep k
 * reads iterate k-1 only, so results are deterministic. */
/* sparse.h:767-771; src/parilu_v0_2.cpp:16-110: the same sweep stopped only
 * by step <= reduction*Ares (the reference has no cap; 1000 here). */
void data_PariLU_v0_2(data_d_matrix *A, data_d_matrix *L, data_d_matrix *U,
                      dataType reduction);

/* Tiled ParILU family.  sparse.h:813-869; src/parilu_v3_0.cpp:15-229 and
 * src/parilu_v3_{1,2,5,6,7,8,9}.cpp (the same fixed point with other tile
 * products).  Runs on the host, for CPU comparison: A is padded to whole
 * tiles in place (pad_rows, pad_cols, one unit entry per padding row, its
 * arrays replaced; control/magma_zmio.cpp:973-1040), L is returned as CSRL
 * (unit diagonal last), U as CSRU (diagonal first).  The tile grid is swept
 * in place in the order of the reference run on one thread, until the step
 * / ||A||_F is at most 1e-15 (at most 1000 sweeps in v3.0-v3.2, 10000 in
 * v3.5-v3.9); the log lines are the reference's.  On failure the reason is
 * in lace_last_error(). */
void data_PariLU_v3_0(data_d_matrix *A, data_d_matrix *L, data_d_matrix *U, int tile);
void data_PariLU_v3_1(data_d_matrix *A, data_d_matrix *L, data_d_matrix *U, int tile);
void data_PariLU_v3_2(data_d_matrix *A, data_d_matrix *L, data_d_matrix *U, int tile);
void data_PariLU_v3_5(data_d_matrix *A, data_d_matrix *L, data_d_matrix *U, int tile);
void data_PariLU_v3_6(data_d_matrix *A, data_d_matrix *L, data_d_matrix *U, int tile);
void data_PariLU_v3_7(data_d_matrix *A, data_d_matrix *L, data_d_matrix *U, int tile);
void data_PariLU_v3_8(data_d_matrix *A, data_d_matrix *L, data_d_matrix *U, int tile);
void data_PariLU_v3_9(data_d_matrix *A, data_d_matrix *L, data_d_matrix *U, int tile);
void data_PariLU_v0_3(data_d_matrix *A, data_d_matrix *L, data_d_matrix *U,
                      dataType reduction, data_d_preconditioner_log *log);
void data_PariLU_v0_3_prescribedSweeps(data_d_matrix *A, data_d_matrix *L,
                                       data_d_matrix *U, dataType reduction,
                                       data_d_preconditioner_log *log);
void data_PariLU_v0_3_gpu(data_d_matrix *A, data_d_matrix *L, data_d_matrix *U,
                          dataType reduction, data_d_preconditioner_log *log,
                          int nthreads);
void data_PariLU_v0_3_gpu_prescribedSweeps(data_d_matrix *A, data_d_matrix *L,
                                           data_d_matrix *U, dataType reduction,
                                           data_d_preconditioner_log *log);

/* Block ParILU(0) on a BCSR matrix.  src/parilu_v0_3_bcsr.cpp:16-276, which
 * defines the symbol data_PariLU_v0_3 a second time (it is in no makefile);
 * the block form is exported under its own name.  Synchronous FP64 sweeps
 * on the GPU until sum((new-old)^2) <= reduction (absolute, :61); at most
 * 1000 sweeps (the reference has no cap), DEV_NOTCONVERGED past it.  L is
 * returned as BCSRL (identity diagonal blocks last; tagged Magma_BCSR as the
 * reference's converter does, magma_zmconverter.cpp:1251), U as BCSCU
 * (block columns, diagonal block last, interiors not transposed).  Every
 * diagonal block must be present. */
int data_PariLU_v0_3_bcsr(data_d_matrix *A, data_d_matrix *L, data_d_matrix *U,
                          dataType reduction, data_d_preconditioner_log *log);

/* Triangular solves.  sparse.h:881-924; src/trisolve.cpp:20-119,256-351.
 * The reference sweeps in place (natural row order for L, reverse for U)
 * until sum((y_new-y_old)^2) <= tol.  Run by one thread -- the only schedule
 * with a defined result -- that is exact substitution in the first sweep and
 * an unchanged second sweep.  These entry points return that result: y is
 * the exact solution, computed on the GPU in level order (bit-identical to
 * substitution), and *iter is 1 when sum((y0-y)^2) over the caller's initial
 * y0 is already <= tol (data_parcsrtrsv: or not finite), else 2; with
 * tol >= 1e8 (the loop's initial step) no sweep runs, y is left as passed
 * and *iter is 0.  The two
 * solve wrappers print the reference's "% iteration = k step = s" lines
 * (option "matlab_log"). */
int data_forward_solve(data_d_matrix *L, data_d_matrix *x, data_d_matrix *rhs,
                       const dataType tol, int *iter);
int data_backward_solve(data_d_matrix *U, data_d_matrix *x, data_d_matrix *rhs,
                        const dataType tol, int *iter);
int data_parcsrtrsv(const data_uplo_t uplo, const data_storage_t storage,
                    const data_diagorder_t diag, const int num_rows,
                    const dataType *Aval, const int *row, const int *col,
                    const dataType *rhsval, dataType *yval,
                    const dataType tol, int *iter);

/* Allocation-free data_parcsrtrsv on device arrays.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and result as data_parcsrtrsv, except that Aval, row, col,
 * rhsval, yval and iter are device pointers and the call is asynchronous on
 * the library stream: *iter (device) receives 0, 1 or 2 as data_parcsrtrsv's
 * does, with the step summed in a fixed tree order, or -1 if the solve gave
 * up waiting.  It allocates nothing: its scratch is a fixed array on each
 * GPU (the caller's initial y for the step, partial sums, a block ticket),
 * so num_rows is at most LACE_NOALLOC_MAX_ROWS (DEV_ERR_ILLEGAL_VALUE
 * beyond).  Calls on one GPU share that array: calls that overlap on one
 * GPU race.  Call lace_init first. */
#define LACE_NOALLOC_MAX_ROWS 4194304
int lace_parcsrtrsv_noalloc(const data_uplo_t uplo, const data_storage_t storage,
                            const data_diagorder_t diag, const int num_rows,
                            const dataType *Aval, const int *row, const int *col,
                            const dataType *rhsval, dataType *yval,
                            const dataType tol, int *iter);

/* FGMRES.  sparse.h:1027-1052; src/fgmres.cpp:64-517,
 * src/fgmres_householder_restart.cpp:63-788.  L: CSRL, U: CSRU.  x0 is
 * freed and re-allocated with the solution (src/fgmres.cpp:491-492).  The
 * device Krylov bases hold at most 63 vectors (MGS) or reflectors
 * (Householder) per cycle: a larger search_max runs as cycles of 63 with the
 * same total number of search directions.  Progress lines go to stdout
 * unless "matlab_log" is 0. */
int data_fgmres(data_d_matrix *A, data_d_matrix *b, data_d_matrix *x0,
                data_d_matrix *L, data_d_matrix *U,
                data_d_gmres_param *gmres_par, data_d_gmres_log *gmres_log);
int data_fgmres_restart(data_d_matrix *A, data_d_matrix *b, data_d_matrix *x0,
                        data_d_matrix *L, data_d_matrix *U,
                        data_d_gmres_param *gmres_par,
                        data_d_gmres_log *gmres_log);
int data_fgmres_householder(data_d_matrix *A, data_d_matrix *b,
                            data_d_matrix *x0, data_d_matrix *L,
                            data_d_matrix *U, data_d_gmres_param *gmres_par,
                            data_d_gmres_log *gmres_log);
int data_fgmres_householder_restart(data_d_matrix *A, data_d_matrix *b,
                                    data_d_matrix *x0, data_d_matrix *L,
                                    data_d_matrix *U,
                                    data_d_gmres_param *gmres_par,
                                    data_d_gmres_log *gmres_log);

/* The reference's MKL RCI FGMRES baseline.  sparse.h:978-980;
 * src/mkl_fgmres.cpp:26-: ParILU v0.2 factors (parilu_reduction), exact
 * triangular solves, MGS FGMRES with restart length restart_max and an
 * iteration cap of search_max (rounded up to whole cycles), rtol relative
 * (tol_type 1) or absolute (0); precondition = 0 drops the preconditioner.
 * Runs on the device; the solution is written into x0->val in place. */
int data_MKL_FGMRES(data_d_matrix *A, data_d_matrix *x0,
                    data_d_matrix *rhs_vector, data_d_gmres_param *param);

/* Orderings.  sparse.h:398-402,662-665; src/cuthill_mckee.cpp:53-265,
 * src/reorder.cpp:167-253.  nn_map[i] = new number of row i; every quirk of
 * the reference ordering is kept (documented in lace_amd/csrc/reorder.cpp);
 * where the reference reads uninitialised memory this returns DEV_ERR.
 * data_sparse_reorder: reorder 1 = Cuthill-McKee, 2 = reversed, 0/other =
 * identity (3 = AMD is not provided: DEV_ERR_NOT_SUPPORTED; 4, an addition,
 * = lace_reorder_rcm); A becomes A(P, P) with rows sorted by column.
 * lace_bool is C++ bool / C _Bool. */
#ifdef __cplusplus
typedef bool lace_bool;
#else
typedef _Bool lace_bool;
#endif
int data_reorder_cuthill_mckee(data_d_matrix *A, int *nn_map, lace_bool reverse);
int data_sparse_reorder(data_d_matrix *A, int *P, int *Pinv, int reorder);
/* Reverse Cuthill-McKee in the textbook form (George & Liu 1981): on the
 * symmetrised graph without self loops, each component from a
 * pseudo-peripheral node, neighbours in increasing (degree, number), the
 * whole order reversed.  P[k] = old number of the row placed at k.  No
 * counterpart in the reference, whose ordering keeps its quirks (above);
 * this one is for callers that want RCM's bandwidth reduction. */
int lace_reorder_rcm(const data_d_matrix *A, int *P);

/* Householder generator and triangularization (host, small dense).
 * sparse.h:1059-1067; src/householder.cpp:49-160.  Dense containers are
 * column-major with leading dimension ld; data_hqrd overwrites X and
 * addresses R with X's leading dimension, as the reference does. */
void data_housegen(int n, dataType *X, dataType *u, dataType *nu);
void data_hqrd(data_d_matrix *X, data_d_matrix *U, data_d_matrix *R);

/* Givens generator.  sparse.h:1096; src/givens.cpp:46-67 */
void givens_rotation(dataType a, dataType b, dataType tol, dataType *c,
                     dataType *s);

/* ======================================================================
 * lace_* device-resident API (no reference counterpart; see header note)
 * ====================================================================== */

typedef enum { LACE_F64 = 0, LACE_F32 = 1, LACE_BF16 = 2, LACE_F16 = 3 } lace_prec_t;

typedef enum {
  LACE_TRSV_LEVEL = 0,   /* exact, level-scheduled */
  LACE_TRSV_JACOBI = 1,  /* k synchronous Jacobi sweeps from y0 = 0 */
  LACE_TRSV_NONE = 2     /* identity preconditioner */
} lace_trsv_kind_t;

typedef struct lace_csr lace_csr;         /* device CSR matrix */
typedef struct lace_parilu lace_parilu;   /* device ParILU factors */
typedef struct lace_solver lace_solver;   /* device FGMRES solver */
typedef struct lace_bcsr lace_bcsr;       /* device BCSR matrix */
typedef struct lace_bilu lace_bilu;       /* device block ParILU factors */

/* Runtime.  lace_init selects the device; lace_last_error returns the text
 * of the most recent failure on this thread. */
int lace_init(int device);
const char *lace_last_error(void);
int lace_device_synchronize(void);

/* Raw device buffers (bytes) and copies, on the library stream. */
int lace_dmalloc(void **ptr, size_t bytes);
int lace_dfree(void *ptr);
int lace_memcpy_h2d(void *dst, const void *src, size_t bytes);
int lace_memcpy_d2h(void *dst, const void *src, size_t bytes);
/* Device-to-device streaming copy kernel (bytes a multiple of 16); used to
 * calibrate hardware traffic counters against a known byte count. */
int lace_stream_copy(void *dst, const void *src, size_t bytes);

/* Device CSR upload (values converted to prec). */
int lace_csr_create(const data_d_matrix *A, lace_prec_t prec, lace_csr **out);
int lace_csr_destroy(lace_csr *A);
int lace_csr_n(const lace_csr *A);
/* y = alpha*A*x + beta*y on device vectors of A's precision. */
int lace_csr_spmv(const lace_csr *A, double alpha, const void *x, double beta,
                  void *y);
/* Storage forms held besides CSR: ndiag > 0 when the entries lie on ndiag
 * (at most 8) diagonals and the diagonal form applies; ell_width > 0 when
 * the ELLPACK form (that many slots per row) applies.  The products use the
 * diagonal form, else ELLPACK, else CSR (options "spmv_dia", "spmv_ell"). */
int lace_csr_forms(const lace_csr *A, int *ndiag, int *ell_width);

/* ParILU(0) on device.  Factor storage in prec: F64 (FP64 iteration), F32,
 * or BF16 / F16 (FP32 iteration; at reset and after every sweep each factor
 * value is rounded to 16 bits relative to the power-of-two scale of its row,
 * 2^e <= max_j |a_ij| < 2^(e+1), with L's unit diagonal kept exact; the
 * Jacobi sweeps of L and U then read 16-bit words).  Vectors passed to the
 * solves are FP64 for F64 and FP32 otherwise. */
int lace_parilu_create(const data_d_matrix *A, lace_prec_t prec,
                       lace_parilu **out);
int lace_parilu_destroy(lace_parilu *P);
/* Re-initialise the factors from A (L0 = strict-lower(A) + I, U0 = upper(A)). */
int lace_parilu_reset(lace_parilu *P);
/* New values of A on the pattern given at create (host array, nnz entries
 * in A's CSR order).  The factors are unchanged until the next
 * lace_parilu_reset, which starts from the new values. */
int lace_parilu_set_values(lace_parilu *P, const double *vals);
/* Run nsweeps synchronous sweeps.  If step is non-NULL, the sum of squared
 * updates of the last sweep divided by ||A||_F is written there (this
 * synchronizes with the host). */
int lace_parilu_sweeps(lace_parilu *P, int nsweeps, double *step);
/* Copy factors to host: L as CSRL (unit diag last), U as CSCU. */
int lace_parilu_get(lace_parilu *P, data_d_matrix *L, data_d_matrix *U);
/* ||A - LU||_F restricted to the pattern of A, evaluated on the device. */
int lace_parilu_nonlinres(lace_parilu *P, double *nonlinres);
/* Per-sweep HBM bytes by the algorithmic count (see DESIGN). */
double lace_parilu_sweep_bytes(const lace_parilu *P);
/* Apply the factors: z = U^{-1} L^{-1} r (trans = 0) or
 * z = L^{-T} U^{-T} r (trans = 1).  r, z are device vectors in the
 * factor's arithmetic precision (F64 or F32). */
int lace_parilu_apply(lace_parilu *P, lace_trsv_kind_t kind, int jacobi_sweeps,
                      int trans, const void *r, void *z);
int lace_parilu_levels(const lace_parilu *P, int *nlev_L, int *nlev_U);
/* Storage forms of the off-diagonal entries of L (upper = 0) or U
 * (upper = 1) for the Jacobi sweeps, as for lace_csr_forms (options
 * "jacobi_dia", "jacobi_ell"). */
int lace_parilu_forms(const lace_parilu *P, int upper, int *ndiag, int *ell_width);
/* Block order of the exact solves with L (upper = 0), U (1), U^T (2) or
 * L^T (3): the box sides dims[0..2] where the rows are taken box by box
 * (option "trsv_exact_tiles"), 0 0 0 where blocks are consecutive rows.
 * A diagnostic; no counterpart in the reference. */
int lace_parilu_tiles(const lace_parilu *P, int which, int *dims);
/* Sweep kernel selection: 0 (default) runs the four-position streaming
 * symmetric kernel (5) when A's pattern is structurally symmetric, else the
 * general one;
 * 1 always runs the general CSR-L / CSC-U kernel; 2 the per-position
 * symmetric kernel that merges rows i and j on every sweep; 3 the per-row
 * symmetric kernel (rows with at most 4 strict-lower entries); 4 the
 * streaming symmetric kernel (merge partners recorded once at set-up); 5 the
 * streaming kernel with four positions per thread (16-byte accesses).
 * All give identical factors. */
int lace_parilu_set_variant(lace_parilu *P, int variant);
/* Non-zero if a sync-free level solve or a pipelined Jacobi launch gave up
 * waiting (a wait exceeded its poll budget); the result of that solve is
 * then invalid.  The flag stays set for the handle's lifetime. */
int lace_syncfree_error(lace_parilu *P, int *err);

/* Device BCSR matrix (A->storage_type Magma_BCSR, num_rows = block rows,
 * blocks row-major, blocksize 1..8); values converted to prec (F64, F32). */
int lace_bcsr_create(const data_d_matrix *A, lace_prec_t prec, lace_bcsr **out);
int lace_bcsr_destroy(lace_bcsr *A);
/* y = A x on device vectors of A's precision (length num_rows*blocksize). */
int lace_bcsr_spmv(const lace_bcsr *A, const void *x, void *y);
/* Algorithmic HBM bytes of one lace_bcsr_spmv. */
double lace_bcsr_spmv_bytes(const lace_bcsr *A);

/* Block ParILU(0) on device (src/parilu_v0_3_bcsr.cpp:118-250 in
 * synchronous form).  prec is the factor storage: F64 (FP64 iteration), F32,
 * BF16 or F16 (FP32 iteration; 16-bit storage keeps a power-of-two scale per
 * scalar row taken from A's row maxima).  Every diagonal block must be
 * present; block columns sorted within each block row. */
int lace_bilu_create(const data_d_matrix *A, lace_prec_t prec, lace_bilu **out);
int lace_bilu_destroy(lace_bilu *B);
/* L0 = strict-lower blocks of A with identity diagonal blocks, U0 = upper
 * blocks of A. */
int lace_bilu_reset(lace_bilu *B);
/* nsweeps synchronous sweeps; step (optional, synchronizes) receives the
 * unscaled sum of squared changes of the last sweep. */
int lace_bilu_sweeps(lace_bilu *B, int nsweeps, double *step);
/* Host copies: L as BCSRL, U as BCSCU (FP64 values). */
int lace_bilu_get(lace_bilu *B, data_d_matrix *L, data_d_matrix *U);
double lace_bilu_sweep_bytes(const lace_bilu *B);
/* z = U^{-1} L^{-1} r with exact level-scheduled block solves or k Jacobi
 * sweeps; r, z device vectors in the iteration precision (double for F64
 * storage, float otherwise). */
int lace_bilu_apply(lace_bilu *B, lace_trsv_kind_t kind, int jacobi_sweeps,
                    const void *r, void *z);
int lace_bilu_levels(const lace_bilu *B, int *nlev_L, int *nlev_U);
/* Sweep kernel: 0 (default) pairs L_IJ with U_JI in one lane group when A's
 * block pattern is structurally symmetric, else one group per block of A;
 * 1 always takes the per-block kernel.  The paired kernel reads
 * per-position tables built at set-up; each wave moves its positions' A, L
 * and U blocks as runs of whole words through its own slice of LDS (where
 * the slices fit 64 KiB), else the workgroup stages L, U and U_JJ through
 * LDS (where they align to 16 bytes and fit), else it merges the two block
 * rows on the device.  2-6 force a form of the paired kernel: 2 the merge,
 * 3 the tables without staging, 4 the tables with A's blocks staged too,
 * 5 the tables with L, U and U_JJ staged, 6 the wave runs.  All give
 * identical factors. */
int lace_bilu_set_variant(lace_bilu *B, int variant);
/* 1 if A's block pattern is structurally symmetric. */
int lace_bilu_is_symmetric(const lace_bilu *B);

/* Tall-skinny FP64 products on MFMA (n x 32 row-major device blocks; the
 * multi-RHS layout).  blocks: host array of nb device pointers (nb <= 40).
 * tn: C[b] (32 x 32 row-major) = X_b^T W over rows [rs_b, n) with
 *     rs_b = rs_main for b < nb-1 and rs_last for the last block.
 * nn: out = base - sum_b X_b D_b on rows [rlo, n), base on rows below;
 *     D holds nb consecutive 32 x 32 row-major blocks; base_mode 0 zero,
 *     1 the block base, 2 the identity block column starting at row e0. */
int lace_tsgemm_tn(int n, int nb, const double *const *blocks, const double *W,
                   int rs_main, int rs_last, double *C);
int lace_tsgemm_nn(int n, int nb, const double *const *blocks, const double *D,
                   const double *base, int base_mode, int e0, int rlo, double *out);

/* Solver configuration. */
typedef struct lace_solver_opts {
  int    krylov;        /* 0 = Householder FGMRES, 1 = MGS FGMRES */
  int    search_max;    /* restart length m */
  int    restart_max;
  double rtol;          /* relative to ||b - A x0|| */
  lace_prec_t prec_precond;   /* factor storage: F64, F32, BF16 or F16 */
  lace_trsv_kind_t trsv;
  int    jacobi_sweeps;
  int    parilu_sweeps;
  int    inner_iters;   /* >0: nested mixed mode -- each outer FP64 step
                           preconditions with inner_iters FP32 GMRES steps */
  int    transpose;     /* 1: solve A^T x = b with transposed factors */
  int    inner_fp64;    /* with inner_iters > 0: 0 (default) inner GMRES in
                           FP32 on an FP32 copy of A (factors at most FP32);
                           1: inner GMRES in FP64 on A with FP64 factors
                           (prec_precond F64), the same nested structure
                           without the precision change */
} lace_solver_opts;

typedef struct lace_solver_stats {
  int    iterations;      /* outer search directions over all restarts */
  int    restarts;
  int    inner_iterations;
  double initial_residual;
  double final_residual_est;  /* |g_{j+1}| */
  double true_residual;       /* ||b - A x||_2 in FP64 */
  double rel_true_residual;   /* true_residual / ||b||_2 */
  double setup_ms;
  double solve_ms;
  int    converged;
} lace_solver_stats;

void lace_solver_default_opts(lace_solver_opts *o);
/* A is CSR (scalar ParILU factors) or BCSR (Magma_BCSR: block ParILU
 * factors, block SpMV; prec_precond may then also be BF16 or F16).
 * Transposed solves, value updates and value gradients take CSR only. */
int lace_solver_create(const data_d_matrix *A, const lace_solver_opts *o,
                       lace_solver **out);
int lace_solver_destroy(lace_solver *S);
/* ParILU setup from A (reset + parilu_sweeps sweeps), on device. */
int lace_solver_setup(lace_solver *S);
/* Solve with device vectors b, x (FP64, length n); x is the initial guess
 * on entry and the solution on exit. */
int lace_solver_solve(lace_solver *S, const double *b, double *x,
                      lace_solver_stats *st);
/* Same solve with A (trans = 0) or A^T (trans = 1).  The transposed solve
 * reuses the ILU factors of A: M^{-T} = L^{-T} U^{-T} (U^T then L^T sweeps)
 * inside the same FGMRES; A^T is assembled on the first transposed call. */
int lace_solver_solve_trans(lace_solver *S, const double *b, double *x, int trans,
                            lace_solver_stats *st);
/* Replace the operator's values on A's pattern (host array of nnz values).
 * The ILU factors in use stay those of the last lace_solver_setup; the next
 * lace_solver_setup factors the new values. */
int lace_solver_set_values(lace_solver *S, const double *vals);
/* dL/dtheta on the pattern of A: dtheta[k] = -lambda[row(k)] * x[col(k)]
 * for CSR position k; device vectors.  With A x = b and A^T lambda = dL/dx
 * this is the gradient of L with respect to A's stored values (and
 * dL/db = lambda). */
int lace_solver_grad_values(lace_solver *S, const double *lambda,
                            const double *x, double *dtheta);

/* ---- 32 right-hand sides at once (no reference counterpart: the
 * reference solves one right-hand side per call, src/fgmres.cpp:64 and
 * src/fgmres_householder_restart.cpp:63).  Blocks are n x 32 device arrays,
 * row-major: row i holds entry i of the 32 right-hand sides. */
#define LACE_NRHS 32

typedef struct lace_block_stats {
  int    iterations;        /* block Arnoldi steps over all restarts */
  int    restarts;          /* restart cycles started */
  int    panel_passes;      /* CholeskyQR passes over all panels */
  int    shifted_panels;    /* panels that needed a shifted first pass */
  int    converged;         /* every column reached rtol */
  double max_rel_true_residual;
  double rel_true_residual[LACE_NRHS];  /* ||b_c - A x_c|| / ||b_c|| */
  double initial_residual[LACE_NRHS];   /* ||b_c - A x0_c|| */
  double solve_ms;
} lace_block_stats;

/* Block Householder FGMRES(search_max) on the operator of S (trans = 0: A;
 * 1: A^T with M^{-T}): 32 right-hand sides share one block Krylov space.
 * Per block step: preconditioned block Z_k = M^{-1} V_k (Jacobi sweeps of
 * the ILU factors in their precision, Z stored in it; with inner_iters > 0
 * and FP32 factors, inner_iters steps of batched FP32 GMRES per column,
 * each preconditioned by the FP32 ILU), W = A Z_k (FP64 A),
 * the FP64 block reflectors applied to W as one compact-WY product (MFMA
 * Gram X^T W and update W - X D), the new panel's Householder QR by
 * CholeskyQR2 (three passes after a shifted first one) and reconstruction,
 * and the block Hessenberg least squares by 64 x 32 Householder windows on
 * device.  Column c converges when its residual estimate is <= rtol *
 * ||b_c - A x0_c||; the solve stops when all have.  X is the initial guess
 * on entry and the solution on exit.  Needs trsv = Jacobi (or none),
 * inner_iters <= 12, n >= 32 (search_max + 2), search_max <= 39. */
int lace_solver_solve_block(lace_solver *S, const double *B, double *X, int trans,
                            lace_block_stats *st);
/* dtheta[k] = -sum_c Lambda[row(k)][c] * X[col(k)][c]: the gradient of
 * sum_c L_c with respect to A's stored values, for A X = B and
 * A^T Lambda = [dL_c/dx_c]. */
int lace_solver_grad_values_block(lace_solver *S, const double *Lambda, const double *X,
                                  double *dtheta);
/* Y = A X (B == NULL) or Y = B - A X on n x 32 blocks; A FP64; X FP64
 * (x_prec LACE_F64) or FP32 (LACE_F32). */
int lace_csr_spmm32(const lace_csr *A, const void *X, lace_prec_t x_prec, const double *B,
                    double *Y);
/* Z = M^{-1} R (trans 0) or M^{-T} R (trans 1) on n x 32 blocks in the
 * factor precision (Jacobi sweeps or kind NONE). */
int lace_parilu_apply_block(lace_parilu *P, lace_trsv_kind_t kind, int jacobi_sweeps, int trans,
                            const void *R, void *Z);
/* out = a x + b y on device arrays of m doubles (a term with a zero
 * coefficient is not read), e.g. the adjoint right-hand sides x - x_t. */
int lace_daxpby(size_t m, double a, const double *x, double b, const double *y, double *out);
/* Householder QR of rows [r0, n) of the n x 32 block W (device): writes
 * Y (n x 32: zero above r0, unit lower triangle at rows r0..r0+31), Tb
 * and Rt (32 x 32 row-major) with (I - Y Tb Y^T) [Rt; 0] = W[r0:, :];
 * passes = CholeskyQR passes used. */
int lace_block_panel_qr(int n, int r0, const double *W, double *Y, double *Tb, double *Rt,
                        int *passes);

/* Kernel-shape options for tuning runs.  "stream_rows_per_thread" (1, 2 or
 * 4): rows each thread owns in the CSR-stream SpMV and Jacobi sweeps
 * (default: 2 for FP32 values, 1 for FP64).  "jacobi_ell" (0 or 1, default
 * 1): Jacobi sweeps of ILU factors with short rows use the ELLPACK form of
 * the off-diagonal entries (0: the off-diagonal CSR-stream form).
 * "spmv_ell" (0 or 1, default 1): SpMV of a matrix with at most 8 entries
 * per row uses its ELLPACK copy (0: the CSR-stream kernel).  "jacobi_dia",
 * "spmv_dia" (0 or 1, default 1): the same with the diagonal form, which
 * takes precedence over ELLPACK where both apply.  "jacobi_fuse" (0 or 1,
 * default 1): with an odd sweep count, the first Jacobi sweep of U rides on
 * the last sweep of L (diagonal form only).  "jacobi_dia_vec" (0 or 1,
 * default 1): Jacobi sweeps on the diagonal form take a 16-byte word of
 * rows per thread (4 FP32, 2 FP64) where n is a multiple of a word and the
 * vectors are 16-byte aligned (0: one row per thread).  "jacobi_pipe" (0 or
 * 1, default 0): a ParILU apply with k >= 3 Jacobi sweeps, both factors on
 * the diagonal form, runs sweeps 2..k of each factor in one launch (tiles
 * in ticket order, each waiting on the tiles it reads through progress
 * flags; the first sweep of U rides on L's launch); slower than a launch
 * per sweep on the 128^3 operator.  "mrhs_plane_order" (0 or 1, default 1):
 * the 32-right-hand-side product and Jacobi sweeps on ELLPACK forms run
 * their row blocks plane segment by plane segment when the pattern's
 * largest diagonal offset splits the rows into planes of more than 1 MB of
 * right-hand rows (3-D grids), so that the rows a row reads at +-1 plane
 * are still in the L2 cache (0: natural row order).  "jacobi_pair" (0 or 1,
 * default 1): the 32-right-hand-side Jacobi sweeps on the diagonal form run
 * two to a launch, each row forming the first sweep's result at the rows
 * it reads (0: one sweep per launch).  Every choice gives bit-identical
 * results.  "matlab_log" (0 or 1, default 1): the
 * drop-in FGMRES entry points print the reference's MATLAB-syntax progress
 * lines to stdout (src/fgmres.cpp:71,178,430; src/fgmres_householder_restart.cpp:
 * 70,81,196,231,707) -- "FGMRES_Householders_partrsv_search(j) = |g_j|;" etc.,
 * the log format its post-processing scripts read; 0 silences them.
 * "jacobi_dia16" (0 or 1, default 1): with 16-bit factor storage the Jacobi
 * sweeps on the diagonal form read the 16-bit words (0: the FP32 copy of
 * the same values).
 * "trsv_exact" (0 or 1, default 1): the exact (level) triangular solves run
 * the persistent kernel (1: blocks of rows taken in a topological ticket
 * order, dependencies inside a block resolved through LDS, across
 * blocks through the solution vector with a NaN sentinel per unsolved row)
 * or the level-ordered sync-free kernel (0: one thread per row).
 * "trsv_exact_rows" (256 or 512, default 512): rows per block of the
 * persistent kernel.  "trsv_exact_wgs_per_cu" (0..16, default 0 = as many
 * as fit): its resident workgroups per CU.  "trsv_exact_poll_sleep" (0..64,
 * default 0): its sleeps of 64 cycles between two polls of the solution
 * vector.  "trsv_exact_idle" (0..16, default 8): the most sleeps of 64
 * cycles a wave whose rows could not move takes before it polls again.  "trsv_exact_tiles" (0 or 1, default 1): where the factor's rows
 * are the points of a grid whose dependencies all step backwards along
 * each axis (7- and 5-point stencils, recognised from the pattern alone),
 * the persistent kernel's blocks are boxes of the grid (8 x 8 x 8, or
 * 32 x 16 in 2-D) taken in order of their diagonal (0: blocks of
 * consecutive rows).  Either way the result is bit-identical.
 * "inner_ortho" (0 or 1, default 1): the inner GMRES of a nested solve
 * (inner_iters > 0) orthogonalises by classical Gram-Schmidt on an
 * unnormalised basis (1: two vector passes and one small launch per step)
 * or by Householder reflectors (0: the outer engine's).  The two are
 * different orthogonalisations of the same Krylov space: results agree to
 * rounding, not bit for bit.
 * "timing_sample" (>= 1, default 1): the timing regions below record events
 * around the first four entries of each region after a reset and every
 * k-th entry after that only. */
int lace_set_option(const char *name, int value);

/* Kernel timing by class, HIP events on the library stream. */
int lace_timing_enable(int on);
int lace_timing_reset(void);
/* Restrict timing to a comma-separated list of region names (NULL or ""
 * times every region).  Per-launch regions of the hot kernels:
 * "parilu_sweep", "spmv" (FP64 matrix), "spmv_f32a" (FP32 matrix),
 * "trsv_jacobi_sweep_lower" / "_upper" (one stream Jacobi sweep). */
int lace_timing_only(const char *name);
/* Summed duration and number of the timed entries of a region. */
int lace_timing_get(const char *name, double *total_ms, long *count);
/* Entries of a region while timing was on, timed or skipped by sampling. */
int lace_timing_get_calls(const char *name, long *calls);
/* Algorithmic work summed over the timed launches of a region: flops and
 * HBM bytes (recorded by the regions whose shape varies between launches:
 * "tsgemm_tn", "tsgemm_nn", and the block kernels "spmm32", "spmm32_f32x",
 * "spmm32_residual", "block_jacobi_sweep_lower" / "_upper"). */
int lace_timing_get_work(const char *name, double *flops, double *bytes);

/* Level sets (host): lev(i) = 1 + max lev(k) over the off-diagonal entries
 * of row i (lower: k < i; upper: k > i), 0 when there are none.  Writes
 * perm (rows sorted by level, ties by row index) and level_ptr
 * (nlevels+1 entries) and returns nlevels, or a negative DEV code. */
int lace_level_sets(int n, const int *row, const int *col, int upper,
                    int *perm, int *level_ptr);

/* Synthetic problem generators (host). */
int lace_gen_laplace2d(int nx, int ny, data_d_matrix *A);
int lace_gen_convdiff3d(int nx, int ny, int nz, double pe_x, double pe_y,
                        double pe_z, data_d_matrix *A);
/* BCSR with bs x bs blocks on the 7-point block pattern of an nx*ny*nz grid
 * (x fastest, Dirichlet: blocks outside the box dropped, block columns
 * ascending).  Entries are sigma * N(0,1) from lace_gen_normal(seed) in
 * storage order (block by block, row-major inside), and shift is added to
 * the diagonal entries of the diagonal blocks. */
int lace_gen_bcsr7(int nx, int ny, int nz, int bs, uint64_t seed, double sigma,
                   double shift, data_d_matrix *A);
int lace_gen_normal(uint64_t seed, int n, double *out);
uint32_t *lace_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                             uint32_t out[4]);

#pragma GCC visibility pop

#ifdef __cplusplus
}
#endif
#endif /* LACE_MI355X_H */
