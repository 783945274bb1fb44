// solver.hip -- ILU-preconditioned FGMRES solver handles (FP64 and mixed
// precision), the adjoint gradient gather, and the drop-in FGMRES entry
// points.
//
// Precision layouts:
//   FP64 path   : A, factors, triangular solves, SpMV, reflectors all FP64.
//   mixed path  : the outer FGMRES (reflectors, Hessenberg/Givens, least
//                 squares, the Arnoldi/true-residual SpMV with A in FP64) is
//                 FP64; the preconditioner runs in FP32 -- FP32 factors and
//                 triangular solves, and with inner_iters > 0 an inner FP32
//                 Householder GMRES on an FP32 copy of A.  The outer FGMRES
//                 therefore acts as flexible iterative refinement of the FP32
//                 solve; preconditioned vectors are stored in FP32 and the
//                 outer SpMV reads them widened, so the Arnoldi relation holds
//                 exactly for the stored vectors.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#include "bcsr.hpp"
#include "bkrylov.hpp"
#include "common.hpp"
#include "device.hpp"
#include "krylov.hpp"

namespace lace {

// ------------------------------------------------------------ operators
struct CsrOp : LinOp {
  const DevCSR* A;
  explicit CsrOp(const DevCSR* a) : A(a) {}
  int apply_dd(const double* x, double* y) const override { return spmv<double, double>(*A, 1.0, x, 0.0, y); }
  int apply_fd(const float* x, double* y) const override { return spmv<float, double>(*A, 1.0, x, 0.0, y); }
  int apply_ff(const float* x, float* y) const override { return spmv<float, float>(*A, 1.0, x, 0.0, y); }
  int residual_dd(const double* x, const double* b, double* r) const override {
    return residual<double, double>(*A, x, b, r);
  }
};

struct BcsrOp : LinOp {
  const DevBCSR* A;
  explicit BcsrOp(const DevBCSR* a) : A(a) {}
  int apply_dd(const double* x, double* y) const override { return bcsr_spmv<double, double>(*A, x, y); }
  int apply_fd(const float* x, double* y) const override { return bcsr_spmv<float, double>(*A, x, y); }
  int apply_ff(const float* x, float* y) const override { return bcsr_spmv<float, float>(*A, x, y); }
  int residual_dd(const double* x, const double* b, double* r) const override {
    return bcsr_residual<double, double>(*A, x, b, r);
  }
};

template <typename K>
int hh_start(HHEngine<K>& E, int set_rtol, double rtol_rel);
template <typename K>
int hh_step(HHEngine<K>& E, int j, bool write_q);
template <typename K, typename XT>
int hh_finish(HHEngine<K>& E, XT* x);
template <typename K>
int cgs_start(HHEngine<K>& E, const double* r);
template <typename K>
int cgs_step(HHEngine<K>& E, int j, bool last, double rtol_rel);
template <typename K>
int read_misc(HHEngine<K>& E);
int norm2_sq_partials(int n, int rpb, const double* x, double* partial, int pstride, int nb);
int sum_slot(int nb, int pstride, const double* partial, double* out);
int cast_d2f(int n, const double* in, float* out);
int cast_f2d(int n, const float* in, double* out);
int set_misc_d(double* small, SmallLayout lay, int slot, double v);

// ------------------------------------------------------------ preconditioners
struct PrecILU : Precond {
  ParILU* P;
  lace_trsv_kind_t kind;
  int sweeps, trans;
  DevBuf qf;
  PrecILU(ParILU* p, lace_trsv_kind_t k, int s, int t) : P(p), kind(k), sweeps(s), trans(t) {}
  bool z_is_float() const override { return P->prec != LACE_F64; }
  int apply_d(const double* q, void* z) override {
    TimedRegion tr("precond");
    if (P->prec == LACE_F64) return P->apply(kind, sweeps, trans, q, z);
    if (qf.bytes < sizeof(float) * P->n) LACE_TRY(qf.alloc(sizeof(float) * P->n));
    LACE_TRY(cast_d2f(P->n, q, qf.as<float>()));
    return P->apply(kind, sweeps, trans, qf.as<float>(), z);
  }
  int apply_f(const float* q, void* z) override {
    TimedRegion tr("precond");
    LACE_REQUIRE(P->prec != LACE_F64, DEV_ERR_NOT_SUPPORTED, "FP32 input to FP64 factors");
    return P->apply(kind, sweeps, trans, q, z);
  }
};

// Block ILU factors: FP64 storage takes FP64 vectors; F32/BF16/F16 storage
// iterates in FP32, so FP64 input is narrowed first and z is FP32.
struct PrecBILU : Precond {
  BILU* B;
  lace_trsv_kind_t kind;
  int sweeps;
  DevBuf qf;
  PrecBILU(BILU* b, lace_trsv_kind_t k, int s) : B(b), kind(k), sweeps(s) {}
  bool z_is_float() const override { return B->store != LACE_F64; }
  int apply_d(const double* q, void* z) override {
    TimedRegion tr("precond");
    if (B->store == LACE_F64) return B->apply(kind, sweeps, q, z);
    if (qf.bytes < sizeof(float) * B->n()) LACE_TRY(qf.alloc(sizeof(float) * B->n()));
    LACE_TRY(cast_d2f(B->n(), q, qf.as<float>()));
    return B->apply(kind, sweeps, qf.as<float>(), z);
  }
  int apply_f(const float* q, void* z) override {
    TimedRegion tr("precond");
    LACE_REQUIRE(B->store != LACE_F64, DEV_ERR_NOT_SUPPORTED, "FP32 input to FP64 block factors");
    return B->apply(kind, sweeps, q, z);
  }
};

struct PrecIdentity : Precond {
  int n;
  explicit PrecIdentity(int n_) : n(n_) {}
  bool z_is_float() const override { return false; }
  int apply_d(const double* q, void* z) override {
    LACE_HIP(hipMemcpyAsync(z, q, sizeof(double) * n, hipMemcpyDeviceToDevice, stream()));
    return DEV_SUCCESS;
  }
};

// Inner Householder GMRES with a fixed ILU preconditioner, `iters` steps
// from x = 0, no host synchronisation.  K = float: FP32 engine on the FP32
// copy of A with FP32 (or 16-bit) factors, output in FP32; K = double: the
// same in FP64 on A with FP64 factors (the nested structure without the
// precision change).
template <typename K>
struct PrecInner : Precond {
  const LinOp* Aop;
  Precond* M;
  HHEngine<K> E;
  int n, iters;
  double rtol;
  PrecInner(const LinOp* a, Precond* m, int n_, int it, double rt)
      : Aop(a), M(m), n(n_), iters(it), rtol(rt) {}
  int init() { return E.init(n, iters, sizeof(K) == 4); }
  bool z_is_float() const override { return sizeof(K) == 4; }
  int apply_d(const double* q, void* z) override {
    TimedRegion tr("inner_gmres");
    if (get_option("inner_ortho", 1) == 1) return apply_cgs(q, static_cast<K*>(z));
    if constexpr (sizeof(K) == 4) LACE_TRY(cast_d2f(n, q, E.r.template as<float>()));
    else LACE_HIP(hipMemcpyAsync(E.r.p, q, sizeof(double) * n, hipMemcpyDeviceToDevice, stream()));
    LACE_HIP(hipMemsetAsync(z, 0, sizeof(K) * n, stream()));
    LACE_TRY(hh_start<K>(E, 1, rtol));
    for (int j = 0; j < iters; j++) {
      if constexpr (sizeof(K) == 4) {
        LACE_TRY(M->apply_f(E.q.template as<float>(), E.zcol(j)));
        LACE_TRY(Aop->apply_ff(static_cast<const float*>(E.zcol(j)), E.w.template as<float>()));
      } else {
        LACE_TRY(M->apply_d(E.q.template as<double>(), E.zcol(j)));
        LACE_TRY(Aop->apply_dd(static_cast<const double*>(E.zcol(j)), E.w.template as<double>()));
      }
      LACE_TRY(hh_step<K>(E, j, j < iters - 1));
    }
    return hh_finish<K, K>(E, static_cast<K*>(z));
  }
  // The same fixed inner GMRES with classical Gram-Schmidt on an
  // unnormalised basis (option "inner_ortho" 1, the default; 0 takes the
  // Householder engine): the preconditioner reads the basis vector W_j as
  // stored.
  int apply_cgs(const double* q, K* z) {
    LACE_HIP(hipMemsetAsync(z, 0, sizeof(K) * n, stream()));
    LACE_TRY(cgs_start<K>(E, q));
    K* W = E.U.template as<K>();
    for (int j = 0; j < iters; j++) {
      const K* wj = W + (size_t)j * E.ld;
      if constexpr (sizeof(K) == 4) {
        LACE_TRY(M->apply_f(wj, E.zcol(j)));
        LACE_TRY(Aop->apply_ff(static_cast<const float*>(E.zcol(j)), E.w.template as<float>()));
      } else {
        LACE_TRY(M->apply_d(wj, E.zcol(j)));
        LACE_TRY(Aop->apply_dd(static_cast<const double*>(E.zcol(j)), E.w.template as<double>()));
      }
      LACE_TRY(cgs_step<K>(E, j, j == iters - 1, rtol));
    }
    return hh_finish<K, K>(E, z);
  }
};

// ------------------------------------------------------------ MATLAB log
// dots[i] = sum_r Q[r, i] * Q[r, j] for i <= j; one workgroup per column,
// fixed reduction order.
__global__ __launch_bounds__(256) void ortho_dots_kernel(int n, const double* __restrict__ Q, size_t ld,
                                                         int j, double* __restrict__ dots) {
  __shared__ double red[256];
  const double* a = Q + (size_t)blockIdx.x * ld;
  const double* b = Q + (size_t)j * ld;
  double s = 0.0;
  for (int r = threadIdx.x; r < n; r += 256) s = fma(a[r], b[r], s);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) dots[blockIdx.x] = red[0];
}

// MATLAB-syntax progress lines of the drop-in FGMRES entry points, in the
// four forms of the reference: src/fgmres.cpp (MGS), src/fgmres_restart.cpp
// (MGS, restarted), src/fgmres_householder.cpp and
// src/fgmres_householder_restart.cpp.  The single-cycle forms number their
// search directions across the cycles of 63 that a larger search_max runs
// as.  With monitorOrthog the orthogonality error ||I - Q_k^T Q_k||_inf of
// the cycle's first k basis vectors (src/orthogonality_error.cpp:46-82;
// Householder: the preconditioned vectors q_j, src/
// fgmres_householder_restart.cpp:628-656) precedes each search line.
struct MatlabLog {
  bool on = false, mgs = false, restarted = false, orthog = false;
  int choice = 0;  // user_csrtrsv_choice
  int offset = 0;  // search index before the current cycle (single-cycle forms)
  DevBuf Q, dots;  // Householder basis copies, dot products
  std::vector<double> G, hd;
  int k = 0;

  const char* tag() const { return choice == 0 ? "mkltrsv" : "partrsv"; }
  const char* fam() const { return mgs ? "" : "Householders_"; }
  int idx(int j) const { return restarted ? j + 1 : offset + j + 1; }

  void cycle_begin(int restart, int m, double rnorm, double rtol_param, double rtol_abs) {
    if (!on) return;
    k = 0;
    offset = restart * m;
    if (restarted) {
      printf("restart = %d\n", restart);           // src/fgmres_householder_restart.cpp:196
      printf("rnorm2 = %e; rtol = %e;\n", rnorm, rtol_abs);  // :231
    } else if (restart == 0) {
      printf("rnorm2 = %e; tol = %e; rtol = %e;\n", rnorm, rtol_param, rtol_param * rnorm);  // src/fgmres.cpp:178
    }
  }
  // src/fgmres.cpp:234,240: sweep counts of the two ParCSRTRSV solves.  The
  // solves here are exact, which is what the in-place natural-order sweep
  // of src/trisolve.cpp:281-292 reaches after one substituting sweep and one
  // confirming sweep.
  void trsv_lines(int j) const {
    if (!on || choice != 1) return;
    printf("ParCSRTRSV_L(%d) = %d;\n", idx(j), 2);
    printf("ParCSRTRSV_U(%d) = %d;\n", idx(j), 2);
  }
  void reorth_line(int j) const {  // src/fgmres.cpp:316, src/fgmres_restart.cpp:338
    if (!on) return;
    if (restarted) printf("Reorthogonalize %d\n", j);
    else printf("Reorthogonalize(%d) = 1;\n", offset + j);
  }
  void breakdown_line() const {  // src/fgmres.cpp:355
    if (on) printf("%%\t******* happy breakdown **********\n");
  }
  // Column j of the basis Qb (leading dimension ld) joins the monitored set.
  int ortho_line(int n, const double* Qb, size_t ld, int j) {
    if (!on || !orthog) return DEV_SUCCESS;
    if (!dots.p) {
      LACE_TRY(dots.alloc(sizeof(double) * 64));
      hd.resize(64);
      G.resize(64 * 64);
    }
    LACE_REQUIRE(j < 64, DEV_ERR_NOT_SUPPORTED, "orthogonality monitor: column %d beyond 63", j);
    ortho_dots_kernel<<<j + 1, 256, 0, stream()>>>(n, Qb, ld, j, dots.as<double>());
    LACE_HIP(hipGetLastError());
    LACE_HIP(hipMemcpyAsync(hd.data(), dots.p, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, stream()));
    LACE_HIP(hipStreamSynchronize(stream()));
    for (int i = 0; i <= j; i++) G[i * 64 + j] = G[j * 64 + i] = hd[i];
    k = j + 1;
    // infinity norm: largest absolute row sum of I - G
    double err = 0.0;
    for (int i = 0; i < k; i++) {
      double s = 0.0;
      for (int l = 0; l < k; l++) s += std::fabs((i == l ? 1.0 : 0.0) - G[i * 64 + l]);
      err = std::max(err, s);
    }
    printf("FGMRES_%s%s_ortherr(%d) = %.16e;\n", fam(), tag(), idx(j), err);
    return DEV_SUCCESS;
  }
  // Householder: the preconditioned vector q_j is copied aside first.
  int ortho_line_copy(int n, const double* q, int j) {
    if (!on || !orthog) return DEV_SUCCESS;
    const size_t ld = ((size_t)n + 63) / 64 * 64;
    if (!Q.p) LACE_TRY(Q.alloc(sizeof(double) * ld * 64));
    LACE_REQUIRE(j < 64, DEV_ERR_NOT_SUPPORTED, "orthogonality monitor: column %d beyond 63", j);
    LACE_HIP(hipMemcpyAsync(Q.as<double>() + (size_t)j * ld, q, sizeof(double) * n, hipMemcpyDeviceToDevice,
                            stream()));
    return ortho_line(n, Q.as<double>(), ld, j);
  }
  void search_line(int j, double res) const {  // src/fgmres.cpp:430,433; ..._restart.cpp:707,710
    if (!on) return;
    printf("FGMRES_%s%s_search(%d) = %.16e;\n", fam(), tag(), idx(j), res);
    fflush(stdout);
  }
};

// ------------------------------------------------------------ solver
struct Solver {
  lace_solver_opts o;
  int n = 0, nnz = 0;
  DevCSR A64;          // A in FP64 (Arnoldi / true-residual products)
  DevCSR A32;          // FP32 copy of A for the inner solver
  DevCSR AT64, AT32;   // A^T, built on the first transposed solve
  bool have_T = false;
  std::vector<int> h_row, h_col;  // pattern of A (for A^T)
  std::vector<int> t_from;        // A^T position -> A position
  std::vector<double> h_val;      // current operator values
  // block (BCSR) operator and block ILU factors
  bool block = false;
  DevBCSR B64, B32;
  BILU* BP = nullptr;
  std::unique_ptr<PrecBILU> bilu;
  // operator wrappers; op64 is the FP64 operator of the current solve
  std::unique_ptr<LinOp> opA64, opA32, opAT64, opAT32;
  const LinOp* op64 = nullptr;
  ParILU* P = nullptr;
  std::unique_ptr<PrecILU> ilu;
  std::unique_ptr<PrecIdentity> ident;
  std::unique_ptr<PrecInner<float>> inner;
  std::unique_ptr<PrecInner<double>> inner64;  // inner_fp64
  Precond* M = nullptr;
  HHEngine<double> E;
  std::unique_ptr<BlockEngine> beng;  // 32-right-hand-side solver, on first use
  MatlabLog mlog;  // progress lines of the drop-in FGMRES entry points
  DevBuf rbuf, partial, scal;
  int nbn = 0, rpb = 4096, pstride = 4;
  double* host_scal = nullptr;
  ~Solver() {
    delete P;
    delete BP;
    if (host_scal) (void)hipHostFree(host_scal);
  }
};

// Host CSR view over a solver's pattern with the given values.
static data_d_matrix csr_view(int n, std::vector<int>& row, std::vector<int>& col, double* val) {
  data_d_matrix M;
  memset(&M, 0, sizeof(M));
  M.storage_type = Magma_CSR;
  M.num_rows = M.num_cols = n;
  M.nnz = row[n];
  M.row = row.data();
  M.col = col.data();
  M.val = val;
  return M;
}

// Build A^T (CSR) on the device; keeps the position map for value updates.
static int ensure_transpose(Solver* S) {
  if (S->have_T) return DEV_SUCCESS;
  const int n = S->n, nnz = S->nnz;
  std::vector<int> tr(n + 1, 0), tc(nnz);
  S->t_from.assign(nnz, 0);
  for (int k = 0; k < nnz; k++) tr[S->h_col[k] + 1]++;
  for (int j = 0; j < n; j++) tr[j + 1] += tr[j];
  std::vector<int> fill(tr.begin(), tr.end() - 1);
  for (int i = 0; i < n; i++)
    for (int k = S->h_row[i]; k < S->h_row[i + 1]; k++) {
      const int q = fill[S->h_col[k]]++;
      tc[q] = i;
      S->t_from[q] = k;
    }
  std::vector<double> tv(nnz);
  for (int q = 0; q < nnz; q++) tv[q] = S->h_val[S->t_from[q]];
  data_d_matrix T = csr_view(n, tr, tc, tv.data());
  LACE_TRY(S->AT64.create(&T, LACE_F64));
  S->opAT64 = std::make_unique<CsrOp>(&S->AT64);
  if (S->inner) {
    LACE_TRY(S->AT32.create(&T, LACE_F32));
    S->opAT32 = std::make_unique<CsrOp>(&S->AT32);
  }
  S->have_T = true;
  return DEV_SUCCESS;
}

// Select the operator (A or A^T) and the matching preconditioner direction.
static int select_direction(Solver* S, int trans) {
  LACE_REQUIRE(!(trans && S->block), DEV_ERR_NOT_SUPPORTED,
               "solver: transposed solves are available for CSR operators only");
  if (trans) LACE_TRY(ensure_transpose(S));
  S->op64 = trans ? S->opAT64.get() : S->opA64.get();
  if (S->ilu) S->ilu->trans = trans ? 1 : 0;
  if (S->inner) S->inner->Aop = trans ? S->opAT32.get() : S->opA32.get();
  if (S->inner64) S->inner64->Aop = trans ? S->opAT64.get() : S->opA64.get();
  return DEV_SUCCESS;
}

static int solver_create(const data_d_matrix* A, const lace_solver_opts* o, Solver** out) {
  LACE_TRY(ensure_init());
  LACE_REQUIRE(A && A->row && A->col && A->val && A->num_rows == A->num_cols, DEV_ERR_ILLEGAL_VALUE,
               "solver: A must be a square CSR or BCSR matrix");
  auto S = std::make_unique<Solver>();
  S->o = *o;
  S->block = A->storage_type == Magma_BCSR;
  // nested FP64: the inner GMRES on A itself with FP64 factors
  const bool in64 = o->inner_iters > 0 && o->inner_fp64 != 0;
  LACE_REQUIRE(!in64 || (!S->block && o->prec_precond == LACE_F64), DEV_ERR_NOT_SUPPORTED,
               "solver: an FP64 inner GMRES needs a CSR operator and FP64 factors");
  if (S->block) {
    // block operator: FP64 BCSR for the outer products, block ILU factors
    LACE_TRY(S->B64.create(A, LACE_F64));
    S->n = S->B64.n();
    S->nnz = S->B64.nblocks * S->B64.bs * S->B64.bs;
    S->opA64 = std::make_unique<BcsrOp>(&S->B64);
    if (o->inner_iters > 0) {
      LACE_TRY(S->B32.create(A, LACE_F32));
      S->opA32 = std::make_unique<BcsrOp>(&S->B32);
    }
  } else {
    S->n = A->num_rows;
    S->nnz = A->row[S->n];
    S->h_row.assign(A->row, A->row + S->n + 1);
    S->h_col.assign(A->col, A->col + S->nnz);
    S->h_val.assign(A->val, A->val + S->nnz);
    LACE_TRY(S->A64.create(A, LACE_F64));
    S->opA64 = std::make_unique<CsrOp>(&S->A64);
    if (o->inner_iters > 0 && !in64) {
      LACE_TRY(S->A32.create(A, LACE_F32));
      S->opA32 = std::make_unique<CsrOp>(&S->A32);
    }
  }
  if (o->trsv == LACE_TRSV_NONE && o->inner_iters <= 0) {
    S->ident = std::make_unique<PrecIdentity>(S->n);
    S->M = S->ident.get();
  } else {
    // the FP32 inner solver iterates in FP32, so its factors are at most FP32
    lace_prec_t fp = o->prec_precond;
    if (o->inner_iters > 0 && !in64 && fp == LACE_F64) fp = LACE_F32;
    Precond* base = nullptr;
    if (S->block) {
      S->BP = new BILU();
      LACE_TRY(S->BP->create(A, fp));
      S->bilu = std::make_unique<PrecBILU>(S->BP, o->trsv, o->jacobi_sweeps);
      base = S->bilu.get();
    } else {
      LACE_REQUIRE(fp == LACE_F64 || fp == LACE_F32 || fp == LACE_BF16 || fp == LACE_F16, DEV_ERR_NOT_SUPPORTED,
                   "solver: CSR factor storage %d not supported", (int)fp);
      LACE_TRY(parilu_create(A, fp, &S->P));
      S->ilu = std::make_unique<PrecILU>(S->P, o->trsv, o->jacobi_sweeps, 0);
      base = S->ilu.get();
    }
    S->M = base;
    if (in64) {
      S->inner64 = std::make_unique<PrecInner<double>>(S->opA64.get(), base, S->n, o->inner_iters, 0.0);
      LACE_TRY(S->inner64->init());
      S->M = S->inner64.get();
    } else if (o->inner_iters > 0) {
      S->inner = std::make_unique<PrecInner<float>>(S->opA32.get(), base, S->n, o->inner_iters, 0.0);
      LACE_TRY(S->inner->init());
      S->M = S->inner.get();
    }
  }
  LACE_TRY(select_direction(S.get(), o->transpose));
  LACE_TRY(S->E.init(S->n, o->search_max, S->M->z_is_float()));
  LACE_TRY(S->rbuf.alloc(sizeof(double) * S->n));
  S->nbn = ceil_div(S->n, S->rpb);
  LACE_TRY(S->partial.alloc(sizeof(double) * S->nbn * S->pstride));
  LACE_TRY(S->scal.alloc(sizeof(double) * 4));
  LACE_HIP(hipHostMalloc((void**)&S->host_scal, sizeof(double) * 4, hipHostMallocDefault));
  LACE_HIP(hipStreamSynchronize(stream()));
  *out = S.release();
  return DEV_SUCCESS;
}

// New values for the operator on A's pattern: updates A, its FP32 copy, A^T
// and the ParILU's own copy of A.  The factors in use stay those of the last
// setup; the next setup factors the new values.
static int solver_set_values(Solver* S, const double* vals) {
  LACE_REQUIRE(!S->block, DEV_ERR_NOT_SUPPORTED, "solver: value updates are for CSR operators");
  LACE_REQUIRE(vals, DEV_ERR_INVALID_PTR, "solver: null values");
  S->h_val.assign(vals, vals + S->nnz);
  // host staging buffers live until the synchronisation below
  std::vector<float> f32;
  std::vector<double> tv;
  std::vector<float> tf32;
  LACE_HIP(hipMemcpyAsync(S->A64.val.p, S->h_val.data(), sizeof(double) * S->nnz, hipMemcpyHostToDevice,
                          stream()));
  if (S->inner) {
    f32.assign(vals, vals + S->nnz);
    LACE_HIP(hipMemcpyAsync(S->A32.val.p, f32.data(), sizeof(float) * S->nnz, hipMemcpyHostToDevice, stream()));
  }
  if (S->have_T) {
    tv.resize(S->nnz);
    for (int q = 0; q < S->nnz; q++) tv[q] = vals[S->t_from[q]];
    LACE_HIP(hipMemcpyAsync(S->AT64.val.p, tv.data(), sizeof(double) * S->nnz, hipMemcpyHostToDevice, stream()));
    if (S->inner) {
      tf32.assign(tv.begin(), tv.end());
      LACE_HIP(hipMemcpyAsync(S->AT32.val.p, tf32.data(), sizeof(float) * S->nnz, hipMemcpyHostToDevice,
                              stream()));
    }
  }
  LACE_TRY(S->A64.refresh_ell());
  if (S->inner) LACE_TRY(S->A32.refresh_ell());
  if (S->have_T) {
    LACE_TRY(S->AT64.refresh_ell());
    if (S->inner) LACE_TRY(S->AT32.refresh_ell());
  }
  if (S->P) LACE_TRY(S->P->set_values(vals));
  LACE_HIP(hipStreamSynchronize(stream()));
  return DEV_SUCCESS;
}

static int solver_setup(Solver* S) {
  if (S->BP) {
    TimedRegion tr("setup");
    LACE_TRY(S->BP->reset());
    for (int s = 0; s < S->o.parilu_sweeps; s++) LACE_TRY(S->BP->sweep(false));
    return S->BP->refresh();
  }
  if (!S->P) return DEV_SUCCESS;
  TimedRegion tr("setup");
  LACE_TRY(S->P->reset());
  for (int s = 0; s < S->o.parilu_sweeps; s++) LACE_TRY(S->P->sweep(false));
  S->P->csr_dirty = true;
  LACE_TRY(S->P->refresh_csr());
  return DEV_SUCCESS;
}

static int norm2(Solver* S, const double* x, double* out) {
  LACE_TRY(norm2_sq_partials(S->n, S->rpb, x, S->partial.as<double>(), S->pstride, S->nbn));
  LACE_TRY(sum_slot(S->nbn, S->pstride, S->partial.as<double>(), S->scal.as<double>()));
  LACE_HIP(hipMemcpyAsync(S->host_scal, S->scal.p, sizeof(double), hipMemcpyDeviceToHost, stream()));
  LACE_HIP(hipStreamSynchronize(stream()));
  *out = std::sqrt(S->host_scal[0]);
  return DEV_SUCCESS;
}

// One solve to rtol (relative to ||b - A x0||) with restarts.
// tol_abs >= 0 selects an absolute tolerance instead.
static int solver_solve(Solver* S, const double* b, double* x, lace_solver_stats* st, double tol_abs) {
  auto t0 = std::chrono::steady_clock::now();
  HHEngine<double>& E = S->E;
  const int m = S->o.search_max;
  lace_solver_stats local;
  memset(&local, 0, sizeof(local));
  bool converged = false;
  double rtol_abs = 0.0;
  for (int restart = 0; restart < S->o.restart_max && !converged; restart++) {
    LACE_TRY(S->op64->residual_dd(x, b, E.r.as<double>()));
    if (restart == 0) {
      if (tol_abs >= 0.0)
        LACE_TRY(set_misc_d(E.small.as<double>(), E.lay, MS_RTOL, tol_abs));
      LACE_TRY(hh_start<double>(E, tol_abs >= 0.0 ? 0 : 1, S->o.rtol));
    } else {
      LACE_TRY(hh_start<double>(E, 0, 0.0));
    }
    LACE_TRY(read_misc<double>(E));
    const double rnorm = E.host_misc[MS_RNORM];
    if (restart == 0) {
      local.initial_residual = rnorm;
      rtol_abs = E.host_misc[MS_RTOL];
    }
    local.final_residual_est = rnorm;
    S->mlog.cycle_begin(restart, m, rnorm, S->o.rtol, rtol_abs);
    if (!std::isfinite(rnorm)) break;
    if (E.host_misc[MS_DONE] != 0.0) {
      converged = rnorm < rtol_abs || rnorm == 0.0;
      break;
    }
    for (int j = 0; j < m; j++) {
      LACE_TRY(S->M->apply_d(E.q.as<double>(), E.zcol(j)));
      S->mlog.trsv_lines(j);
      LACE_TRY(S->mlog.ortho_line_copy(S->n, E.q.as<double>(), j));
      if (E.z_float)
        LACE_TRY(S->op64->apply_fd(static_cast<const float*>(E.zcol(j)), E.w.as<double>()));
      else
        LACE_TRY(S->op64->apply_dd(static_cast<const double*>(E.zcol(j)), E.w.as<double>()));
      LACE_TRY(hh_step<double>(E, j, j < m - 1));
      LACE_TRY(read_misc<double>(E));
      S->mlog.search_line(j, E.host_misc[MS_RES]);
      if (E.host_misc[MS_DONE] != 0.0) break;
    }
    LACE_TRY((hh_finish<double, double>(E, x)));
    LACE_TRY(read_misc<double>(E));
    const int nc = (int)E.host_misc[MS_NCOLS];
    local.iterations += nc;
    local.restarts = restart;
    local.final_residual_est = E.host_misc[MS_RES];
    if (E.host_misc[MS_BREAK] != 0.0) break;
    converged = E.host_misc[MS_RES] < rtol_abs;
  }
  if (S->inner || S->inner64) local.inner_iterations = local.iterations * S->o.inner_iters;
  // true residual in FP64 (against the operator in use)
  LACE_TRY(S->op64->residual_dd(x, b, S->rbuf.as<double>()));
  LACE_TRY(norm2(S, S->rbuf.as<double>(), &local.true_residual));
  double bn = 0.0;
  LACE_TRY(norm2(S, b, &bn));
  local.rel_true_residual = bn > 0.0 ? local.true_residual / bn : local.true_residual;
  local.converged = converged ? 1 : 0;
  auto t1 = std::chrono::steady_clock::now();
  local.solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (st) *st = local;
  return DEV_SUCCESS;
}

// dtheta[k] = -lambda[i] * x[col[k]] for the CSR entries k of row i.
__global__ void grad_gather_kernel(int n, const int* __restrict__ row, const int* __restrict__ col,
                                   const double* __restrict__ lambda, const double* __restrict__ x,
                                   double* __restrict__ dtheta) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double li = lambda[i];
  for (int k = row[i]; k < row[i + 1]; k++) dtheta[k] = -li * x[col[k]];
}

// Block form over 32 right-hand sides: dtheta[k] = -sum_c Lam[i][c] X[j][c]
// for position k = (i, j); 32 lanes per row, lane c one right-hand side,
// the sum over c in a fixed butterfly order.
__global__ __launch_bounds__(256) void grad_gather_block_kernel(int n, const int* __restrict__ row,
                                                                const int* __restrict__ col,
                                                                const double* __restrict__ Lam,
                                                                const double* __restrict__ X,
                                                                double* __restrict__ dtheta) {
  const int i = xcd_block(blockIdx.x, gridDim.x) * 8 + (threadIdx.x >> 5);
  const int c = threadIdx.x & 31;
  if (i >= n) return;
  const double li = Lam[(size_t)i * 32 + c];
  for (int k = row[i]; k < row[i + 1]; k++) {
    const double s = group_sum<32>(li * X[(size_t)col[k] * 32 + c]);
    if (c == 0) dtheta[k] = -s;
  }
}


// ------------------------------------------------------------ MGS FGMRES
// src/fgmres.cpp:64-517 and src/fgmres_restart.cpp: modified Gram-Schmidt
// Arnoldi with the Brown/Hindmarsh reorthogonalisation test (:314-338),
// Givens rotations in the form of :385-411, convergence on |g_{j+1}|.
// Vector passes run on the device; the scalar recurrences run in one-thread
// kernels on a small device state, so the only host reads are the
// per-step convergence/reorthogonalisation flags.
struct MgsState {  // device doubles
  int m;
  __host__ __device__ int H() const { return 0; }                        // (m+1) x m
  __host__ __device__ int cs() const { return (m + 1) * m; }
  __host__ __device__ int sn() const { return cs() + m; }
  __host__ __device__ int g() const { return sn() + m; }
  __host__ __device__ int y() const { return g() + m + 1; }
  __host__ __device__ int sc() const { return y() + m + 1; }  // scratch scalars
  __host__ __device__ int total() const { return sc() + 16; }
};
enum { MG_H = 0, MG_NORMAV = 1, MG_NORM2 = 2, MG_RES = 3, MG_DONE = 4, MG_REORTH = 5,
       MG_NC = 6, MG_RTOL = 7, MG_RNORM = 8, MG_ALPHA = 9 };

__global__ __launch_bounds__(256) void dot_partial_kernel(int n, int rpb, const double* __restrict__ a,
                                                          const double* __restrict__ b,
                                                          double* __restrict__ partial) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int r0 = blk * rpb, r1 = min(r0 + rpb, n);
  double v = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += 256) v += a[i] * b[i];
  __shared__ double sm[4];
  v = block_sum<256>(v, sm);
  if (threadIdx.x == 0) partial[blk] = v;
}

__global__ __launch_bounds__(256) void sum_to_slot_kernel(int nb, const double* partial, double* dst,
                                                          int take_sqrt) {
  __shared__ double sm[4];
  double v = 0.0;
  for (int t = threadIdx.x; t < nb; t += 256) v += partial[t];
  v = block_sum<256>(v, sm);
  if (threadIdx.x == 0) *dst = take_sqrt ? sqrt(v) : v;
}

// y <- y - (*alpha) * x   or  y <- y + (*alpha) * x (sign = -1 / +1)
__global__ void axpy_dev_kernel(int n, const double* __restrict__ alpha, double sign,
                                const double* __restrict__ x, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = y[i] + sign * (*alpha) * x[i];
}

__global__ void store_scalar_kernel(double* p, double v) { *p = v; }

__global__ void h_accum_kernel(double* h, const double* v, int set) {
  *h = set ? *v : *h + *v;
}

__global__ void scale_into_kernel(int n, const double* __restrict__ u, const double* __restrict__ h,
                                  const double* __restrict__ flag, double* __restrict__ v) {
  if (*flag == 0.0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = u[i] / (*h);
}

// After the MGS pass of step j: reorthogonalisation decision, breakdown
// test, Givens rotations, convergence.  sc[MG_H] holds the last dot.
__global__ void mgs_step_small_kernel(double* st, MgsState L, int j, int last, int stage) {
  if (threadIdx.x != 0) return;
  double* H = st + L.H();
  double* sc = st + L.sc();
  const int ldH = L.m + 1;
  if (stage == 0) {  // after the first norm of u: reorthogonalisation test
    const double normav = sc[MG_NORMAV], normav2 = sc[MG_NORM2];
    H[j + 1 + j * ldH] = normav2;
    const double hr = (normav + 0.001 * normav2) - normav;
    sc[MG_REORTH] = (hr <= 4.9406564584124654e-324) ? 1.0 : 0.0;
    return;
  }
  if (stage == 1) {  // after a reorthogonalisation pass
    H[j + 1 + j * ldH] = sc[MG_NORM2];
    return;
  }
  // stage 2: breakdown flag, rotations, convergence
  double* cs = st + L.cs();
  double* sn = st + L.sn();
  double* g = st + L.g();
  const double hj1 = H[j + 1 + j * ldH];
  sc[MG_ALPHA] = fabs(hj1) > 2.220446049250313e-16 ? 1.0 : 0.0;
  for (int i = 0; i < j; i++) {
    const double delta = H[i + j * ldH];
    H[i + j * ldH] = cs[i] * delta - sn[i] * H[i + 1 + j * ldH];
    H[i + 1 + j * ldH] = sn[i] * delta + cs[i] * H[i + 1 + j * ldH];
  }
  const double a = H[j + j * ldH], bb = H[j + 1 + j * ldH];
  const double gamma = sqrt(a * a + bb * bb);
  if (gamma > 0.0) {
    cs[j] = a / gamma;
    sn[j] = -bb / gamma;
    H[j + j * ldH] = cs[j] * a - sn[j] * bb;
    H[j + 1 + j * ldH] = 0.0;
    const double delta = g[j];
    g[j] = cs[j] * delta - sn[j] * g[j + 1];
    g[j + 1] = sn[j] * delta + cs[j] * g[j + 1];
  }
  const double res = fabs(g[j + 1]);
  sc[MG_RES] = res;
  sc[MG_NC] = j + 1;
  sc[MG_DONE] = (res < sc[MG_RTOL] || last || !isfinite(res)) ? 1.0 : 0.0;
}

__global__ void mgs_lsq_kernel(double* st, MgsState L) {
  if (threadIdx.x != 0) return;
  const double* H = st + L.H();
  const double* g = st + L.g();
  double* y = st + L.y();
  const int ldH = L.m + 1, nc = (int)st[L.sc() + MG_NC];
  for (int i = 0; i < nc; i++) y[i] = g[i] / H[i + i * ldH];
  for (int jj = nc - 1; jj > 0; jj--)
    for (int i = jj - 1; i > -1; i--) y[i] = y[i] - H[i + jj * ldH] * y[jj] / H[i + i * ldH];
}

__global__ void mgs_start_kernel(double* st, MgsState L, int set_rtol, double rtol) {
  if (threadIdx.x != 0) return;
  double* sc = st + L.sc();
  const double rn = sc[MG_RNORM];
  if (set_rtol) sc[MG_RTOL] = rtol * rn;
  for (int t = 0; t < L.total() - 16; t++) st[t] = 0.0;
  st[L.g()] = rn;
  sc[MG_DONE] = 0.0;
}

__global__ void mgs_xupdate_kernel(int n, int ld, const double* __restrict__ Z, const double* st,
                                   MgsState L, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int nc = (int)st[L.sc() + MG_NC];
  const double* y = st + L.y();
  double t = 0.0;
  for (int k = 0; k < nc; k++) t = t + Z[(size_t)k * ld + i] * y[k];
  x[i] = x[i] + t;
}

static int solver_solve_mgs(Solver* S, const double* b, double* x, lace_solver_stats* out,
                            double tol_abs, int reorth) {
  auto t0 = std::chrono::steady_clock::now();
  const int n = S->n, m = S->o.search_max, ld = ((n + 63) / 64) * 64;
  const int rpb = 4096, nb = ceil_div(n, rpb), nbv = ceil_div(n, 256);
  MgsState L{m};
  DevBuf V, Z, u, st, part;
  LACE_TRY(V.alloc(sizeof(double) * (size_t)ld * (m + 1)));
  LACE_TRY(Z.alloc(sizeof(double) * (size_t)ld * m));
  LACE_TRY(u.alloc(sizeof(double) * ld));
  LACE_TRY(st.alloc(sizeof(double) * L.total()));
  LACE_TRY(part.alloc(sizeof(double) * nb));
  LACE_HIP(hipMemsetAsync(st.p, 0, sizeof(double) * L.total(), stream()));
  double* sd = st.as<double>();
  double* sc = sd + L.sc();
  double hostsc[16];
  auto Vc = [&](int k) { return V.as<double>() + (size_t)k * ld; };
  auto Zc = [&](int k) { return Z.as<double>() + (size_t)k * ld; };
  auto nrm_into = [&](const double* a, double* dst) -> int {
    dot_partial_kernel<<<nb, 256, 0, stream()>>>(n, rpb, a, a, part.as<double>());
    sum_to_slot_kernel<<<1, 256, 0, stream()>>>(nb, part.as<double>(), dst, 1);
    LACE_HIP(hipGetLastError());
    return DEV_SUCCESS;
  };
  auto read_sc = [&]() -> int {
    LACE_HIP(hipMemcpyAsync(hostsc, sc, sizeof(hostsc), hipMemcpyDeviceToHost, stream()));
    LACE_HIP(hipStreamSynchronize(stream()));
    return DEV_SUCCESS;
  };
  lace_solver_stats local;
  memset(&local, 0, sizeof(local));
  bool converged = false;
  double rtol_abs = 0.0;
  if (tol_abs >= 0.0) {
    store_scalar_kernel<<<1, 1, 0, stream()>>>(sc + MG_RTOL, tol_abs);
    LACE_HIP(hipGetLastError());
  }
  for (int restart = 0; restart < S->o.restart_max && !converged; restart++) {
    double* r = Vc(0);
    LACE_TRY(S->op64->residual_dd(x, b, r));
    LACE_TRY(nrm_into(r, sc + MG_RNORM));
    mgs_start_kernel<<<1, 64, 0, stream()>>>(sd, L, restart == 0 && tol_abs < 0.0, S->o.rtol);
    LACE_TRY(read_sc());
    if (restart == 0) {
      local.initial_residual = hostsc[MG_RNORM];
      rtol_abs = hostsc[MG_RTOL];
    }
    S->mlog.cycle_begin(restart, m, hostsc[MG_RNORM], S->o.rtol, rtol_abs);
    if (!(hostsc[MG_RNORM] >= rtol_abs) || hostsc[MG_RNORM] == 0.0) {
      converged = std::isfinite(hostsc[MG_RNORM]);
      local.final_residual_est = hostsc[MG_RNORM];
      break;
    }
    // v_0 = r / ||r||
    scale_into_kernel<<<nbv, 256, 0, stream()>>>(n, r, sc + MG_RNORM, sc + MG_RNORM, Vc(0));
    for (int j = 0; j < m; j++) {
      LACE_TRY(S->M->apply_d(Vc(j), Zc(j)));
      S->mlog.trsv_lines(j);
      LACE_TRY(S->op64->apply_dd(Zc(j), u.as<double>()));
      LACE_TRY(nrm_into(u.as<double>(), sc + MG_NORMAV));
      const int passes = 1;
      for (int pass = 0; pass < passes + 1; pass++) {
        for (int k = 0; k <= j; k++) {
          dot_partial_kernel<<<nb, 256, 0, stream()>>>(n, rpb, Vc(k), u.as<double>(), part.as<double>());
          sum_to_slot_kernel<<<1, 256, 0, stream()>>>(nb, part.as<double>(), sc + MG_H, 0);
          // H[k, j] = dot (pass 0); += dot (reorthogonalisation pass, :327)
          h_accum_kernel<<<1, 1, 0, stream()>>>(sd + L.H() + k + j * (m + 1), sc + MG_H, pass == 0);
          axpy_dev_kernel<<<nbv, 256, 0, stream()>>>(n, sc + MG_H, -1.0, Vc(k), u.as<double>());
        }
        LACE_TRY(nrm_into(u.as<double>(), sc + MG_NORM2));
        mgs_step_small_kernel<<<1, 64, 0, stream()>>>(sd, L, j, 0, pass == 0 ? 0 : 1);
        if (pass == 0) {
          LACE_TRY(read_sc());
          const bool again = (reorth == 0 && hostsc[MG_REORTH] != 0.0) || reorth == 2;
          if (!again) break;
          S->mlog.reorth_line(j);
        }
      }
      mgs_step_small_kernel<<<1, 64, 0, stream()>>>(sd, L, j, j == m - 1, 2);
      // v_{j+1} = u / h_{j+1,j} unless breakdown (:341-356); the H entry was
      // saved before rotation in sc[MG_NORM2]
      scale_into_kernel<<<nbv, 256, 0, stream()>>>(n, u.as<double>(), sc + MG_NORM2, sc + MG_ALPHA,
                                                   Vc(j + 1));
      LACE_HIP(hipGetLastError());
      LACE_TRY(read_sc());
      if (!(std::fabs(hostsc[MG_NORM2]) > std::numeric_limits<double>::epsilon())) S->mlog.breakdown_line();
      LACE_TRY(S->mlog.ortho_line(n, V.as<double>(), ld, j));
      S->mlog.search_line(j, hostsc[MG_RES]);
      if (hostsc[MG_DONE] != 0.0) break;
    }
    mgs_lsq_kernel<<<1, 64, 0, stream()>>>(sd, L);
    mgs_xupdate_kernel<<<nbv, 256, 0, stream()>>>(n, ld, Z.as<double>(), sd, L, x);
    LACE_HIP(hipGetLastError());
    LACE_TRY(read_sc());
    local.iterations += (int)hostsc[MG_NC];
    local.restarts = restart;
    local.final_residual_est = hostsc[MG_RES];
    if (!std::isfinite(hostsc[MG_RES])) break;
    converged = hostsc[MG_RES] < rtol_abs;
  }
  LACE_TRY(S->op64->residual_dd(x, b, S->rbuf.as<double>()));
  LACE_TRY(norm2(S, S->rbuf.as<double>(), &local.true_residual));
  double bn = 0.0;
  LACE_TRY(norm2(S, b, &bn));
  local.rel_true_residual = bn > 0.0 ? local.true_residual / bn : local.true_residual;
  local.converged = converged ? 1 : 0;
  local.solve_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (out) *out = local;
  return DEV_SUCCESS;
}

}  // namespace lace

using namespace lace;

extern "C" {

void lace_solver_default_opts(lace_solver_opts* o) {
  memset(o, 0, sizeof(*o));
  o->krylov = 0;
  o->search_max = 50;
  o->restart_max = 100;
  o->rtol = 1e-8;
  o->prec_precond = LACE_F64;
  o->trsv = LACE_TRSV_LEVEL;
  o->jacobi_sweeps = LACE_JACOBI_SWEEPS_DEFAULT;
  o->parilu_sweeps = 3;
  o->inner_iters = 0;
  o->transpose = 0;
}

int lace_solver_create(const data_d_matrix* A, const lace_solver_opts* o, lace_solver** out) {
  Solver* S = nullptr;
  LACE_TRY(solver_create(A, o, &S));
  *out = reinterpret_cast<lace_solver*>(S);
  return DEV_SUCCESS;
}

int lace_solver_destroy(lace_solver* S) {
  delete reinterpret_cast<Solver*>(S);
  return DEV_SUCCESS;
}

int lace_solver_setup(lace_solver* Sh) {
  LACE_TRY(ensure_init());
  return solver_setup(reinterpret_cast<Solver*>(Sh));
}

int lace_solver_solve(lace_solver* Sh, const double* b, double* x, lace_solver_stats* st) {
  LACE_TRY(ensure_init());
  auto* S = reinterpret_cast<Solver*>(Sh);
  LACE_TRY(select_direction(S, S->o.transpose));
  return solver_solve(S, b, x, st, -1.0);
}

int lace_solver_solve_trans(lace_solver* Sh, const double* b, double* x, int trans,
                            lace_solver_stats* st) {
  LACE_TRY(ensure_init());
  auto* S = reinterpret_cast<Solver*>(Sh);
  LACE_TRY(select_direction(S, trans));
  const int rc = solver_solve(S, b, x, st, -1.0);
  LACE_TRY(select_direction(S, S->o.transpose));
  return rc;
}

int lace_solver_set_values(lace_solver* Sh, const double* vals) {
  LACE_TRY(ensure_init());
  return solver_set_values(reinterpret_cast<Solver*>(Sh), vals);
}

int lace_solver_grad_values(lace_solver* Sh, const double* lambda, const double* x, double* dtheta) {
  LACE_TRY(ensure_init());
  auto* S = reinterpret_cast<Solver*>(Sh);
  LACE_REQUIRE(!S->block, DEV_ERR_NOT_SUPPORTED, "solver: value gradients are for CSR operators");
  grad_gather_kernel<<<ceil_div(S->n, 256), 256, 0, stream()>>>(
      S->n, S->A64.row.as<int>(), S->A64.col.as<int>(), lambda, x, dtheta);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

int lace_solver_solve_block(lace_solver* Sh, const double* B, double* X, int trans, lace_block_stats* st) {
  LACE_TRY(ensure_init());
  auto* S = reinterpret_cast<Solver*>(Sh);
  const int m = S->o.search_max;
  LACE_REQUIRE(!S->block && S->P, DEV_ERR_NOT_SUPPORTED, "block solve: CSR operator with ILU factors");
  LACE_REQUIRE(!S->inner64, DEV_ERR_NOT_SUPPORTED, "block solve: the inner GMRES runs in FP32 only");
  LACE_REQUIRE(S->o.inner_iters >= 0 && S->o.inner_iters <= kMaxInner, DEV_ERR_NOT_SUPPORTED,
               "block solve: inner iterations %d outside 0..%d", S->o.inner_iters, kMaxInner);
  LACE_REQUIRE(S->o.trsv == LACE_TRSV_JACOBI || S->o.trsv == LACE_TRSV_NONE, DEV_ERR_NOT_SUPPORTED,
               "block solve: Jacobi (or no) triangular solves");
  LACE_REQUIRE(m >= 1 && m + 1 <= kMaxBlocks, DEV_ERR_ILLEGAL_VALUE, "block solve: search_max %d", m);
  LACE_REQUIRE((long)S->n >= 32L * (m + 2), DEV_ERR_ILLEGAL_VALUE,
               "block solve: n = %d below 32 (search_max + 2)", S->n);
  if (trans) LACE_TRY(ensure_transpose(S));
  const bool zf = S->P->prec != LACE_F64;
  if (!S->beng || S->beng->m != m || S->beng->zf != zf) {
    S->beng.reset();
    auto E = std::make_unique<BlockEngine>();
    LACE_TRY(E->init(S->n, m, zf));
    S->beng = std::move(E);
  }
  BlockSolveArgs args;
  args.A = trans ? &S->AT64 : &S->A64;
  args.A32 = S->o.inner_iters > 0 ? (trans ? &S->AT32 : &S->A32) : nullptr;
  args.inner_iters = S->o.inner_iters;
  args.P = S->P;
  args.kind = S->o.trsv;
  args.sweeps = S->o.jacobi_sweeps;
  args.trans = trans ? 1 : 0;
  args.restart_max = S->o.restart_max;
  args.rtol = S->o.rtol;
  lace_block_stats local;
  LACE_TRY(block_fgmres(*S->beng, args, B, X, &local));
  if (st) *st = local;
  return DEV_SUCCESS;
}

int lace_solver_grad_values_block(lace_solver* Sh, const double* Lambda, const double* X, double* dtheta) {
  LACE_TRY(ensure_init());
  auto* S = reinterpret_cast<Solver*>(Sh);
  LACE_REQUIRE(!S->block, DEV_ERR_NOT_SUPPORTED, "solver: value gradients are for CSR operators");
  grad_gather_block_kernel<<<ceil_div(S->n, 8), 256, 0, stream()>>>(
      S->n, S->A64.row.as<int>(), S->A64.col.as<int>(), Lambda, X, dtheta);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

}  // extern "C"

// ------------------------------------------------------------ drop-ins
namespace {

// Shared body of data_fgmres_householder[_restart]: L as CSRL, U as CSRU
// (the test path converts the CSCU factor to CSR, testing/
// test_linearSolver.cpp:670-672).  The factors are used as given.
int fgmres_dropin_common(data_d_matrix* A, data_d_matrix* b, data_d_matrix* x0, data_d_matrix* L,
                         data_d_matrix* U, data_d_gmres_param* par, data_d_gmres_log* log,
                         bool restarted, bool mgs, bool x_in_place = false) {
  LACE_TRY(ensure_init());
  auto t0 = std::chrono::steady_clock::now();
  const int n = A->num_rows;
  lace_solver_opts o;
  lace_solver_default_opts(&o);
  o.search_max = par->search_max;
  o.restart_max = restarted ? par->restart_max : 1;
  o.rtol = par->rtol;
  const int choice = par->user_csrtrsv_choice;
  // 0: exact solves (mkl_dcsrtrsv in the reference); 1: ParCSRTRSV, which
  // with the reference's natural row order and one thread is exact forward/
  // backward substitution plus a confirming sweep (src/trisolve.cpp:281-292),
  // so it also maps to the exact solve; 2: none; 3: Jacobi-k.
  o.trsv = choice == 2 ? LACE_TRSV_NONE : (choice == 3 ? LACE_TRSV_JACOBI : LACE_TRSV_LEVEL);
  o.jacobi_sweeps = LACE_JACOBI_SWEEPS_DEFAULT;
  o.parilu_sweeps = 0;
  if (o.search_max > 63) {
    // the device Krylov engines hold at most 63 basis vectors per cycle:
    // the same budget of search directions as cycles of 63
    const long total = (long)o.search_max * std::max(1, o.restart_max);
    o.search_max = 63;
    o.restart_max = (int)((total + 62) / 63);
  }
  Solver* S = nullptr;
  LACE_TRY(solver_create(A, &o, &S));
  std::unique_ptr<Solver> hold(S);
  if (get_option("matlab_log", 1) != 0) {
    MatlabLog& ml = S->mlog;
    ml.on = true;
    ml.mgs = mgs;
    ml.restarted = restarted;
    ml.orthog = par->monitorOrthog == 1;
    ml.choice = choice;
    // src/fgmres.cpp:71, src/fgmres_restart.cpp:70,81,
    // src/fgmres_householder.cpp:70, src/fgmres_householder_restart.cpp:70,81
    const char* name = mgs ? "data_fgmres" : "data_fgmres_householder";
    printf("%% %s%s begin\n", name, restarted && !mgs ? "_restart" : "");
    if (restarted)
      printf("%% %s restart_max = %d search_max =%d\n", name, par->restart_max, par->search_max);
  }
  if (S->P) {
    // load the caller's factors: L (CSRL) into the L buffer, U (CSRU) into
    // the CSC buffer through the CSR->CSC slot map
    ParILU* P = S->P;
    LACE_REQUIRE(L->row && U->row && L->row[n] == P->nnzL && U->row[n] == P->nnzU,
                 DEV_ERR_ILLEGAL_VALUE, "FGMRES: L/U patterns do not match A's ILU(0) pattern");
    std::vector<double> uc(P->nnzU);
    // U CSR slot q of row i col j -> CSC slot
    std::vector<int> fill(P->h_ucp.begin(), P->h_ucp.end() - 1);
    for (int i = 0; i < n; i++)
      for (int k = U->row[i]; k < U->row[i + 1]; k++) uc[fill[U->col[k]]++] = U->val[k];
    LACE_HIP(hipMemcpyAsync(P->lval[0].p, L->val, sizeof(double) * P->nnzL, hipMemcpyHostToDevice, stream()));
    LACE_HIP(hipMemcpyAsync(P->uval[0].p, uc.data(), sizeof(double) * P->nnzU, hipMemcpyHostToDevice, stream()));
    LACE_HIP(hipStreamSynchronize(stream()));
    P->cur = 0;
    P->csr_dirty = true;
    LACE_TRY(P->refresh_csr());
  }
  DevBuf db, dx;
  LACE_TRY(upload(db, b->val, sizeof(double) * n));
  LACE_TRY(upload(dx, x0->val, sizeof(double) * n));
  lace_solver_stats st;
  const double tol_abs = par->tol_type == 1 ? -1.0 : par->rtol;
  if (mgs)
    LACE_TRY(solver_solve_mgs(S, db.as<double>(), dx.as<double>(), &st, tol_abs, par->reorth));
  else
    LACE_TRY(solver_solve(S, db.as<double>(), dx.as<double>(), &st, tol_abs));
  if (x_in_place) {
    LACE_TRY(download(x0->val, dx, sizeof(double) * n));
  } else {
    // x0 is replaced by the solution (src/fgmres.cpp:491-492)
    data_d_matrix xs;
    memset(&xs, 0, sizeof(xs));
    LACE_TRY(data_zvinit(&xs, n, 1, 0.0, x0->major ? x0->major : MagmaColMajor));
    LACE_TRY(download(xs.val, dx, sizeof(double) * n));
    data_zmfree(x0);
    *x0 = xs;
  }
  auto t1 = std::chrono::steady_clock::now();
  log->restarts = st.restarts;
  log->search_directions = st.iterations;
  log->solve_time = std::chrono::duration<double>(t1 - t0).count();
  log->initial_residual = st.initial_residual;
  log->final_residual = st.final_residual_est;
  return st.converged ? 0 : DEV_NOTCONVERGED;
}

}  // namespace

extern "C" {

int data_fgmres_householder(data_d_matrix* A, data_d_matrix* b, data_d_matrix* x0, data_d_matrix* L,
                            data_d_matrix* U, data_d_gmres_param* par, data_d_gmres_log* log) {
  return fgmres_dropin_common(A, b, x0, L, U, par, log, false, false);
}

int data_fgmres_householder_restart(data_d_matrix* A, data_d_matrix* b, data_d_matrix* x0,
                                    data_d_matrix* L, data_d_matrix* U, data_d_gmres_param* par,
                                    data_d_gmres_log* log) {
  return fgmres_dropin_common(A, b, x0, L, U, par, log, true, false);
}

}  // extern "C"

namespace {

int fgmres_mgs_dropin(data_d_matrix* A, data_d_matrix* b, data_d_matrix* x0, data_d_matrix* L,
                      data_d_matrix* U, data_d_gmres_param* par, data_d_gmres_log* log,
                      bool restarted) {
  return fgmres_dropin_common(A, b, x0, L, U, par, log, restarted, true);
}

}  // namespace

extern "C" {

int data_fgmres(data_d_matrix* A, data_d_matrix* b, data_d_matrix* x0, data_d_matrix* L,
                data_d_matrix* U, data_d_gmres_param* par, data_d_gmres_log* log) {
  return fgmres_mgs_dropin(A, b, x0, L, U, par, log, false);
}

int data_fgmres_restart(data_d_matrix* A, data_d_matrix* b, data_d_matrix* x0, data_d_matrix* L,
                        data_d_matrix* U, data_d_gmres_param* par, data_d_gmres_log* log) {
  return fgmres_mgs_dropin(A, b, x0, L, U, par, log, true);
}

}  // extern "C"

extern "C" {

// src/mkl_fgmres.cpp:26-: the reference's CPU baseline, MKL's RCI FGMRES
// (modified Gram-Schmidt Arnoldi) with ILU(0) factors from ParILU v0.2
// (parilu_reduction) and exact triangular solves on the merged LU
// (mkl_dcsrtrsv), ipar[14] = restart_max as the restart length, ipar[4] =
// search_max as the iteration cap, dpar[0]/dpar[1] = rtol relative
// (tol_type 1) or absolute (0).  Here the same method runs on the device:
// ParILU v0.2 sweeps, level-scheduled exact solves, MGS FGMRES with
// restart length restart_max and ceil(search_max / restart_max) cycles
// (the last cycle may run past the cap by less than one cycle);
// precondition = 0 drops the preconditioner.  The solution is written into
// x0->val in place, as the reference does; returns 0, or DEV_NOTCONVERGED.
int data_MKL_FGMRES(data_d_matrix* A, data_d_matrix* x0, data_d_matrix* rhs_vector,
                    data_d_gmres_param* par) {
  LACE_TRY(ensure_init());
  LACE_REQUIRE(A && x0 && rhs_vector && par && x0->val && rhs_vector->val, DEV_ERR_INVALID_PTR,
               "data_MKL_FGMRES: null argument");
  data_d_matrix L, Ucsc, U;
  memset(&L, 0, sizeof(L));
  memset(&Ucsc, 0, sizeof(Ucsc));
  memset(&U, 0, sizeof(U));
  data_PariLU_v0_2(A, &L, &Ucsc, par->parilu_reduction);
  LACE_REQUIRE(L.row && Ucsc.row, DEV_ERR, "data_MKL_FGMRES: ParILU failed: %s", get_error());
  int rc = data_zmconvert(Ucsc, &U, Magma_CSCU, Magma_CSRU);
  data_d_gmres_param p = *par;
  p.search_max = std::max(1, par->restart_max);
  p.restart_max = std::max(1, (par->search_max + p.search_max - 1) / p.search_max);
  p.user_csrtrsv_choice = par->precondition ? 0 : 2;
  data_d_gmres_log log;
  memset(&log, 0, sizeof(log));
  if (rc == DEV_SUCCESS)
    rc = fgmres_dropin_common(A, rhs_vector, x0, &L, &U, &p, &log, true, true, true);
  data_zmfree(&L);
  data_zmfree(&Ucsc);
  data_zmfree(&U);
  return rc;
}

}  // extern "C"
