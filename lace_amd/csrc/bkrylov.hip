// bkrylov.hip -- block Householder FGMRES for 32 right-hand sides.
//
// The single-vector solver follows Walker's Householder Arnoldi
// (src/fgmres_householder_restart.cpp:63-788): reflector from the residual,
// per step v_j = P_0..P_j e_j, z_j = M^{-1} v_j, w = A z_j, w <- P_j..P_0 w,
// a new reflector from w's tail, Givens on the Hessenberg column.  Here the
// same recurrence runs on blocks of 32 columns:
//   cycle start: R = B - A X; block reflector H_0 = I - Y_0 T_0 Y_0^T with
//                H_0 [R~_0; 0] = R (panel QR below), G = [R~_0; 0].
//   step k:      V_k = Q E_k with Q = H_0 ... H_k = I - Y T Y^T (compact WY),
//                Z_k = M^{-1} V_k, W = A Z_k, W <- Q^T W = W - Y T^T (Y^T W),
//                H_{k+1} from rows >= 32(k+1) of W: its R~ is the
//                subdiagonal block, rows < 32(k+1) of W the block column
//                above it; T grows by [[T, -T (Y^T Y_{k+1}) T_{k+1}], [0,
//                T_{k+1}]].  The block Hessenberg column is reduced by the
//                stored 64 x 32 Householder windows plus a new one, which
//                also updates G; the residual norm of right-hand side c is the
//                norm of column c of G's new last block.
//   cycle end:   back substitution for the block least-squares solution
//                Yls, X += Z Yls.
// Panels use CholeskyQR2 (Yamamoto et al. 2015) with a shifted first pass
// when the Gram matrix is numerically indefinite (Fukaya et al. 2020),
// and Householder reconstruction (Ballard et al. 2015).  The n x 32
// products are the MFMA kernels of tsgemm.hip; the host synchronises once
// per Cholesky (to read its status) and once per step (residual norms).
#include <chrono>
#include <cmath>
#include <cstring>

#include "bkrylov.hpp"

namespace lace {

BlockEngine::~BlockEngine() {
  if (host) (void)hipHostFree(host);
  if (host_status) (void)hipHostFree(host_status);
}

int BlockEngine::init(int n_, int m_, bool zf_) {
  n = n_;
  m = m_;
  zf = zf_;
  const size_t blk = (size_t)n * kNrhs;
  const size_t es = zf ? sizeof(float) : sizeof(double);
  Y.clear();
  Z.clear();
  for (int i = 0; i <= m; i++) {
    Y.emplace_back(new DevBuf());
    LACE_TRY(Y.back()->alloc(blk * sizeof(double)));
  }
  for (int i = 0; i < m; i++) {
    Z.emplace_back(new DevBuf());
    LACE_TRY(Z.back()->alloc(blk * es));
  }
  LACE_TRY(W.alloc(blk * sizeof(double)));
  LACE_TRY(Qa.alloc(blk * sizeof(double)));
  LACE_TRY(Qb.alloc(blk * sizeof(double)));
  LACE_TRY(R.alloc(blk * sizeof(double)));
  LACE_TRY(Vt.alloc(blk * es));
  LACE_TRY(t1.alloc(blk * es));
  LACE_TRY(t2.alloc(blk * es));
  LACE_TRY(tg.init(n));
  LACE_TRY(partial.alloc(sizeof(double) * sd_colnorms_partial_elems(n)));
  // small matrices
  const int M = (m + 1) * kNrhs;
  ldT = M;
  H.m = m;
  H.ldh = m * kNrhs;
  const size_t s32 = 1024;
  size_t off = 0;
  auto take = [&](size_t elems) {
    size_t o = off;
    off += (elems + 31) / 32 * 32;
    return o;
  };
  const size_t oT = take((size_t)M * M), oH = take((size_t)M * H.ldh), oVw = take((size_t)(m + 1) * 2048),
               oTw = take((size_t)(m + 1) * s32), oRd = take((size_t)(m + 1) * s32), oG = take((size_t)M * 32),
               oa = take((size_t)M * 32), oD = take((size_t)M * 32), oGk = take((size_t)M * 32),
               oGp = take((size_t)M * 32), oK = take((size_t)M * 32), otm = take((size_t)M * 32),
               oDn = take((size_t)M * 32), oG1 = take(s32), oRl = take(s32), oRli = take(s32),
               onR = take(s32), oRp0 = take(s32), oRp1 = take(s32), oI = take(s32), oL = take(s32),
               oTb = take(s32), oRt = take(s32), oDy = take(s32), oZ = take(s32), ores = take(32),
               orn = take(32), obn = take(32);
  LACE_TRY(small.alloc(off * sizeof(double)));
  double* b = small.as<double>();
  Tall = b + oT;
  H.Hm = b + oH;
  H.Vw = b + oVw;
  H.Tw = b + oTw;
  H.Rdinv = b + oRd;
  H.Gls = b + oG;
  a = b + oa;
  D = b + oD;
  Gk = b + oGk;
  Gp = b + oGp;
  Km = b + oK;
  tmpM = b + otm;
  Dneg = b + oDn;
  G1 = b + oG1;
  Rl = b + oRl;
  Rlinv = b + oRli;
  negRlinv = b + onR;
  Rp[0] = b + oRp0;
  Rp[1] = b + oRp1;
  Ident = b + oI;
  Ltop = b + oL;
  Tb = b + oTb;
  Rt = b + oRt;
  Dy = b + oDy;
  Zero = b + oZ;
  res = b + ores;
  rn = b + orn;
  bn = b + obn;
  LACE_HIP(hipMemsetAsync(small.p, 0, off * sizeof(double), stream()));
  LACE_TRY(sd_copy32(1.0, nullptr, Ident, 32));
  LACE_TRY(ints.alloc(sizeof(int) * 64));
  if (!host) LACE_HIP(hipHostMalloc((void**)&host, sizeof(double) * 96, hipHostMallocDefault));
  if (!host_status) LACE_HIP(hipHostMalloc((void**)&host_status, sizeof(int) * 16, hipHostMallocDefault));
  status = ints.as<int>();
  LACE_HIP(hipStreamSynchronize(stream()));
  return DEV_SUCCESS;
}

static BlockList one_block(const void* p) {
  BlockList L;
  L.nb = 1;
  L.p[0] = p;
  return L;
}

static int read_status(BlockEngine& E, int* st, unsigned* zero_mask = nullptr) {
  LACE_HIP(hipMemcpyAsync(E.host_status, E.status, 2 * sizeof(int), hipMemcpyDeviceToHost, stream()));
  LACE_HIP(hipStreamSynchronize(stream()));
  *st = E.host_status[0];
  if (zero_mask) *zero_mask = (unsigned)E.host_status[1];
  return DEV_SUCCESS;
}

// CholeskyQR passes on rows [r0, n) of W until two consecutive unshifted
// passes; then reconstruction.  On return E.Rt, E.Tb, E.Dy, E.Ltop hold the
// panel's factors and *qlast the last formed orthogonalised block (null
// when the panel is zero: *breakdown = 1, Rt = 0).  Exactly zero columns of
// a nonzero panel are filled (rows >= r0 of W change) and zeroed in Rt.
static int panel(BlockEngine& E, double* W, int r0, const double** qlast, int* passes, int* shifted,
                 int* breakdown) {
  TimedRegion tr("block_panel");
  *breakdown = 0;
  const long rows = (long)E.n - r0;
  LACE_TRY(E.tg.tn(one_block(W), W, r0, r0, E.G1));
  LACE_TRY(sd_chol32(E.G1, rows, E.Rl, E.Rlinv, E.status));
  int st = 0;
  unsigned zmask = 0;
  LACE_TRY(read_status(E, &st, &zmask));
  if (zmask != 0 && zmask != 0xFFFFFFFFu) {
    LACE_TRY(sd_fill_cols(E.n, r0, zmask, W));
    LACE_TRY(E.tg.tn(one_block(W), W, r0, r0, E.G1));
    LACE_TRY(sd_chol32(E.G1, rows, E.Rl, E.Rlinv, E.status));
    LACE_TRY(read_status(E, &st));
  } else {
    zmask = 0;
  }
  int np = 1, clean = 0, rp = 0;
  bool have_prev = false;
  const double* src = W;
  double* qbuf[2] = {E.Qa.as<double>(), E.Qb.as<double>()};
  int qi = 0;
  while (true) {
    if (st == 2) {
      *breakdown = 1;
      *qlast = nullptr;
      LACE_TRY(sd_copy32(0.0, nullptr, E.Rt, 32));
      *passes += np;
      return DEV_SUCCESS;
    }
    if (st == 1) {
      clean = 0;
      (*shifted)++;
    } else {
      clean++;
    }
    if (clean >= 2 || np >= 5) break;
    // Q <- src Rl^{-1}; cumulative R <- Rl Rprev
    LACE_TRY(sd_copy32(-1.0, E.Rlinv, E.negRlinv, 32));
    double* q = qbuf[qi];
    qi ^= 1;
    LACE_TRY((E.tg.nn<double, double>(one_block(src), E.negRlinv, nullptr, 0, 0, r0, q)));
    if (!have_prev) {
      LACE_TRY(sd_copy32(1.0, E.Rl, E.Rp[rp], 32));
      have_prev = true;
    } else {
      LACE_TRY(sd_gemm(32, 32, 32, 1.0, E.Rl, 32, false, E.Rp[rp], 32, false, 0.0, E.Rp[rp ^ 1], 32));
      rp ^= 1;
    }
    src = q;
    LACE_TRY(E.tg.tn(one_block(src), src, r0, r0, E.G1));
    LACE_TRY(sd_chol32(E.G1, rows, E.Rl, E.Rlinv, E.status));
    LACE_TRY(read_status(E, &st));
    np++;
  }
  *passes += np;
  LACE_TRY(sd_recon32(src, r0, E.Rl, E.Rlinv, have_prev ? E.Rp[rp] : E.Ident, E.Ltop, E.Tb, E.Rt, E.Dy));
  LACE_TRY(sd_zero_cols32(zmask, E.Rt));
  *qlast = src;
  return DEV_SUCCESS;
}

// Y (block) = [0; L; -Qlast_2 Dy] for the panel at r0.
static int form_y(BlockEngine& E, const double* qlast, int r0, double* Yb) {
  LACE_TRY((E.tg.nn<double, double>(one_block(qlast), E.Dy, nullptr, 0, 0, r0 + 32, Yb)));
  LACE_TRY(sd_write_top(E.Ltop, r0, Yb));
  return DEV_SUCCESS;
}

int block_panel_qr(BlockEngine& E, const double* W, int r0, double* Y, int* passes, int* shifted) {
  const double* q = nullptr;
  int brk = 0;
  LACE_HIP(hipMemcpyAsync(E.W.p, W, sizeof(double) * (size_t)E.n * kNrhs, hipMemcpyDeviceToDevice,
                          stream()));
  LACE_TRY(panel(E, E.W.as<double>(), r0, &q, passes, shifted, &brk));
  if (brk) {
    LACE_HIP(hipMemsetAsync(Y, 0, sizeof(double) * (size_t)E.n * kNrhs, stream()));
    return DEV_SUCCESS;
  }
  return form_y(E, q, r0, Y);
}

static BlockList block_list(const std::vector<std::unique_ptr<DevBuf>>& v, int nb) {
  BlockList L;
  L.nb = nb;
  for (int b = 0; b < nb; b++) L.p[b] = v[b]->p;
  return L;
}

template <typename T>
static int fgmres_t(BlockEngine& E, const BlockSolveArgs& g, const double* B, double* X,
                    lace_block_stats* st) {
  const int n = E.n, m = E.m;
  const size_t blk = (size_t)n * kNrhs;
  const DevCSR& A = *g.A;
  double* W = E.W.as<double>();
  double* R = E.R.as<double>();
  double tol[kNrhs];
  double* h_res = E.host;
  double* h_rn = E.host + 32;
  double* h_bn = E.host + 64;
  // ||b_c||
  LACE_TRY(sd_colnorms(n, B, E.partial.as<double>(), E.bn));
  int cycles = 0;
  bool done = false;
  bool first = true;
  while (!done && cycles < std::max(1, g.restart_max)) {
    // true residual of the current X
    LACE_TRY(spmm32<double>(A, X, B, R, true));
    LACE_TRY(sd_colnorms(n, R, E.partial.as<double>(), E.rn));
    LACE_HIP(hipMemcpyAsync(h_rn, E.rn, sizeof(double) * 32, hipMemcpyDeviceToHost, stream()));
    LACE_HIP(hipStreamSynchronize(stream()));
    if (first) {
      for (int c = 0; c < kNrhs; c++) {
        tol[c] = g.rtol * h_rn[c];
        st->initial_residual[c] = h_rn[c];
      }
      first = false;
    }
    bool all = true;
    for (int c = 0; c < kNrhs; c++) all = all && (h_rn[c] <= tol[c]);
    if (all) {
      done = true;
      break;
    }
    cycles++;
    st->restarts = cycles;
    // H_0 from the residual
    LACE_HIP(hipMemsetAsync(E.Tall, 0, sizeof(double) * (size_t)E.ldT * E.ldT, stream()));
    LACE_HIP(hipMemsetAsync(E.H.Gls, 0, sizeof(double) * (size_t)E.ldT * 32, stream()));
    const double* q = nullptr;
    int brk = 0;
    LACE_TRY(panel(E, R, 0, &q, &st->panel_passes, &st->shifted_panels, &brk));
    LACE_REQUIRE(!brk, DEV_ERR, "block FGMRES: zero residual panel with unconverged columns");
    LACE_TRY(form_y(E, q, 0, E.Y[0]->as<double>()));
    LACE_TRY(sd_copy32(1.0, E.Rt, E.H.Gls, 32));
    LACE_TRY(sd_copy32(1.0, E.Tb, E.Tall, E.ldT));
    int nblk = 0;
    for (int k = 0; k < m; k++) {
      const int nY = k + 1, K = nY * 32;
      BlockList Yl = block_list(E.Y, nY);
      // V_k = Q E_k
      LACE_TRY(sd_gather_rows(Yl, k * 32, E.a));
      LACE_TRY(sd_gemm(K, 32, K, 1.0, E.Tall, E.ldT, false, E.a, 32, false, 0.0, E.D, 32));
      LACE_TRY((E.tg.nn<double, T>(Yl, E.D, nullptr, 2, k * 32, 0, E.Vt.as<T>())));
      // Z_k = M^{-1} V_k; W = A Z_k
      T* zk = E.Z[k]->as<T>();
      if constexpr (std::is_same<T, float>::value) {
        if (g.inner_iters > 0) {
          LACE_TRY(E.inner->solve(*g.A32, g.P, g.kind, g.sweeps, g.trans, E.Vt.as<float>(), zk));
        } else {
          LACE_TRY(block_precond<T>(g.P, g.kind, g.sweeps, g.trans, E.Vt.as<T>(), zk, E.t1.as<T>(),
                                    E.t2.as<T>()));
        }
      } else {
        LACE_TRY(block_precond<T>(g.P, g.kind, g.sweeps, g.trans, E.Vt.as<T>(), zk, E.t1.as<T>(),
                                  E.t2.as<T>()));
      }
      LACE_TRY(spmm32<T>(A, zk, nullptr, W, false));
      // W <- Q^T W
      LACE_TRY(E.tg.tn(Yl, W, 0, 0, E.Gk));
      LACE_TRY(sd_gemm(K, 32, K, 1.0, E.Tall, E.ldT, true, E.Gk, 32, false, 0.0, E.Gp, 32));
      LACE_TRY((E.tg.nn<double, double>(Yl, E.Gp, W, 1, 0, 0, W)));
      // new panel, Hessenberg column, residual estimate
      LACE_TRY(panel(E, W, K, &q, &st->panel_passes, &st->shifted_panels, &brk));
      LACE_TRY(sd_hess_step(E.H, k, W, E.Rt, E.res));
      LACE_HIP(hipMemcpyAsync(h_res, E.res, sizeof(double) * 32, hipMemcpyDeviceToHost, stream()));
      LACE_HIP(hipStreamSynchronize(stream()));
      st->iterations++;
      nblk = k + 1;
      bool conv = true;
      for (int c = 0; c < kNrhs; c++) conv = conv && (h_res[c] <= tol[c]);
      if (conv || brk || k == m - 1) break;
      // Y_{k+1} and the grown T
      double* ynew = E.Y[k + 1]->as<double>();
      LACE_TRY(form_y(E, q, K, ynew));
      LACE_TRY(E.tg.tn(Yl, ynew, K, K, E.Km));
      LACE_TRY(sd_gemm(K, 32, K, 1.0, E.Tall, E.ldT, false, E.Km, 32, false, 0.0, E.tmpM, 32));
      LACE_TRY(sd_gemm(K, 32, 32, -1.0, E.tmpM, 32, false, E.Tb, 32, false, 0.0, E.Tall + K, E.ldT));
      LACE_TRY(sd_copy32(1.0, E.Tb, E.Tall + (size_t)K * E.ldT + K, E.ldT));
    }
    // least squares and X += Z Yls
    LACE_TRY(sd_backsolve(E.H, nblk, E.Dneg));
    BlockList Zl = block_list(E.Z, nblk);
    LACE_TRY((E.tg.nn<T, double>(Zl, E.Dneg, X, 1, 0, 0, X)));
  }
  // final true residuals
  LACE_TRY(spmm32<double>(A, X, B, R, true));
  LACE_TRY(sd_colnorms(n, R, E.partial.as<double>(), E.rn));
  LACE_HIP(hipMemcpyAsync(h_rn, E.rn, sizeof(double) * 32, hipMemcpyDeviceToHost, stream()));
  LACE_HIP(hipMemcpyAsync(h_bn, E.bn, sizeof(double) * 32, hipMemcpyDeviceToHost, stream()));
  LACE_HIP(hipStreamSynchronize(stream()));
  (void)blk;
  bool all = true;
  double mx = 0.0;
  for (int c = 0; c < kNrhs; c++) {
    const double rel = h_bn[c] > 0.0 ? h_rn[c] / h_bn[c] : h_rn[c];
    st->rel_true_residual[c] = rel;
    mx = std::max(mx, rel);
    all = all && (h_rn[c] <= tol[c]);
  }
  st->max_rel_true_residual = mx;
  st->converged = all ? 1 : 0;
  return DEV_SUCCESS;
}

int block_fgmres(BlockEngine& E, const BlockSolveArgs& g, const double* B, double* X,
                 lace_block_stats* st) {
  memset(st, 0, sizeof(*st));
  if (g.inner_iters > 0) {
    LACE_REQUIRE(E.zf && g.A32, DEV_ERR_NOT_SUPPORTED, "inner iterations need FP32 factors and A");
    if (!E.inner || E.inner->m != g.inner_iters) {
      E.inner.reset();
      auto I = std::make_unique<InnerBatch>();
      LACE_TRY(I->init(E.n, g.inner_iters));
      E.inner = std::move(I);
    }
  }
  const auto t0 = std::chrono::steady_clock::now();
  int rc;
  if (E.zf) rc = fgmres_t<float>(E, g, B, X, st);
  else rc = fgmres_t<double>(E, g, B, X, st);
  const auto t1 = std::chrono::steady_clock::now();
  st->solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return rc;
}

}  // namespace lace

using namespace lace;

extern "C" {

int lace_block_panel_qr(int n, int r0, const double* W, double* Y, double* Tb, double* Rt, int* passes) {
  LACE_TRY(ensure_init());
  LACE_REQUIRE(n >= 64 && r0 >= 0 && r0 + 32 <= n, DEV_ERR_ILLEGAL_VALUE, "panel qr: n %d r0 %d", n, r0);
  BlockEngine E;
  LACE_TRY(E.init(n, 1, false));
  int np = 0, sh = 0;
  LACE_TRY(block_panel_qr(E, W, r0, Y, &np, &sh));
  LACE_HIP(hipMemcpyAsync(Tb, E.Tb, sizeof(double) * 1024, hipMemcpyDeviceToDevice, stream()));
  LACE_HIP(hipMemcpyAsync(Rt, E.Rt, sizeof(double) * 1024, hipMemcpyDeviceToDevice, stream()));
  LACE_HIP(hipStreamSynchronize(stream()));
  if (passes) *passes = np;
  return DEV_SUCCESS;
}

}  // extern "C"
