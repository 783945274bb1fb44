// bkrylov.hpp -- block Householder FGMRES for 32 right-hand sides.
#pragma once

#include <memory>
#include <vector>

#include "binner.hpp"
#include "common.hpp"
#include "mrhs.hpp"
#include "parilu.hpp"
#include "smalldense.hpp"
#include "spmv.hpp"
#include "tsgemm.hpp"

namespace lace {

// Device state of one block solver (restart length m block steps).  Krylov
// data are n x 32 row-major blocks: Y[0..m] the Householder vectors of the
// block reflectors (FP64), Z[0..m-1] the preconditioned blocks (factor
// precision T: FP32 for FP32 factors).
struct BlockEngine {
  int n = 0, m = 0;
  bool zf = true;
  std::vector<std::unique_ptr<DevBuf>> Y, Z;
  DevBuf W, Qa, Qb, R, Vt, t1, t2;
  DevBuf small, ints, partial;
  TsGemm tg;
  HessState H;
  std::unique_ptr<InnerBatch> inner;  // nested FP32 inner iterations, when used
  int ldT = 0;  // leading dimension of the accumulated T, (m + 1) * 32
  double *Tall = nullptr, *a = nullptr, *D = nullptr, *Gk = nullptr, *Gp = nullptr, *Km = nullptr,
         *tmpM = nullptr, *Dneg = nullptr, *G1 = nullptr, *Rl = nullptr, *Rlinv = nullptr,
         *negRlinv = nullptr, *Rp[2] = {nullptr, nullptr}, *Ident = nullptr, *Ltop = nullptr,
         *Tb = nullptr, *Rt = nullptr, *Dy = nullptr, *Zero = nullptr, *res = nullptr, *rn = nullptr,
         *bn = nullptr;
  int* status = nullptr;
  double* host = nullptr;  // pinned: res[32], rn[32], bn[32]
  int* host_status = nullptr;
  ~BlockEngine();
  int init(int n_, int m_, bool zf_);
};

struct BlockSolveArgs {
  const DevCSR* A;    // operator (A or A^T), FP64 values
  const DevCSR* A32;  // its FP32 copy (inner iterations only)
  ParILU* P;          // ILU factors (precision decides T)
  int inner_iters;    // > 0: each outer step preconditions with this many
                      // batched inner FP32 GMRES steps (FP32 factors)
  lace_trsv_kind_t kind;
  int sweeps, trans, restart_max;
  double rtol;
};

int block_fgmres(BlockEngine& E, const BlockSolveArgs& args, const double* B, double* X,
                 lace_block_stats* st);

// Householder QR of rows [r0, n) of the n x 32 block W through CholeskyQR
// passes and reconstruction (test entry as well): Y (n x 32, zero above
// r0), Tb, Rt = S R.  Returns the number of passes in *passes.
int block_panel_qr(BlockEngine& E, const double* W, int r0, double* Y, int* passes, int* shifted);

}  // namespace lace
