// binner.hpp -- batched inner FP32 GMRES for the 32-right-hand-side solver.
#pragma once

#include "common.hpp"
#include "parilu.hpp"
#include "spmv.hpp"

namespace lace {

constexpr int kMaxInner = 12;

// z_c ~ A^{-1} v_c for the 32 columns independently: `m` steps of
// right-preconditioned GMRES from zero with the fixed FP32 ILU
// preconditioner and the FP32 copy of A, classical Gram-Schmidt with
// FP64 dot products on a basis kept unnormalised past its first vector,
// Givens least squares per column, z = sum_j M^{-1} W_j y_j.
// Used as the preconditioner of the outer FP64 block FGMRES (flexible in
// the preconditioner, so the inner solve may change from step to step).
struct InnerBatch {
  int n = 0, m = 0, rpb = 1024, nslab = 0;
  DevBuf Q, P, w, t1, t2;  // FP32 n x 32 blocks: Q (m+1), P (m), w, Jacobi scratch
  DevBuf state, partial, red;
  int init(int n_, int m_);
  int solve(const DevCSR& A32, ParILU* F, lace_trsv_kind_t kind, int sweeps, int trans, const float* V,
            float* Z);
};

}  // namespace lace
