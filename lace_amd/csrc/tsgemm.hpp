// tsgemm.hpp -- tall-skinny FP64 products (n x 32 row-major blocks) on MFMA.
#pragma once

#include "common.hpp"

namespace lace {

constexpr int kMaxBlocks = 40;

// A list of n x 32 row-major blocks (FP64, or FP32 where a product says
// so), passed to kernels by value.
struct BlockList {
  const void* p[kMaxBlocks];
  int nb;
};

struct TsGemm {
  int n = 0, rpb = 0, nslab = 0;
  DevBuf partial, mid;  // per-slab partial products; per-slab-group sums
  int init(int n_);
  // C[b] (32 x 32, row-major, b < L.nb) = X_b^T W over rows [rs_b, n):
  // rs_b = rs_main for b < nb-1, rs_last for the last block.
  int tn(const BlockList& L, const double* W, int rs_main, int rs_last, double* C);
  // out = base - sum_b X_b D_b on rows [rlo, n) (D: nb consecutive 32 x 32
  // row-major blocks); rows below rlo get base.  base_mode 0: zero, 1: the
  // block `base`, 2: identity block column (1 at row e0 + c, column c).
  // out may be base; it must not be one of the X_b.  TX: element type of
  // the X_b (double or float, widened exactly); TO: element type of out.
  template <typename TX, typename TO>
  int nn(const BlockList& L, const double* D, const double* base, int base_mode, int e0, int rlo,
         TO* out);
};

}  // namespace lace
