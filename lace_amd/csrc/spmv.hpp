// spmv.hpp -- device CSR matrix and sparse products.
#pragma once

#include "common.hpp"
#include "dia.hpp"

namespace lace {

struct DevCSR {
  int n = 0, ncols = 0, nnz = 0;
  int group = 8;  // lanes per row
  lace_prec_t prec = LACE_F64;
  DevBuf row, col, val;
  // ELLPACK copy for short rows (at most 8 per row, padding under 25 %):
  // ew slots per row, slot-major, column -1 for padding; eval is gathered
  // from val through emap by refresh_ell (call it after changing val).
  int ew = 0;
  DevBuf ecol, emap, eval;
  // Diagonal form of a square matrix with at most 8 diagonals (stencils),
  // refreshed with the ELLPACK copy.
  DiaForm dia;
  int create(const data_d_matrix* A, lace_prec_t p);
  int refresh_ell();
  // Algorithmic bytes of one product: values+columns, row pointers, x, y
  // (and b for the residual form); sx, sy are vector element sizes.
  double bytes(int sx, int sy, bool residual) const;
};

template <typename TX, typename TY>
int spmv(const DevCSR& A, double alpha, const TX* x, double beta, TY* y);
template <typename TX, typename TY>
int residual(const DevCSR& A, const TX* x, const TY* b, TY* r);

}  // namespace lace
