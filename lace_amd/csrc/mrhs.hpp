// mrhs.hpp -- multi-right-hand-side sparse products and ILU triangular
// sweeps on n x 32 row-major blocks (row i holds the 32 right-hand sides'
// entries i).
#pragma once

#include "common.hpp"
#include "parilu.hpp"
#include "spmv.hpp"

namespace lace {

constexpr int kNrhs = 32;

// Y = A X (residual = false) or Y = B - A X (residual = true); A FP64 CSR.
// Column c of Y equals the single-vector CSR product of column c bit for
// bit (row sums in column order from zero, no contraction).
template <typename TX>
int spmm32(const DevCSR& A, const TX* X, const double* B, double* Y, bool residual);
// Y = A X with FP32 values, X and Y (FP32 accumulation).
int spmm32_f32(const DevCSR& A, const float* X, float* Y);

// Z = M^{-1} R (trans = 0) or M^{-T} R (trans = 1) with `sweeps` Jacobi
// sweeps per triangle, in the factor precision T; tmp and tmp2 are n x 32
// scratch blocks.  Column c equals ParILU::apply (Jacobi kind) on column c
// bit for bit.
template <typename T>
int block_precond(ParILU* P, lace_trsv_kind_t kind, int sweeps, int trans, const T* R, T* Z, T* tmp,
                  T* tmp2);

// Element-wise precision change of an n x 32 block.
int block_cast_d2f(int n, const double* in, float* out);

}  // namespace lace
