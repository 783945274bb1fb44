// krylov.hpp -- device Householder FGMRES engine and preconditioners.
#pragma once

#include "common.hpp"
#include "parilu.hpp"
#include "spmv.hpp"

namespace lace {

// The operator of a solve: y = A x with the vector precisions the engines
// need (FP64 Krylov vectors, FP32 preconditioned vectors widened, and the
// FP32 operator of the inner solver), and the FP64 residual r = b - A x.
struct LinOp {
  virtual ~LinOp() = default;
  virtual int apply_dd(const double* x, double* y) const = 0;
  virtual int apply_fd(const float* x, double* y) const = 0;
  virtual int apply_ff(const float* x, float* y) const = 0;
  virtual int residual_dd(const double* x, const double* b, double* r) const = 0;
};

// z = M^{-1} q.  q has the Krylov precision of the calling engine; z is a
// column of the engine's Z store (precision ZT of the engine).
struct Precond {
  virtual ~Precond() = default;
  virtual int apply_d(const double* q, void* z) { (void)q; (void)z; return DEV_ERR_NOT_SUPPORTED; }
  virtual int apply_f(const float* q, void* z) { (void)q; (void)z; return DEV_ERR_NOT_SUPPORTED; }
  virtual bool z_is_float() const = 0;
};

// Small-state layout (doubles) of one engine, m = restart length.
struct SmallLayout {
  int m;
  __host__ __device__ int ldT() const { return m + 2; }
  __host__ __device__ int T() const { return 0; }
  __host__ __device__ int R() const { return T() + (m + 2) * (m + 2); }
  __host__ __device__ int cs() const { return R() + (m + 1) * (m + 1); }
  __host__ __device__ int sn() const { return cs() + (m + 1); }
  __host__ __device__ int g() const { return sn() + (m + 1); }
  __host__ __device__ int y() const { return g() + (m + 2); }
  __host__ __device__ int d() const { return y() + (m + 2); }
  __host__ __device__ int misc() const { return d() + (m + 2); }
  __host__ __device__ int total() const { return misc() + 32; }
};

// misc slots
enum {
  MS_DD = 0, MS_SCALE = 1, MS_ALPHA = 2, MS_RNORM = 3, MS_RES = 4, MS_DONE = 5,
  MS_RTOL = 6, MS_NCOLS = 7, MS_BREAK = 8
};

// One restart cycle engine (Walker's Householder Arnoldi in compact-WY
// form).  K: precision of the Krylov vectors and reflectors.
template <typename K>
struct HHEngine {
  int n = 0, m = 0, ld = 0, nb = 0, rpb = 4096, pstride = 0;
  bool z_float = false;
  DevBuf U, Z, w, q, r, small, partial, hostflag_dev;
  DevBuf ptot;  // column totals of partial (one row, pstride doubles)
  SmallLayout lay{0};
  double* host_misc = nullptr;  // pinned copy of the misc slots
  int init(int n_, int m_, bool z_float_);
  ~HHEngine();
  void* zcol(int j) const;
};

}  // namespace lace
