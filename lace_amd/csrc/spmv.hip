// spmv.hip -- CSR sparse matrix-vector products on the GPU.
//
// Reference: blas/zspmv.cpp:48-78 (mkl_dcsrmv wrapper) and the inline
// Arnoldi/residual loops of src/fgmres.cpp:259-268 and
// src/fgmres_householder_restart.cpp:210-221,357-367.
//
// Kernels: matrices with at most 16 nonzeros per row on average take the
// CSR-stream kernel (the block's nonzeros staged through LDS with
// coalesced loads, each row summed in CSR order by one thread).  Longer
// rows take the lane-group kernel: G lanes per row (G = 4, 8, 16 or 32
// from the mean row length), each lane strides the row's nonzeros, partial
// sums meet in a shuffle tree inside the lane group.  Blocks are remapped
// so each XCD streams one contiguous slab of rows.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "spmv.hpp"

namespace lace {

constexpr int kSpmvBS = 256;
constexpr int kSpmvRows = 4;  // rows per lane group

// MODE 0: y = alpha*A*x + beta*y  (beta == 0 never reads y)
// MODE 1: y = b - A*x             (residual form)
// Each lane group of G lanes handles R rows of the block's slab; the first
// G nonzeros of all R rows are loaded before any is used, so a wave keeps
// 3R independent loads per lane in flight (row bounds, column/value, x).
template <typename TV, typename TX, typename TY, int G, int MODE>
__global__ __launch_bounds__(kSpmvBS) void spmv_kernel(int n, const int* __restrict__ row,
                                                       const int* __restrict__ col,
                                                       const TV* __restrict__ val,
                                                       const TX* __restrict__ x, TY* __restrict__ y,
                                                       const TY* __restrict__ b, double alpha,
                                                       double beta) {
  using ACC = typename std::conditional<sizeof(TV) == 8 || sizeof(TY) == 8, double, float>::type;
  constexpr int R = kSpmvRows;
  constexpr int GPB = kSpmvBS / G;  // lane groups per block
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int lane = threadIdx.x & (G - 1);
  const int g = threadIdx.x / G;
  const int base = blk * GPB * R + g;
  int kb[R], ke[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int i = base + r * GPB;
    kb[r] = i < n ? row[i] : 0;
    ke[r] = i < n ? row[i + 1] : 0;
  }
  ACC s[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int k = kb[r] + lane;
    s[r] = k < ke[r] ? (ACC)val[k] * (ACC)x[col[k]] : ACC(0);
  }
#pragma unroll
  for (int r = 0; r < R; r++)
    for (int k = kb[r] + lane + G; k < ke[r]; k += G) s[r] += (ACC)val[k] * (ACC)x[col[k]];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int i = base + r * GPB;
    const ACC t = group_sum<G>(s[r]);
    if (i < n && lane == 0) {
      if (MODE == 1) {
        y[i] = (TY)((ACC)b[i] - t);
      } else {
        ACC v = (ACC)alpha * t;
        if (beta != 0.0) v += (ACC)beta * (ACC)y[i];
        y[i] = (TY)v;
      }
    }
  }
}

// CSR-stream form for short rows: a block owns 256*R consecutive rows
// (thread t: rows r0 + t + 256 q, q < R); their nonzeros form one
// contiguous range of col/val, which the block reads with consecutive
// threads on consecutive entries (chunks of CAP), forming val*x[col] into
// LDS.  Each thread then sums its rows' products in column order, so each
// row is summed sequentially from zero exactly as a serial CSR loop does.
// The per-row operands (row bounds, b or y) are loaded before the staging
// so their latency overlaps it.
constexpr int kStreamThreads = 256;

template <typename ACC, int R>
struct StreamCap {
  static constexpr int rows = kStreamThreads * R;
  static constexpr int cap = (rows * 8 * (int)sizeof(ACC) <= 32768) ? rows * 8 : 32768 / (int)sizeof(ACC);
};

template <typename TV, typename TX, typename TY, int MODE, int R>
__global__ __launch_bounds__(kStreamThreads) void spmv_stream_kernel(
    int n, const int* __restrict__ row, const int* __restrict__ col, const TV* __restrict__ val,
    const TX* __restrict__ x, TY* __restrict__ y, const TY* __restrict__ b, double alpha, double beta) {
#pragma clang fp contract(off)
  using ACC = typename std::conditional<sizeof(TV) == 8 || sizeof(TY) == 8, double, float>::type;
  constexpr int ROWS = StreamCap<ACC, R>::rows, CAP = StreamCap<ACC, R>::cap;
  __shared__ ACC prod[CAP];
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int r0 = blk * ROWS;
  const int r1 = min(r0 + ROWS, n);
  int rs[R], re[R];
  ACC add[R], acc[R];
#pragma unroll
  for (int q = 0; q < R; q++) {
    const int i = r0 + (int)threadIdx.x + q * kStreamThreads;
    rs[q] = re[q] = 0;
    add[q] = acc[q] = ACC(0);
    if (i < r1) {
      rs[q] = row[i];
      re[q] = row[i + 1];
      if (MODE == 1) add[q] = (ACC)b[i];
      else if (beta != 0.0) add[q] = (ACC)beta * (ACC)y[i];
    }
  }
  const int k0 = row[r0], k1 = row[r1];
  for (int c0 = k0; c0 < k1; c0 += CAP) {
    const int cn = min(CAP, k1 - c0);
    stage_products<kStreamThreads, ACC>(col, val, x, c0, cn, prod);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R; q++) {
      const int lo = max(rs[q], c0), hi = min(re[q], c0 + cn);
      for (int k = lo; k < hi; k++) acc[q] = acc[q] + prod[k - c0];
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < R; q++) {
    const int i = r0 + (int)threadIdx.x + q * kStreamThreads;
    if (i < r1) {
      if (MODE == 1) {
        y[i] = (TY)(add[q] - acc[q]);
      } else {
        ACC v = (ACC)alpha * acc[q];
        if (beta != 0.0) v = v + add[q];
        y[i] = (TY)v;
      }
    }
  }
}

template <typename TV, typename TX, typename TY, int MODE, int R>
static void launch_stream(const DevCSR& A, const TX* x, TY* y, const TY* b, double alpha, double beta) {
  using ACC = typename std::conditional<sizeof(TV) == 8 || sizeof(TY) == 8, double, float>::type;
  spmv_stream_kernel<TV, TX, TY, MODE, R>
      <<<ceil_div(A.n, StreamCap<ACC, R>::rows), kStreamThreads, 0, stream()>>>(
          A.n, A.row.as<int>(), A.col.as<int>(), A.val.as<TV>(), x, y, b, alpha, beta);
}

// ELLPACK form: thread per row, the row's slots read slot-major (coalesced),
// padding skipped, products summed in CSR order from zero -- the same
// arithmetic as the stream kernel, so the results are bit-identical.
template <typename TV, typename TX, typename TY, int MODE, int W>
__global__ __launch_bounds__(256) void spmv_ell_kernel(int n, const int* __restrict__ ecol,
                                                       const TV* __restrict__ eval,
                                                       const TX* __restrict__ x, TY* __restrict__ y,
                                                       const TY* __restrict__ b, double alpha, double beta) {
#pragma clang fp contract(off)
  using ACC = typename std::conditional<sizeof(TV) == 8 || sizeof(TY) == 8, double, float>::type;
  const int i = xcd_block(blockIdx.x, gridDim.x) * 256 + (int)threadIdx.x;
  if (i >= n) return;
  int c[W];
  TV v[W];
  TX xv[W];
#pragma unroll
  for (int q = 0; q < W; q++) {
    c[q] = ecol[(size_t)q * n + i];
    v[q] = eval[(size_t)q * n + i];
  }
#pragma unroll
  for (int q = 0; q < W; q++) xv[q] = c[q] >= 0 ? x[c[q]] : TX(0);
  ACC add = ACC(0);
  if (MODE == 1) add = (ACC)b[i];
  else if (beta != 0.0) add = (ACC)beta * (ACC)y[i];
  ACC s = ACC(0);
#pragma unroll
  for (int q = 0; q < W; q++)
    if (c[q] >= 0) s = s + (ACC)v[q] * (ACC)xv[q];
  if (MODE == 1) {
    y[i] = (TY)(add - s);
  } else {
    ACC o = (ACC)alpha * s;
    if (beta != 0.0) o = o + add;
    y[i] = (TY)o;
  }
}

// Diagonal form: thread per row, the row's value on diagonal d read at
// d*n + i and x at i + off[d] (contiguous across the wavefront; rows without
// an entry there read x[i] and skip it); the products are summed in
// ascending offset = CSR order, the same arithmetic as the other forms.
template <typename TV, typename TX, typename TY, int MODE, int D>
__global__ __launch_bounds__(256) void spmv_dia_kernel(int n, DiaOffsets off, const uint8_t* __restrict__ mask,
                                                       const TV* __restrict__ dia,
                                                       const TX* __restrict__ x, TY* __restrict__ y,
                                                       const TY* __restrict__ b, double alpha, double beta) {
#pragma clang fp contract(off)
  using ACC = typename std::conditional<sizeof(TV) == 8 || sizeof(TY) == 8, double, float>::type;
  const int i = xcd_block(blockIdx.x, gridDim.x) * 256 + (int)threadIdx.x;
  if (i >= n) return;
  const unsigned m = mask[i];
  TV v[D];
  TX xv[D];
#pragma unroll
  for (int q = 0; q < D; q++) v[q] = dia[(size_t)q * n + i];
#pragma unroll
  for (int q = 0; q < D; q++) xv[q] = x[((m >> q) & 1u) ? i + off.o[q] : i];
  ACC add = ACC(0);
  if (MODE == 1) add = (ACC)b[i];
  else if (beta != 0.0) add = (ACC)beta * (ACC)y[i];
  ACC s = ACC(0);
#pragma unroll
  for (int q = 0; q < D; q++)
    if ((m >> q) & 1u) s = s + (ACC)v[q] * (ACC)xv[q];
  if (MODE == 1) {
    y[i] = (TY)(add - s);
  } else {
    ACC o = (ACC)alpha * s;
    if (beta != 0.0) o = o + add;
    y[i] = (TY)o;
  }
}

// Which form launch_spmv uses: 2 diagonal, 1 ELLPACK, 0 CSR.
static int spmv_form(const DevCSR& A) {
  if (A.dia.nd > 0 && get_option("spmv_dia", 1) != 0) return 2;
  if (A.ew > 0 && get_option("spmv_ell", 1) != 0) return 1;
  return 0;
}

template <typename TV, typename TX, typename TY, int MODE>
static int launch_spmv(const DevCSR& A, const TX* x, TY* y, const TY* b, double alpha, double beta,
                       TimedRegion& tr) {
  const int form = spmv_form(A);
  if (form == 2) {
    const int nb = ceil_div(A.n, 256);
    const uint8_t* mk = A.dia.mask.as<uint8_t>();
    const TV* dv = A.dia.val.as<TV>();
    switch (A.dia.nd) {
#define LACE_SPMV_DIA(D)                                                                                     \
  case D:                                                                                                    \
    tr.launch(spmv_dia_kernel<TV, TX, TY, MODE, D>, dim3(nb), dim3(256), A.n, A.dia.off, mk, dv, x, y, b, \
              alpha, beta);                                                                                 \
    break;
      LACE_SPMV_DIA(1)
      LACE_SPMV_DIA(2)
      LACE_SPMV_DIA(3)
      LACE_SPMV_DIA(4)
      LACE_SPMV_DIA(5)
      LACE_SPMV_DIA(6)
      LACE_SPMV_DIA(7)
      LACE_SPMV_DIA(8)
#undef LACE_SPMV_DIA
      default:
        break;
    }
    LACE_HIP(hipGetLastError());
    return DEV_SUCCESS;
  }
  if (form == 1) {
    const int nb = ceil_div(A.n, 256);
    const int* ec = A.ecol.as<int>();
    const TV* ev = A.eval.as<TV>();
    switch (A.ew) {
#define LACE_SPMV_ELL(W)                                                                               \
  case W:                                                                                              \
    spmv_ell_kernel<TV, TX, TY, MODE, W><<<nb, 256, 0, stream()>>>(A.n, ec, ev, x, y, b, alpha, beta); \
    break;
      LACE_SPMV_ELL(1)
      LACE_SPMV_ELL(2)
      LACE_SPMV_ELL(3)
      LACE_SPMV_ELL(4)
      LACE_SPMV_ELL(5)
      LACE_SPMV_ELL(6)
      LACE_SPMV_ELL(7)
      LACE_SPMV_ELL(8)
#undef LACE_SPMV_ELL
      default:
        break;
    }
    LACE_HIP(hipGetLastError());
    return DEV_SUCCESS;
  }
  if (A.n > 0 && A.nnz <= 16LL * A.n) {
    // short rows (at most 16 per row on average): the stream kernel
    // measured at 128^3: two rows per thread for FP32 values, one for FP64
    const int R = get_option("stream_rows_per_thread", sizeof(TV) == 8 ? 1 : 2);
    if (R == 1) launch_stream<TV, TX, TY, MODE, 1>(A, x, y, b, alpha, beta);
    else if (R == 4) launch_stream<TV, TX, TY, MODE, 4>(A, x, y, b, alpha, beta);
    else launch_stream<TV, TX, TY, MODE, 2>(A, x, y, b, alpha, beta);
    LACE_HIP(hipGetLastError());
    return DEV_SUCCESS;
  }
  const int g = A.group;
  const int rows_per_block = (kSpmvBS / g) * kSpmvRows;
  const int nb = ceil_div(A.n, rows_per_block);
  const int* row = A.row.as<int>();
  const int* col = A.col.as<int>();
  const TV* val = A.val.as<TV>();
  switch (g) {
    case 4: spmv_kernel<TV, TX, TY, 4, MODE><<<nb, kSpmvBS, 0, stream()>>>(A.n, row, col, val, x, y, b, alpha, beta); break;
    case 8: spmv_kernel<TV, TX, TY, 8, MODE><<<nb, kSpmvBS, 0, stream()>>>(A.n, row, col, val, x, y, b, alpha, beta); break;
    case 16: spmv_kernel<TV, TX, TY, 16, MODE><<<nb, kSpmvBS, 0, stream()>>>(A.n, row, col, val, x, y, b, alpha, beta); break;
    default: spmv_kernel<TV, TX, TY, 32, MODE><<<nb, kSpmvBS, 0, stream()>>>(A.n, row, col, val, x, y, b, alpha, beta); break;
  }
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

int DevCSR::create(const data_d_matrix* A, lace_prec_t p) {
  LACE_TRY(ensure_init());
  LACE_REQUIRE(A && A->row && A->col && A->val, DEV_ERR_INVALID_PTR, "DevCSR: null arrays");
  LACE_REQUIRE(p == LACE_F64 || p == LACE_F32, DEV_ERR_NOT_SUPPORTED, "DevCSR: precision %d", (int)p);
  n = A->num_rows;
  ncols = A->num_cols;
  nnz = A->row[n];
  prec = p;
  const double mean = n ? (double)nnz / n : 0.0;
  group = mean <= 4.0 ? 4 : (mean <= 10.0 ? 8 : (mean <= 24.0 ? 16 : 32));
  LACE_TRY(upload(row, A->row, sizeof(int) * (n + 1)));
  LACE_TRY(upload(col, A->col, sizeof(int) * std::max(nnz, 1)));
  if (p == LACE_F64) {
    LACE_TRY(upload(val, A->val, sizeof(double) * std::max(nnz, 1)));
  } else {
    std::vector<float> f(std::max(nnz, 1));
    for (int k = 0; k < nnz; k++) f[k] = (float)A->val[k];
    LACE_TRY(upload(val, f.data(), sizeof(float) * f.size()));
  }
  int wmax = 0;
  for (int i = 0; i < n; i++) wmax = std::max(wmax, A->row[i + 1] - A->row[i]);
  ew = 0;
  if (n > 0 && wmax >= 1 && wmax <= 8 && (double)wmax * n <= 1.25 * nnz) {
    ew = wmax;
    std::vector<int> ec((size_t)ew * n, -1), em((size_t)ew * n, -1);
    for (int i = 0; i < n; i++)
      for (int k = A->row[i]; k < A->row[i + 1]; k++) {
        ec[(size_t)(k - A->row[i]) * n + i] = A->col[k];
        em[(size_t)(k - A->row[i]) * n + i] = k;
      }
    LACE_TRY(upload(ecol, ec.data(), sizeof(int) * ec.size()));
    LACE_TRY(upload(emap, em.data(), sizeof(int) * em.size()));
  }
  if (n == ncols) LACE_TRY(dia.build(n, A->row, A->col, [](int, int) { return true; }));
  return refresh_ell();
}

template <typename T>
__global__ void csr_ell_gather_kernel(size_t m, const int* __restrict__ emap, const T* __restrict__ val,
                                      T* __restrict__ eval) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < m) eval[e] = emap[e] >= 0 ? val[emap[e]] : T(0);
}

int DevCSR::refresh_ell() {
  const size_t es = prec == LACE_F64 ? sizeof(double) : sizeof(float);
  LACE_TRY(dia.gather(val.p, es));
  if (ew == 0) return DEV_SUCCESS;
  const size_t m = (size_t)ew * n;
  if (eval.bytes < es * m) LACE_TRY(eval.alloc(es * m));
  if (prec == LACE_F64)
    csr_ell_gather_kernel<double><<<ceil_div((long)m, 256), 256, 0, stream()>>>(
        m, emap.as<int>(), val.as<double>(), eval.as<double>());
  else
    csr_ell_gather_kernel<float><<<ceil_div((long)m, 256), 256, 0, stream()>>>(
        m, emap.as<int>(), val.as<float>(), eval.as<float>());
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

// Bytes of the form launch_spmv uses, each element once: the matrix
// (values and masks, values and columns, or CSR), x, y, and b for the
// residual form.
double DevCSR::bytes(int sx, int sy, bool residual) const {
  const double sv = prec == LACE_F64 ? 8.0 : 4.0;
  const int form = spmv_form(*this);
  const double mat = form == 2 ? dia.bytes((size_t)sv)
                     : form == 1 ? (double)ew * n * (sv + 4.0)
                                 : (double)nnz * (sv + 4.0) + 4.0 * (n + 1);
  return mat + (double)sx * ncols + (double)sy * n + (residual ? (double)sy * n : 0.0);
}

// y = alpha*A*x + beta*y; vector precisions follow the template arguments.
template <typename TX, typename TY>
int spmv(const DevCSR& A, double alpha, const TX* x, double beta, TY* y) {
  // the diagonal-form launch carries the events itself
  TimedRegion tr(A.prec == LACE_F64 ? "spmv" : "spmv_f32a", 0.0,
                 A.bytes((int)sizeof(TX), (int)sizeof(TY), false) + (beta != 0.0 ? sizeof(TY) * (double)A.n : 0.0),
                 spmv_form(A) == 2);
  if (A.prec == LACE_F64) return launch_spmv<double, TX, TY, 0>(A, x, y, nullptr, alpha, beta, tr);
  return launch_spmv<float, TX, TY, 0>(A, x, y, nullptr, alpha, beta, tr);
}

template <typename TX, typename TY>
int residual(const DevCSR& A, const TX* x, const TY* b, TY* r) {
  TimedRegion tr("spmv_residual", 0.0, A.bytes((int)sizeof(TX), (int)sizeof(TY), true), spmv_form(A) == 2);
  if (A.prec == LACE_F64) return launch_spmv<double, TX, TY, 1>(A, x, r, b, 1.0, 0.0, tr);
  return launch_spmv<float, TX, TY, 1>(A, x, r, b, 1.0, 0.0, tr);
}

template int spmv<double, double>(const DevCSR&, double, const double*, double, double*);
template int spmv<float, double>(const DevCSR&, double, const float*, double, double*);
template int spmv<float, float>(const DevCSR&, double, const float*, double, float*);
template int residual<double, double>(const DevCSR&, const double*, const double*, double*);
template int residual<float, float>(const DevCSR&, const float*, const float*, float*);

}  // namespace lace

using namespace lace;

extern "C" {

int lace_csr_create(const data_d_matrix* A, lace_prec_t prec, lace_csr** out) {
  auto* M = new DevCSR();
  int rc = M->create(A, prec);
  if (rc != DEV_SUCCESS) {
    delete M;
    return rc;
  }
  *out = reinterpret_cast<lace_csr*>(M);
  return DEV_SUCCESS;
}

int lace_csr_destroy(lace_csr* A) {
  delete reinterpret_cast<DevCSR*>(A);
  return DEV_SUCCESS;
}

int lace_csr_n(const lace_csr* A) { return reinterpret_cast<const DevCSR*>(A)->n; }

int lace_csr_forms(const lace_csr* Ah, int* ndiag, int* ell_width) {
  LACE_REQUIRE(Ah && ndiag && ell_width, DEV_ERR_INVALID_PTR, "lace_csr_forms: null argument");
  const DevCSR* A = reinterpret_cast<const DevCSR*>(Ah);
  *ndiag = A->dia.nd;
  *ell_width = A->ew;
  return DEV_SUCCESS;
}

int lace_csr_spmv(const lace_csr* Ah, double alpha, const void* x, double beta, void* y) {
  LACE_TRY(ensure_init());
  const DevCSR& A = *reinterpret_cast<const DevCSR*>(Ah);
  if (A.prec == LACE_F64)
    return spmv<double, double>(A, alpha, static_cast<const double*>(x), beta, static_cast<double*>(y));
  return spmv<float, float>(A, alpha, static_cast<const float*>(x), beta, static_cast<float*>(y));
}

// blas/zspmv.cpp:48-78: y = alpha*A*x + beta*y with host arrays.
int data_z_spmv(dataType alpha, data_d_matrix* A, data_d_matrix* x, dataType beta, data_d_matrix* y) {
  LACE_TRY(ensure_init());
  LACE_REQUIRE(A->storage_type == Magma_CSR || A->storage_type == Magma_CSRL ||
                   A->storage_type == Magma_CSRU,
               DEV_ERR_NOT_SUPPORTED, "data_z_spmv: storage %d not supported", (int)A->storage_type);
  DevCSR M;
  LACE_TRY(M.create(A, LACE_F64));
  DevBuf dx, dy;
  LACE_TRY(upload(dx, x->val, sizeof(double) * A->num_cols));
  if (beta != 0.0) LACE_TRY(upload(dy, y->val, sizeof(double) * A->num_rows));
  else LACE_TRY(dy.alloc(sizeof(double) * A->num_rows));
  LACE_TRY((spmv<double, double>(M, alpha, dx.as<double>(), beta, dy.as<double>())));
  LACE_TRY(download(y->val, dy, sizeof(double) * A->num_rows));
  return DEV_SUCCESS;
}

}  // extern "C"
