// mrhs.hip -- multi-right-hand-side kernels for the block Krylov path.
//
// Blocks are n x 32 and row-major, so one row of a block is 32 contiguous
// values (256 bytes in FP64, 128 in FP32).  A wave covers two matrix rows
// with one lane per right-hand side: each half-wave reads the row's column
// indices and values once (the 32 lanes share the addresses) and gathers
// whole 256- or 128-byte rows of the right-hand block, so every access to
// the dense blocks is a full, aligned row.  The matrix is read once per
// product for all 32 right-hand sides.
//
// Arithmetic per column is the single-vector arithmetic of the CSR-stream
// kernels (products val*x in column order summed from zero, then b - s or
// (r - s) / diag, contraction off), so each column of a block result is the
// single-vector result of that column bit for bit.  The sparse triangular
// sweep formula is src/trisolve.cpp:281-290 (data_parcsrtrsv) with all reads
// from the previous sweep; the product is the row loop of
// blas/zspmv.cpp:48-78 applied to 32 vectors at once (the reference has no
// sparse-times-dense product: blas/zspmm.cpp is sparse times sparse).
#include <algorithm>

#include "common.hpp"
#include "device.hpp"
#include "mrhs.hpp"

namespace lace {

constexpr int kChunk = 8;  // nonzeros whose operands are loaded ahead

// Lane geometry: each lane owns W = 16 / sizeof(TX) consecutive right-hand
// sides of one row (one 16-byte load of X per nonzero), 32 / W lanes per
// row, 256 W / 32 rows per workgroup.
template <typename TX>
struct Lanes {
  static constexpr int W = 16 / (int)sizeof(TX);
  static constexpr int LPR = 32 / W;
  static constexpr int RPB = 256 / LPR;
};

// Order in which the row blocks of the ELLPACK kernels run.  On a 3-D grid
// (rows x + nx (y + ny z)) row i reads the block rows i +- nx ny, so in
// natural order a row of the right-hand block is read again two planes
// after its first read -- 4 MB of an FP32 n x 32 block at 128^3, an XCD's
// whole L2 -- and comes from HBM up to three times.  When the largest
// diagonal offset m of the pattern divides n into planes of more than 1 MB
// of right-hand rows, each plane is cut into segments of about 256 KB of
// rows and the blocks run segment by segment, each segment through all
// planes: an XCD (which runs a contiguous range of the logical blocks)
// then marches one segment through z, and a right-hand row's reads all
// fall within a few segment-planes of each other.  Only the order changes:
// every row is computed by the same code from the same operands.
struct PlaneOrder {
  int nplane = 0;  // rows per plane; 0: natural order
  int seg = 0;     // rows per segment (a multiple of the rows per block)
  int useg = 0;    // blocks per segment and plane
  int nseg = 0;    // segments per plane
};

// The row that thread group g of logical block lb computes, or -1.
template <int RPB>
__device__ __forceinline__ int plane_row(const PlaneOrder& po, int lb, int g, int n) {
  if (po.nplane == 0) {
    const int i = lb * RPB + g;
    return i < n ? i : -1;
  }
  const int per_seg = (n / po.nplane) * po.useg;  // blocks of one segment over all planes
  const int sg = lb / per_seg, rem = lb - sg * per_seg;
  const int z = rem / po.useg, k = rem - z * po.useg;
  const int p = sg * po.seg + k * RPB + g;  // row within the plane
  return (sg < po.nseg && p < po.nplane) ? z * po.nplane + p : -1;
}

// The order for a pattern with diagonal offsets dia (its form need not be
// gathered), rows of rpb per block and right-hand rows of row_bytes, and
// the number of blocks to launch.
static PlaneOrder plane_order(const DiaForm& dia, int n, int rpb, size_t row_bytes, int* nblocks) {
  PlaneOrder po;
  *nblocks = ceil_div(n, rpb);
  if (dia.nd < 3 || n <= 0 || get_option("mrhs_plane_order", 1) == 0) return po;
  int m = 0;
  for (int d = 0; d < dia.nd; d++) m = std::max(m, std::abs(dia.off.o[d]));
  if (m <= 0 || n % m != 0 || n / m < 2 || (double)m * (double)row_bytes <= 1048576.0) return po;
  int seg = std::max(rpb, (int)(262144 / row_bytes) / rpb * rpb);
  seg = std::min(seg, ceil_div(m, rpb) * rpb);
  po.nplane = m;
  po.seg = seg;
  po.useg = seg / rpb;
  po.nseg = ceil_div(m, seg);
  *nblocks = po.nseg * (n / m) * po.useg;
  return po;
}

// Y = A X (MODE 0) or Y = B - A X (MODE 1).  Accumulation in double when the
// values or the output are double (the single-vector rule).
template <typename TV, typename TX, typename TY, int MODE>
__global__ __launch_bounds__(256) void spmm32_kernel(int n, const int* __restrict__ row,
                                                     const int* __restrict__ col,
                                                     const TV* __restrict__ val,
                                                     const TX* __restrict__ X,
                                                     const TY* __restrict__ B,
                                                     TY* __restrict__ Y) {
#pragma clang fp contract(off)
  using ACC = typename std::conditional<sizeof(TV) == 8 || sizeof(TY) == 8, double, float>::type;
  constexpr int W = Lanes<TX>::W, LPR = Lanes<TX>::LPR, RPB = Lanes<TX>::RPB;
  const int i = xcd_block(blockIdx.x, gridDim.x) * RPB + (int)threadIdx.x / LPR;
  const int c0 = ((int)threadIdx.x % LPR) * W;
  if (i >= n) return;
  const int k0 = row[i], k1 = row[i + 1];
  TY add[W];
  if (MODE == 1) vload<TY, W>(B + (size_t)i * 32 + c0, add);
  ACC s[W];
#pragma unroll
  for (int w = 0; w < W; w++) s[w] = ACC(0);
  for (int kb = k0; kb < k1; kb += kChunk) {
    int cc[kChunk];
    TV vv[kChunk];
    TX xv[kChunk][W];
#pragma unroll
    for (int e = 0; e < kChunk; e++) {
      const bool in = kb + e < k1;
      cc[e] = in ? col[kb + e] : 0;
      vv[e] = in ? val[kb + e] : TV(0);
    }
#pragma unroll
    for (int e = 0; e < kChunk; e++)
      if (kb + e < k1) vload<TX, W>(X + (size_t)cc[e] * 32 + c0, xv[e]);
#pragma unroll
    for (int e = 0; e < kChunk; e++)
      if (kb + e < k1) {
#pragma unroll
        for (int w = 0; w < W; w++) s[w] = s[w] + (ACC)vv[e] * (ACC)xv[e][w];
      }
  }
  TY out[W];
#pragma unroll
  for (int w = 0; w < W; w++) out[w] = MODE == 1 ? (TY)((ACC)add[w] - s[w]) : (TY)s[w];
  vstore<TY, W>(Y + (size_t)i * 32 + c0, out);
}

// The product on A's ELLPACK copy (slot-major, column -1 for padding).
template <typename TV, typename TX, typename TY, int MODE, int WS>
__global__ __launch_bounds__(256) void spmm32_ell_kernel(int n, const int* __restrict__ ecol,
                                                         const TV* __restrict__ eval,
                                                         const TX* __restrict__ X,
                                                         const TY* __restrict__ B, TY* __restrict__ Y,
                                                         PlaneOrder po) {
#pragma clang fp contract(off)
  using ACC = typename std::conditional<sizeof(TV) == 8 || sizeof(TY) == 8, double, float>::type;
  constexpr int W = Lanes<TX>::W, LPR = Lanes<TX>::LPR, RPB = Lanes<TX>::RPB;
  const int i = plane_row<RPB>(po, xcd_block(blockIdx.x, gridDim.x), (int)threadIdx.x / LPR, n);
  const int c0 = ((int)threadIdx.x % LPR) * W;
  if (i < 0) return;
  int cc[WS];
  TV vv[WS];
  TX xv[WS][W];
#pragma unroll
  for (int q = 0; q < WS; q++) {
    cc[q] = ecol[(size_t)q * n + i];
    vv[q] = eval[(size_t)q * n + i];
  }
#pragma unroll
  for (int q = 0; q < WS; q++)
    if (cc[q] >= 0) vload<TX, W>(X + (size_t)cc[q] * 32 + c0, xv[q]);
  TY add[W];
  if (MODE == 1) vload<TY, W>(B + (size_t)i * 32 + c0, add);
  ACC s[W];
#pragma unroll
  for (int w = 0; w < W; w++) s[w] = ACC(0);
#pragma unroll
  for (int q = 0; q < WS; q++)
    if (cc[q] >= 0) {
#pragma unroll
      for (int w = 0; w < W; w++) s[w] = s[w] + (ACC)vv[q] * (ACC)xv[q][w];
    }
  TY out[W];
#pragma unroll
  for (int w = 0; w < W; w++) out[w] = MODE == 1 ? (TY)((ACC)add[w] - s[w]) : (TY)s[w];
  vstore<TY, W>(Y + (size_t)i * 32 + c0, out);
}

// The product on A's diagonal form (offsets ascending, values diagonal-
// major, a mask bit per entry): the value and right-hand loads of a row do
// not wait on column indices, so a row takes one round trip to memory.
// Every diagonal's right-hand row is loaded (clamped into the block; only
// the masked entries are added), and a row's entries by ascending offset
// are its CSR entries by ascending column, so the sums are the ELLPACK
// kernel's, term by term.
template <typename TV, typename TX, typename TY, int MODE, int ND>
__global__ __launch_bounds__(256) void spmm32_dia_kernel(int n, DiaOffsets off, const uint8_t* __restrict__ mask,
                                                         const TV* __restrict__ dval,
                                                         const TX* __restrict__ X,
                                                         const TY* __restrict__ B, TY* __restrict__ Y,
                                                         PlaneOrder po) {
#pragma clang fp contract(off)
  using ACC = typename std::conditional<sizeof(TV) == 8 || sizeof(TY) == 8, double, float>::type;
  constexpr int W = Lanes<TX>::W, LPR = Lanes<TX>::LPR, RPB = Lanes<TX>::RPB;
  const int i = plane_row<RPB>(po, xcd_block(blockIdx.x, gridDim.x), (int)threadIdx.x / LPR, n);
  const int c0 = ((int)threadIdx.x % LPR) * W;
  if (i < 0) return;
  const unsigned mk = mask[i];
  TV vv[ND];
  TX xv[ND][W];
#pragma unroll
  for (int d = 0; d < ND; d++) vv[d] = dval[(size_t)d * n + i];
#pragma unroll
  for (int d = 0; d < ND; d++) {
    const int c = min(max(i + off.o[d], 0), n - 1);
    vload<TX, W>(X + (size_t)c * 32 + c0, xv[d]);
  }
  TY add[W];
  if (MODE == 1) vload<TY, W>(B + (size_t)i * 32 + c0, add);
  ACC s[W];
#pragma unroll
  for (int w = 0; w < W; w++) s[w] = ACC(0);
#pragma unroll
  for (int d = 0; d < ND; d++)
    if ((mk >> d) & 1u) {
#pragma unroll
      for (int w = 0; w < W; w++) s[w] = s[w] + (ACC)vv[d] * (ACC)xv[d][w];
    }
  TY out[W];
#pragma unroll
  for (int w = 0; w < W; w++) out[w] = MODE == 1 ? (TY)((ACC)add[w] - s[w]) : (TY)s[w];
  vstore<TY, W>(Y + (size_t)i * 32 + c0, out);
}

template <typename TV, typename TX, typename TY, int MODE>
static bool spmm32_dia(const DevCSR& A, const TX* X, const TY* B, TY* Y) {
  const uint8_t* mk = A.dia.mask.as<uint8_t>();
  const TV* dv = A.dia.val.as<TV>();
  int nb = 0;
  const PlaneOrder po = plane_order(A.dia, A.n, Lanes<TX>::RPB, 32 * sizeof(TX), &nb);
  switch (A.dia.nd) {
#define LACE_SPMM_DIA(ND)                                                                                \
  case ND:                                                                                               \
    spmm32_dia_kernel<TV, TX, TY, MODE, ND><<<nb, 256, 0, stream()>>>(A.n, A.dia.off, mk, dv, X, B, Y, po); \
    return true;
    LACE_SPMM_DIA(1)
    LACE_SPMM_DIA(2)
    LACE_SPMM_DIA(3)
    LACE_SPMM_DIA(4)
    LACE_SPMM_DIA(5)
    LACE_SPMM_DIA(6)
    LACE_SPMM_DIA(7)
    LACE_SPMM_DIA(8)
#undef LACE_SPMM_DIA
    default:
      return false;
  }
}

template <typename TV, typename TX, typename TY, int MODE>
static bool spmm32_ell(const DevCSR& A, const TX* X, const TY* B, TY* Y) {
  const int* ec = A.ecol.as<int>();
  const TV* ev = A.eval.as<TV>();
  int nb = 0;
  const PlaneOrder po = plane_order(A.dia, A.n, Lanes<TX>::RPB, 32 * sizeof(TX), &nb);
  switch (A.ew) {
#define LACE_SPMM_ELL(WS)                                                                                \
  case WS:                                                                                               \
    spmm32_ell_kernel<TV, TX, TY, MODE, WS><<<nb, 256, 0, stream()>>>(A.n, ec, ev, X, B, Y, po);         \
    return true;
    LACE_SPMM_ELL(1)
    LACE_SPMM_ELL(2)
    LACE_SPMM_ELL(3)
    LACE_SPMM_ELL(4)
    LACE_SPMM_ELL(5)
    LACE_SPMM_ELL(6)
    LACE_SPMM_ELL(7)
    LACE_SPMM_ELL(8)
#undef LACE_SPMM_ELL
    default:
      return false;
  }
}

template <typename TV, typename TX, typename TY>
static int spmm32_launch(const DevCSR& A, const TX* X, const TY* B, TY* Y, bool residual) {
  if (A.n == 0) return DEV_SUCCESS;
  const int nb = ceil_div(A.n, Lanes<TX>::RPB);
  const bool dia = A.dia.nd > 0 && get_option("spmv_dia", 1) != 0;
  // algorithmic work: 2 flops per nonzero and right-hand side; the matrix
  // form the launch reads (diagonal form: values and masks; else values and
  // columns, row pointers), X, Y (and B) once
  const double flops = 64.0 * A.nnz;
  const double mat = dia ? A.dia.bytes(sizeof(TV)) : (4.0 + sizeof(TV)) * A.nnz + 4.0 * (A.n + 1);
  const double bytes =
      mat + 32.0 * A.n * (sizeof(TX) + sizeof(TY)) + (residual ? 32.0 * sizeof(TY) * A.n : 0.0);
  const char* name = residual ? "spmm32_residual"
                              : (sizeof(TV) == 4 ? "spmm32_f32a" : (sizeof(TX) == 4 ? "spmm32_f32x" : "spmm32"));
  TimedRegion tr(name, flops, bytes);
  if (dia && (residual ? spmm32_dia<TV, TX, TY, 1>(A, X, B, Y) : spmm32_dia<TV, TX, TY, 0>(A, X, B, Y))) {
    LACE_HIP(hipGetLastError());
    return DEV_SUCCESS;
  }
  const bool ell = A.ew > 0 && get_option("spmv_ell", 1) != 0;
  if (ell && (residual ? spmm32_ell<TV, TX, TY, 1>(A, X, B, Y) : spmm32_ell<TV, TX, TY, 0>(A, X, B, Y))) {
    LACE_HIP(hipGetLastError());
    return DEV_SUCCESS;
  }
  if (residual)
    spmm32_kernel<TV, TX, TY, 1><<<nb, 256, 0, stream()>>>(A.n, A.row.as<int>(), A.col.as<int>(),
                                                           A.val.as<TV>(), X, B, Y);
  else
    spmm32_kernel<TV, TX, TY, 0><<<nb, 256, 0, stream()>>>(A.n, A.row.as<int>(), A.col.as<int>(),
                                                           A.val.as<TV>(), X, nullptr, Y);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

template <typename TX>
int spmm32(const DevCSR& A, const TX* X, const double* B, double* Y, bool residual) {
  LACE_REQUIRE(A.prec == LACE_F64, DEV_ERR_NOT_SUPPORTED, "spmm32: FP64 matrix values expected");
  return spmm32_launch<double, TX, double>(A, X, B, Y, residual);
}

int spmm32_f32(const DevCSR& A, const float* X, float* Y) {
  LACE_REQUIRE(A.prec == LACE_F32, DEV_ERR_NOT_SUPPORTED, "spmm32_f32: FP32 matrix values expected");
  return spmm32_launch<float, float, float>(A, X, nullptr, Y, false);
}

template int spmm32<double>(const DevCSR&, const double*, const double*, double*, bool);
template int spmm32<float>(const DevCSR&, const float*, const double*, double*, bool);

// One synchronous Jacobi sweep on 32 right-hand sides:
//   yout[i][c] = (r[i][c] - sum_{k != diag} T_ik yin[k][c]) / T_ii,
// yin == nullptr meaning yin = 0.  Upper triangles store the diagonal first
// in each row, lower ones last (the TriPlan layouts).
template <typename T, bool UPPER>
__global__ __launch_bounds__(256) void jacobi32_kernel(int n, const int* __restrict__ row,
                                                       const int* __restrict__ col,
                                                       const T* __restrict__ val,
                                                       const T* __restrict__ r,
                                                       const T* __restrict__ yin,
                                                       T* __restrict__ yout) {
#pragma clang fp contract(off)
  constexpr int W = Lanes<T>::W, LPR = Lanes<T>::LPR, RPB = Lanes<T>::RPB;
  const int i = xcd_block(blockIdx.x, gridDim.x) * RPB + (int)threadIdx.x / LPR;
  const int c0 = ((int)threadIdx.x % LPR) * W;
  if (i >= n) return;
  int k0, k1, kd;
  if (UPPER) {
    kd = row[i];
    k0 = kd + 1;
    k1 = row[i + 1];
  } else {
    k0 = row[i];
    kd = row[i + 1] - 1;
    k1 = kd;
  }
  T ri[W], s[W];
  vload<T, W>(r + (size_t)i * 32 + c0, ri);
  const T di = val[kd];
#pragma unroll
  for (int w = 0; w < W; w++) s[w] = T(0);
  if (yin) {
    for (int kb = k0; kb < k1; kb += kChunk) {
      int cc[kChunk];
      T vv[kChunk], yv[kChunk][W];
#pragma unroll
      for (int e = 0; e < kChunk; e++) {
        const bool in = kb + e < k1;
        cc[e] = in ? col[kb + e] : 0;
        vv[e] = in ? val[kb + e] : T(0);
      }
#pragma unroll
      for (int e = 0; e < kChunk; e++)
        if (kb + e < k1) vload<T, W>(yin + (size_t)cc[e] * 32 + c0, yv[e]);
#pragma unroll
      for (int e = 0; e < kChunk; e++)
        if (kb + e < k1) {
#pragma unroll
          for (int w = 0; w < W; w++) s[w] = s[w] + vv[e] * yv[e][w];
        }
    }
  }
  T out[W];
#pragma unroll
  for (int w = 0; w < W; w++) out[w] = (ri[w] - s[w]) / di;
  vstore<T, W>(yout + (size_t)i * 32 + c0, out);
}

// The sweep on the factor's off-diagonal ELLPACK form (slot-major, column -1
// for padding) with the diagonal in its own array: the row's slots are read
// once per lane group, no row pointers; same sums in the same order.
template <typename T, bool UNIT, int WS>
__global__ __launch_bounds__(256) void jacobi32_ell_kernel(int n, const int* __restrict__ ecol,
                                                           const T* __restrict__ eval,
                                                           const T* __restrict__ dval,
                                                           const T* __restrict__ r,
                                                           const T* __restrict__ yin,
                                                           T* __restrict__ yout, PlaneOrder po) {
#pragma clang fp contract(off)
  constexpr int W = Lanes<T>::W, LPR = Lanes<T>::LPR, RPB = Lanes<T>::RPB;
  const int i = plane_row<RPB>(po, xcd_block(blockIdx.x, gridDim.x), (int)threadIdx.x / LPR, n);
  const int c0 = ((int)threadIdx.x % LPR) * W;
  if (i < 0) return;
  int cc[WS];
  T vv[WS], yv[WS][W], ri[W], s[W];
#pragma unroll
  for (int q = 0; q < WS; q++) {
    cc[q] = ecol[(size_t)q * n + i];
    vv[q] = eval[(size_t)q * n + i];
  }
#pragma unroll
  for (int q = 0; q < WS; q++)
    if (cc[q] >= 0) vload<T, W>(yin + (size_t)cc[q] * 32 + c0, yv[q]);
  vload<T, W>(r + (size_t)i * 32 + c0, ri);
  const T di = UNIT ? T(1) : dval[i];
#pragma unroll
  for (int w = 0; w < W; w++) s[w] = T(0);
#pragma unroll
  for (int q = 0; q < WS; q++)
    if (cc[q] >= 0) {
#pragma unroll
      for (int w = 0; w < W; w++) s[w] = s[w] + vv[q] * yv[q][w];
    }
  T out[W];
#pragma unroll
  for (int w = 0; w < W; w++) out[w] = UNIT ? ri[w] - s[w] : (ri[w] - s[w]) / di;
  vstore<T, W>(yout + (size_t)i * 32 + c0, out);
}

// The sweep on the factor's off-diagonal diagonal form: as the product on
// A's diagonal form, one round trip per row, the same sums in the same
// order as the ELLPACK sweep.
template <typename T, bool UNIT, int ND>
__global__ __launch_bounds__(256) void jacobi32_dia_kernel(int n, DiaOffsets off,
                                                           const uint8_t* __restrict__ mask,
                                                           const T* __restrict__ oval,
                                                           const T* __restrict__ dval,
                                                           const T* __restrict__ r,
                                                           const T* __restrict__ yin,
                                                           T* __restrict__ yout, PlaneOrder po) {
#pragma clang fp contract(off)
  constexpr int W = Lanes<T>::W, LPR = Lanes<T>::LPR, RPB = Lanes<T>::RPB;
  const int i = plane_row<RPB>(po, xcd_block(blockIdx.x, gridDim.x), (int)threadIdx.x / LPR, n);
  const int c0 = ((int)threadIdx.x % LPR) * W;
  if (i < 0) return;
  const unsigned mk = mask[i];
  T vv[ND], yv[ND][W], ri[W], s[W];
#pragma unroll
  for (int d = 0; d < ND; d++) vv[d] = oval[(size_t)d * n + i];
#pragma unroll
  for (int d = 0; d < ND; d++) {
    const int c = min(max(i + off.o[d], 0), n - 1);
    vload<T, W>(yin + (size_t)c * 32 + c0, yv[d]);
  }
  vload<T, W>(r + (size_t)i * 32 + c0, ri);
  const T di = UNIT ? T(1) : dval[i];
#pragma unroll
  for (int w = 0; w < W; w++) s[w] = T(0);
#pragma unroll
  for (int d = 0; d < ND; d++)
    if ((mk >> d) & 1u) {
#pragma unroll
      for (int w = 0; w < W; w++) s[w] = s[w] + vv[d] * yv[d][w];
    }
  T out[W];
#pragma unroll
  for (int w = 0; w < W; w++) out[w] = UNIT ? ri[w] - s[w] : (ri[w] - s[w]) / di;
  vstore<T, W>(yout + (size_t)i * 32 + c0, out);
}

// Two sweeps in one launch on the diagonal form: row i's second-sweep sum
// reads the first sweep's result at rows i + off_q, which the thread forms
// itself from the previous iterate (the rows of a neighbour are formed by
// each row that reads them).  Every load is a whole 128- or 256-byte block
// row, so the repeated first-sweep reads are L2 hits, and the factor
// values, r and the previous iterate come from HBM once for two sweeps
// instead of twice, and the intermediate iterate is never written.  Each
// first-sweep row is formed exactly as jacobi32_dia_kernel forms it (the
// same sums in the same order; an entry that a mask excludes is computed
// from a clamped row and not added), so the result is bit-identical to two
// launches of that kernel.
template <typename T, bool UNIT, int ND>
__global__ __launch_bounds__(256) void jacobi32_dia_pair_kernel(int n, DiaOffsets off,
                                                                const uint8_t* __restrict__ mask,
                                                                const T* __restrict__ oval,
                                                                const T* __restrict__ dval,
                                                                const T* __restrict__ r,
                                                                const T* __restrict__ yin,
                                                                T* __restrict__ yout, PlaneOrder po) {
#pragma clang fp contract(off)
  constexpr int W = Lanes<T>::W, LPR = Lanes<T>::LPR, RPB = Lanes<T>::RPB;
  const int i = plane_row<RPB>(po, xcd_block(blockIdx.x, gridDim.x), (int)threadIdx.x / LPR, n);
  const int c0 = ((int)threadIdx.x % LPR) * W;
  if (i < 0) return;
  // every load first (they depend on i alone), so a row waits on memory
  // once: for each neighbour j = i + off_q its mask, values, r row,
  // diagonal and the previous iterate's rows j + off_p
  const unsigned mi = mask[i];
  unsigned mj[ND];
  T vq[ND], dj[ND], vp[ND][ND], rj[ND][W], yk[ND][ND][W], ri[W];
#pragma unroll
  for (int q = 0; q < ND; q++) {
    const int j = min(max(i + off.o[q], 0), n - 1);
    mj[q] = mask[j];
    vq[q] = oval[(size_t)q * n + i];
    dj[q] = UNIT ? T(1) : dval[j];
    vload<T, W>(r + (size_t)j * 32 + c0, rj[q]);
#pragma unroll
    for (int p = 0; p < ND; p++) {
      const int k = min(max(j + off.o[p], 0), n - 1);
      vp[q][p] = oval[(size_t)p * n + j];
      vload<T, W>(yin + (size_t)k * 32 + c0, yk[q][p]);
    }
  }
  vload<T, W>(r + (size_t)i * 32 + c0, ri);
  const T di = UNIT ? T(1) : dval[i];
  T s[W];
#pragma unroll
  for (int w = 0; w < W; w++) s[w] = T(0);
#pragma unroll
  for (int q = 0; q < ND; q++) {
    // the first sweep at row j = i + off_q
    T t[W];
#pragma unroll
    for (int w = 0; w < W; w++) t[w] = T(0);
#pragma unroll
    for (int p = 0; p < ND; p++) {
      const bool on = (mj[q] >> p) & 1u;
#pragma unroll
      for (int w = 0; w < W; w++) {
        const T a = t[w] + vp[q][p] * yk[q][p][w];
        t[w] = on ? a : t[w];
      }
    }
    const bool onq = (mi >> q) & 1u;
#pragma unroll
    for (int w = 0; w < W; w++) {
      const T y1 = UNIT ? rj[q][w] - t[w] : (rj[q][w] - t[w]) / dj[q];
      const T b = s[w] + vq[q] * y1;
      s[w] = onq ? b : s[w];
    }
  }
  T out[W];
#pragma unroll
  for (int w = 0; w < W; w++) out[w] = UNIT ? ri[w] - s[w] : (ri[w] - s[w]) / di;
  vstore<T, W>(yout + (size_t)i * 32 + c0, out);
}

// Whether two sweeps of P with the value array val can share a launch:
// the diagonal form is current for val (the 32-column sweeps would read it).
template <typename T>
static bool jacobi32_pair_applies(const TriPlan& P, const T* val) {
  return P.dia.nd > 0 && P.strict_src == (const void*)val && !(P.stale & TriPlan::kFormDia) &&
         get_option("jacobi_dia", 1) != 0 && get_option("jacobi_pair", 1) != 0;
}

// yout = two sweeps from yin (yin, r and yout distinct).
template <typename T>
static int jacobi32_pair(const TriPlan& P, const T* r, const T* yin, T* yout) {
  // algorithmic work: the diagonal form (and the diagonal unless it is a
  // unit one), r and the previous iterate once, the result once; flops of
  // two sweeps
  const double offd = (double)P.nnz - P.n;
  const double bytes = P.dia.bytes(sizeof(T)) + (P.unit_diag ? 0.0 : sizeof(T) * (double)P.n) +
                       32.0 * sizeof(T) * P.n * 3.0;
  TimedRegion tr(P.upper ? "block_jacobi_pair_upper" : "block_jacobi_pair_lower", 2.0 * (64.0 * offd + 32.0 * P.n),
                 bytes);
  const uint8_t* mk = P.dia.mask.as<uint8_t>();
  const T* ov = P.dia.val.as<T>();
  const T* dv = P.dval.as<T>();
  int nb = 0;
  const PlaneOrder po = plane_order(P.dia, P.n, Lanes<T>::RPB, 32 * sizeof(T), &nb);
  switch (P.dia.nd) {
#define LACE_J32_PAIR(ND)                                                                                    \
  case ND:                                                                                                   \
    if (P.unit_diag)                                                                                         \
      jacobi32_dia_pair_kernel<T, true, ND><<<nb, 256, 0, stream()>>>(P.n, P.dia.off, mk, ov, dv, r, yin, yout, po); \
    else                                                                                                     \
      jacobi32_dia_pair_kernel<T, false, ND><<<nb, 256, 0, stream()>>>(P.n, P.dia.off, mk, ov, dv, r, yin, yout, po); \
    break;
    LACE_J32_PAIR(1)
    LACE_J32_PAIR(2)
    LACE_J32_PAIR(3)
    LACE_J32_PAIR(4)
    LACE_J32_PAIR(5)
    LACE_J32_PAIR(6)
    LACE_J32_PAIR(7)
    LACE_J32_PAIR(8)
#undef LACE_J32_PAIR
    default:
      break;
  }
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

template <typename T>
static int jacobi32_sweep(const TriPlan& P, const T* val, const T* r, const T* yin, T* yout) {
  const int nb = ceil_div(P.n, Lanes<T>::RPB);
  const bool dia = yin && P.dia.nd > 0 && P.strict_src == (const void*)val && !(P.stale & TriPlan::kFormDia) &&
                   get_option("jacobi_dia", 1) != 0;
  // algorithmic work: the off-diagonal form the launch reads (diagonal
  // form: values and masks, and the diagonal unless it is a unit one; else
  // values and columns, the diagonal and row pointers), r, the previous
  // iterate (when read) and the new one
  const double offd = (double)P.nnz - P.n;
  const double mat = dia ? P.dia.bytes(sizeof(T)) + (P.unit_diag ? 0.0 : sizeof(T) * (double)P.n)
                         : (4.0 + sizeof(T)) * (yin ? offd : 0.0) + sizeof(T) * (double)P.n + 4.0 * (P.n + 1);
  const double bytes = mat + 32.0 * sizeof(T) * P.n * (yin ? 3.0 : 2.0);
  TimedRegion tr(P.upper ? "block_jacobi_sweep_upper" : "block_jacobi_sweep_lower",
                 64.0 * (yin ? offd : 0.0) + 32.0 * P.n, bytes);
  if (dia) {
    const uint8_t* mk = P.dia.mask.as<uint8_t>();
    const T* ov = P.dia.val.as<T>();
    const T* dv = P.dval.as<T>();
    int nbd = 0;
    const PlaneOrder po = plane_order(P.dia, P.n, Lanes<T>::RPB, 32 * sizeof(T), &nbd);
    switch (P.dia.nd) {
#define LACE_J32_DIA(ND)                                                                                     \
  case ND:                                                                                                   \
    if (P.unit_diag)                                                                                         \
      jacobi32_dia_kernel<T, true, ND><<<nbd, 256, 0, stream()>>>(P.n, P.dia.off, mk, ov, dv, r, yin, yout, po); \
    else                                                                                                     \
      jacobi32_dia_kernel<T, false, ND><<<nbd, 256, 0, stream()>>>(P.n, P.dia.off, mk, ov, dv, r, yin, yout, po); \
    break;
      LACE_J32_DIA(1)
      LACE_J32_DIA(2)
      LACE_J32_DIA(3)
      LACE_J32_DIA(4)
      LACE_J32_DIA(5)
      LACE_J32_DIA(6)
      LACE_J32_DIA(7)
      LACE_J32_DIA(8)
#undef LACE_J32_DIA
      default:
        break;
    }
    LACE_HIP(hipGetLastError());
    return DEV_SUCCESS;
  }
  if (yin && P.ew > 0 && P.strict_src == (const void*)val && !(P.stale & TriPlan::kFormEll) &&
      get_option("jacobi_ell", 1) != 0) {
    const int* ec = P.ecol.as<int>();
    const T* ev = P.eval.as<T>();
    const T* dv = P.dval.as<T>();
    int nbe = 0;
    const PlaneOrder po = plane_order(P.dia, P.n, Lanes<T>::RPB, 32 * sizeof(T), &nbe);
    switch (P.ew) {
#define LACE_J32_ELL(W)                                                                                      \
  case W:                                                                                                    \
    if (P.unit_diag)                                                                                         \
      jacobi32_ell_kernel<T, true, W><<<nbe, 256, 0, stream()>>>(P.n, ec, ev, dv, r, yin, yout, po);         \
    else                                                                                                     \
      jacobi32_ell_kernel<T, false, W><<<nbe, 256, 0, stream()>>>(P.n, ec, ev, dv, r, yin, yout, po);        \
    break;
      LACE_J32_ELL(1)
      LACE_J32_ELL(2)
      LACE_J32_ELL(3)
      LACE_J32_ELL(4)
      LACE_J32_ELL(5)
      LACE_J32_ELL(6)
      LACE_J32_ELL(7)
      LACE_J32_ELL(8)
#undef LACE_J32_ELL
      default:
        break;
    }
    LACE_HIP(hipGetLastError());
    return DEV_SUCCESS;
  }
  if (P.upper)
    jacobi32_kernel<T, true><<<nb, 256, 0, stream()>>>(P.n, P.row.as<int>(), P.col.as<int>(), val, r,
                                                       yin, yout);
  else
    jacobi32_kernel<T, false><<<nb, 256, 0, stream()>>>(P.n, P.row.as<int>(), P.col.as<int>(), val, r,
                                                        yin, yout);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

// `sweeps` Jacobi sweeps from y = 0 -- the same sequence as the
// single-vector solve, including the unit-diagonal shortcut (the first sweep
// from zero returns r exactly, so the second reads r).
template <typename T>
static int jacobi32_solve(const TriPlan& P, const T* val, const T* r, T* y, T* tmp, int sweeps) {
  const size_t bytes = sizeof(T) * (size_t)P.n * kNrhs;
  if (sweeps <= 0) {
    LACE_HIP(hipMemsetAsync(y, 0, bytes, stream()));
    return DEV_SUCCESS;
  }
  if (P.unit_diag && sweeps == 1) {
    LACE_HIP(hipMemcpyAsync(y, r, bytes, hipMemcpyDeviceToDevice, stream()));
    return DEV_SUCCESS;
  }
  // The first sweep from zero is r (unit diagonal: no launch) or r / diag;
  // the K sweeps after it read the previous iterate and run two to a
  // launch where they can (a single one last when K is odd).  Launch idx
  // of L writes y when L - 1 - idx is even, tmp otherwise, so the last
  // lands in y and no launch reads its own output.
  const int K = sweeps - 1;
  const bool pair = jacobi32_pair_applies<T>(P, val);
  const int L = pair ? (K + 1) / 2 : K;
  auto out_of = [&](int idx) { return ((L - 1 - idx) % 2 == 0) ? y : tmp; };
  const T* src = r;
  if (!P.unit_diag) {
    T* first = L == 0 ? y : (out_of(0) == y ? tmp : y);
    LACE_TRY(jacobi32_sweep<T>(P, val, r, nullptr, first));
    src = first;
  }
  int left = K;
  for (int idx = 0; idx < L; idx++) {
    T* o = out_of(idx);
    if (pair && left >= 2) {
      LACE_TRY(jacobi32_pair<T>(P, r, src, o));
      left -= 2;
    } else {
      LACE_TRY(jacobi32_sweep<T>(P, val, r, src, o));
      left -= 1;
    }
    src = o;
  }
  return DEV_SUCCESS;
}

template <typename T>
int block_precond(ParILU* P, lace_trsv_kind_t kind, int sweeps, int trans, const T* R, T* Z, T* tmp,
                  T* tmp2) {
  LACE_REQUIRE((P->prec == LACE_F64) == (sizeof(T) == 8), DEV_ERR_ILLEGAL_VALUE,
               "block_precond: vector precision differs from the factors'");
  if (kind == LACE_TRSV_NONE) {
    LACE_HIP(hipMemcpyAsync(Z, R, sizeof(T) * (size_t)P->n * kNrhs, hipMemcpyDeviceToDevice, stream()));
    return DEV_SUCCESS;
  }
  LACE_REQUIRE(kind == LACE_TRSV_JACOBI, DEV_ERR_NOT_SUPPORTED,
               "block solves take Jacobi (or no) triangular solves");
  LACE_TRY(P->prepare(kind, trans, true));
  const TriPlan& first = trans ? P->lowerT : P->lower;
  const TriPlan& second = trans ? P->upperT : P->upper;
  const T* v1 = trans ? P->uval[P->cur].as<T>() : P->lval[P->cur].as<T>();
  const T* v2 = trans ? P->lval_t.as<T>() : P->uval_csr.as<T>();
  TimedRegion tr("block_precond");
  LACE_TRY(jacobi32_solve<T>(first, v1, R, tmp, tmp2, sweeps));
  LACE_TRY(jacobi32_solve<T>(second, v2, tmp, Z, tmp2, sweeps));
  return DEV_SUCCESS;
}

template int block_precond<double>(ParILU*, lace_trsv_kind_t, int, int, const double*, double*,
                                   double*, double*);
template int block_precond<float>(ParILU*, lace_trsv_kind_t, int, int, const float*, float*, float*,
                                  float*);

__global__ void block_cast_d2f_kernel(size_t m, const double* __restrict__ in, float* __restrict__ out) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e < m) out[e] = (float)in[e];
}

int block_cast_d2f(int n, const double* in, float* out) {
  const size_t m = (size_t)n * kNrhs;
  if (m == 0) return DEV_SUCCESS;
  block_cast_d2f_kernel<<<ceil_div((long)m, 256), 256, 0, stream()>>>(m, in, out);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

}  // namespace lace

using namespace lace;

extern "C" {

int lace_csr_spmm32(const lace_csr* Ah, const void* X, lace_prec_t x_prec, const double* B, double* Y) {
  LACE_TRY(ensure_init());
  const DevCSR& A = *reinterpret_cast<const DevCSR*>(Ah);
  if (x_prec == LACE_F64) return spmm32<double>(A, static_cast<const double*>(X), B, Y, B != nullptr);
  LACE_REQUIRE(x_prec == LACE_F32, DEV_ERR_ILLEGAL_VALUE, "spmm32: X precision %d", (int)x_prec);
  return spmm32<float>(A, static_cast<const float*>(X), B, Y, B != nullptr);
}

int lace_parilu_apply_block(lace_parilu* Ph, lace_trsv_kind_t kind, int jacobi_sweeps, int trans,
                            const void* R, void* Z) {
  LACE_TRY(ensure_init());
  auto* P = reinterpret_cast<ParILU*>(Ph);
  const size_t elems = (size_t)P->n * kNrhs;
  DevBuf t1, t2;
  const size_t es = P->prec == LACE_F64 ? sizeof(double) : sizeof(float);
  LACE_TRY(t1.alloc(elems * es));
  LACE_TRY(t2.alloc(elems * es));
  int rc;
  if (P->prec == LACE_F64)
    rc = block_precond<double>(P, kind, jacobi_sweeps, trans, static_cast<const double*>(R),
                               static_cast<double*>(Z), t1.as<double>(), t2.as<double>());
  else
    rc = block_precond<float>(P, kind, jacobi_sweeps, trans, static_cast<const float*>(R),
                              static_cast<float*>(Z), t1.as<float>(), t2.as<float>());
  LACE_TRY(rc);
  LACE_HIP(hipStreamSynchronize(stream()));
  return DEV_SUCCESS;
}

}  // extern "C"

namespace lace {

__global__ void daxpby_kernel(size_t m, double a, const double* __restrict__ x, double b,
                              const double* __restrict__ y, double* __restrict__ out) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= m) return;
  double v = 0.0;
  if (a != 0.0) v = a * x[e];
  if (b != 0.0) v = v + b * y[e];
  out[e] = v;
}

}  // namespace lace

extern "C" {

int lace_daxpby(size_t m, double a, const double* x, double b, const double* y, double* out) {
  LACE_TRY(ensure_init());
  if (m == 0) return DEV_SUCCESS;
  daxpby_kernel<<<ceil_div((long)m, 256), 256, 0, stream()>>>(m, a, x, b, y, out);
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

}  // extern "C"
