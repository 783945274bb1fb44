// host_formats.cpp -- host-side containers, format conversions, MatrixMarket
// I/O, norms, the sparse x sparse product, exact ILU(0), level sets and the
// synthetic problem generators.  These are set-up and reporting routines of
// the drop-in ABI; the hot path itself (sweeps, triangular solves, SpMV,
// FGMRES) runs in the HIP translation units.
#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

#include "common.hpp"

using namespace lace;

namespace {

void free_arrays(data_d_matrix* A) {
  free(A->val); A->val = nullptr;
  free(A->diag); A->diag = nullptr;
  free(A->row); A->row = nullptr;
  free(A->rowidx); A->rowidx = nullptr;
  free(A->col); A->col = nullptr;
  free(A->list); A->list = nullptr;
  free(A->blockinfo); A->blockinfo = nullptr;
}

bool is_csr_like(data_storage_t s) {
  return s == Magma_CSR || s == Magma_CSRL || s == Magma_CSRU || s == Magma_CSRD ||
         s == Magma_CSRCOO || s == Magma_CUCSR;
}
bool is_csc_like(data_storage_t s) {
  return s == Magma_CSC || s == Magma_CSCL || s == Magma_CSCU || s == Magma_CSCD ||
         s == Magma_CSCCOO;
}
bool is_dense(data_storage_t s) {
  return s == Magma_DENSE || s == Magma_DENSEL || s == Magma_DENSEU || s == Magma_DENSED;
}

// Lower/upper split of control/magma_zmconverter.cpp:303-460.  Rows are
// assumed sorted, so CSRL keeps the diagonal last and CSRU first.
int split(const data_d_matrix& A, data_d_matrix* B, bool upper) {
  const int n = A.num_rows;
  const data_diagorder_t dmode = B->diagorder_type;
  int cnt = 0;
  for (int i = 0; i < n; i++)
    for (int k = A.row[i]; k < A.row[i + 1]; k++) {
      int j = A.col[k];
      if ((upper && j > i) || (!upper && j < i)) cnt++;
      else if (j == i && dmode != Magma_NODIAG) cnt++;
    }
  B->storage_type = upper ? Magma_CSRU : Magma_CSRL;
  B->major = MagmaRowMajor;
  B->fill_mode = upper ? MagmaUpper : MagmaLower;
  B->num_rows = A.num_rows;
  B->num_cols = A.num_cols;
  B->pad_rows = A.pad_rows;
  B->pad_cols = A.pad_cols;
  B->diameter = A.diameter;
  B->nnz = cnt;
  B->true_nnz = cnt;
  LACE_TRY(host_calloc(&B->val, cnt));
  LACE_TRY(host_calloc(&B->row, n + 1));
  LACE_TRY(host_calloc(&B->col, cnt));
  int z = 0;
  for (int i = 0; i < n; i++) {
    B->row[i] = z;
    for (int k = A.row[i]; k < A.row[i + 1]; k++) {
      int j = A.col[k];
      if ((upper && j > i) || (!upper && j < i)) {
        B->val[z] = A.val[k];
        B->col[z++] = j;
      } else if (j == i && dmode != Magma_NODIAG) {
        // UNITY -> 1, ZERO -> 0, VALUE (and any other order tag) -> a_ii.
        double v = dmode == Magma_UNITY ? 1.0 : (dmode == Magma_ZERO ? 0.0 : A.val[k]);
        B->val[z] = v;
        B->col[z++] = j;
      }
    }
  }
  B->row[n] = z;
  B->nnz = z;
  return DEV_SUCCESS;
}

// CSR -> CSC by counting sort (control/magma_zmtranspose.cpp:85-235).
int transpose_csr(const data_d_matrix& A, data_d_matrix* B) {
  const int m = A.num_rows, n = A.num_cols, nnz = A.row[m];
  data_int_t *row = nullptr, *col = nullptr;
  dataType* val = nullptr;
  LACE_TRY(host_calloc(&row, n + 1));
  LACE_TRY(host_calloc(&col, nnz));
  LACE_TRY(host_calloc(&val, nnz));
  for (int k = 0; k < nnz; k++) row[A.col[k] + 1]++;
  for (int j = 0; j < n; j++) row[j + 1] += row[j];
  std::vector<int> fill(row, row + n);
  for (int i = 0; i < m; i++)
    for (int k = A.row[i]; k < A.row[i + 1]; k++) {
      int p = fill[A.col[k]]++;
      col[p] = i;
      val[p] = A.val ? A.val[k] : 0.0;
    }
  B->row = row;
  B->col = col;
  B->val = val;
  B->nnz = nnz;
  B->true_nnz = A.true_nnz;
  B->fill_mode = A.fill_mode;
  B->num_rows = n;
  B->num_cols = m;
  B->pad_rows = A.pad_cols;
  B->pad_cols = A.pad_rows;
  B->storage_type = A.storage_type;
  return DEV_SUCCESS;
}

bool is_bcsr_like(data_storage_t s) {
  return s == Magma_BCSR || s == Magma_BCSRD || s == Magma_BCSRL || s == Magma_BCSRU ||
         s == Magma_BCSC || s == Magma_BCSCD || s == Magma_BCSCL || s == Magma_BCSCU;
}

void copy_block_header(const data_d_matrix& A, data_d_matrix* B) {
  B->num_rows = A.num_rows;
  B->num_cols = A.num_cols;
  B->pad_rows = A.pad_rows;
  B->pad_cols = A.pad_cols;
  B->diameter = A.diameter;
  B->blocksize = A.blocksize;
  B->ldblock = A.ldblock;
}

// CSR -> BCSR with bs = B->blocksize (the reference calls mkl_dcsrbsr with
// zero-based indexing, which keeps blocks row-major,
// control/magma_zmconverter.cpp:639-677): a block is stored when any entry
// of A falls in it; block columns ascending; absent entries are zero.
int csr_to_bcsr(const data_d_matrix& A, data_d_matrix* B) {
  const int bs = B->blocksize;
  LACE_REQUIRE(bs > 0, DEV_ERR_NOT_SUPPORTED,
               "data_zmconvert: conversion to BCSR requires blocksize > 0 (have %d)", bs);
  const int nbr = (A.num_rows + bs - 1) / bs, nbc = (A.num_cols + bs - 1) / bs;
  const int lb = bs * bs;
  std::vector<int> brow(nbr + 1, 0), bcol, mark(nbc, -1);
  for (int I = 0; I < nbr; I++) {
    brow[I] = (int)bcol.size();
    const size_t start = bcol.size();
    for (int i = I * bs; i < std::min(A.num_rows, (I + 1) * bs); i++)
      for (int k = A.row[i]; k < A.row[i + 1]; k++) {
        const int J = A.col[k] / bs;
        if (mark[J] != I) {
          mark[J] = I;
          bcol.push_back(J);
        }
      }
    std::sort(bcol.begin() + start, bcol.end());
  }
  brow[nbr] = (int)bcol.size();
  const int nb = brow[nbr];
  B->storage_type = Magma_BCSR;
  B->major = MagmaRowMajor;
  B->fill_mode = A.fill_mode;
  B->num_rows = nbr;
  B->num_cols = nbc;
  B->ldblock = lb;
  B->numblocks = nb;
  B->nnz = nb * lb;
  B->true_nnz = A.nnz;
  LACE_TRY(host_calloc(&B->row, nbr + 1));
  LACE_TRY(host_calloc(&B->col, std::max(nb, 1)));
  LACE_TRY(host_calloc(&B->val, (size_t)std::max(nb, 1) * lb));
  memcpy(B->row, brow.data(), sizeof(int) * (nbr + 1));
  if (nb) memcpy(B->col, bcol.data(), sizeof(int) * nb);
  for (int I = 0; I < nbr; I++)
    for (int i = I * bs; i < std::min(A.num_rows, (I + 1) * bs); i++)
      for (int k = A.row[i]; k < A.row[i + 1]; k++) {
        const int J = A.col[k] / bs;
        const int* it = std::lower_bound(B->col + brow[I], B->col + brow[I + 1], J);
        const size_t q = (size_t)(it - B->col);
        B->val[q * lb + (i - I * bs) * bs + (A.col[k] - J * bs)] += A.val[k];
      }
  return DEV_SUCCESS;
}

// BCSR -> BCSRL / BCSRU (control/magma_zmconverter.cpp:1248-1436): strict
// lower (upper) blocks and the diagonal block per B->diagorder_type --
// UNITY: identity, ZERO: zeros, NODIAG: omitted, any other tag: A's block.
// BCSRL comes back tagged Magma_BCSR as in the reference (:1251).
int bcsr_split(const data_d_matrix& A, data_d_matrix* B, bool upper) {
  const int nbr = A.num_rows, bs = A.blocksize, lb = bs * bs;
  const data_diagorder_t dmode = B->diagorder_type;
  int cnt = 0;
  for (int I = 0; I < nbr; I++)
    for (int k = A.row[I]; k < A.row[I + 1]; k++) {
      const int J = A.col[k];
      if ((upper && J > I) || (!upper && J < I) || (J == I && dmode != Magma_NODIAG)) cnt++;
    }
  copy_block_header(A, B);
  B->storage_type = upper ? Magma_BCSRU : Magma_BCSR;
  B->major = MagmaRowMajor;
  B->fill_mode = upper ? MagmaUpper : MagmaLower;
  B->ldblock = lb;
  B->numblocks = cnt;
  B->nnz = cnt * lb;
  LACE_TRY(host_calloc(&B->row, nbr + 1));
  LACE_TRY(host_calloc(&B->col, std::max(cnt, 1)));
  LACE_TRY(host_calloc(&B->val, (size_t)std::max(cnt, 1) * lb));
  int z = 0;
  for (int I = 0; I < nbr; I++) {
    B->row[I] = z;
    for (int k = A.row[I]; k < A.row[I + 1]; k++) {
      const int J = A.col[k];
      const bool off = (upper && J > I) || (!upper && J < I);
      if (!off && !(J == I && dmode != Magma_NODIAG)) continue;
      double* dst = B->val + (size_t)z * lb;
      if (off || (dmode != Magma_UNITY && dmode != Magma_ZERO)) {
        memcpy(dst, A.val + (size_t)k * lb, sizeof(double) * lb);
      } else if (dmode == Magma_UNITY) {
        for (int a = 0; a < bs; a++) dst[a * bs + a] = 1.0;
      }
      B->col[z++] = J;
    }
  }
  B->row[nbr] = z;
  return DEV_SUCCESS;
}

// Block transpose by counting sort: blocks move, their interiors are kept
// as stored (control/magma_zmtranspose.cpp:237-380).
int transpose_bcsr(const data_d_matrix& A, data_d_matrix* B) {
  const int m = A.num_rows, n = A.num_cols, nb = A.row[m], lb = A.blocksize * A.blocksize;
  data_int_t *row = nullptr, *col = nullptr;
  dataType* val = nullptr;
  LACE_TRY(host_calloc(&row, n + 1));
  LACE_TRY(host_calloc(&col, std::max(nb, 1)));
  LACE_TRY(host_calloc(&val, (size_t)std::max(nb, 1) * lb));
  for (int k = 0; k < nb; k++) row[A.col[k] + 1]++;
  for (int j = 0; j < n; j++) row[j + 1] += row[j];
  std::vector<int> fill(row, row + n);
  for (int i = 0; i < m; i++)
    for (int k = A.row[i]; k < A.row[i + 1]; k++) {
      const int p = fill[A.col[k]]++;
      col[p] = i;
      memcpy(val + (size_t)p * lb, A.val + (size_t)k * lb, sizeof(double) * lb);
    }
  copy_block_header(A, B);
  B->num_rows = n;
  B->num_cols = m;
  B->pad_rows = A.pad_cols;
  B->pad_cols = A.pad_rows;
  B->row = row;
  B->col = col;
  B->val = val;
  B->numblocks = nb;
  B->nnz = A.nnz;
  B->true_nnz = A.true_nnz;
  B->fill_mode = A.fill_mode;
  B->storage_type = A.storage_type;
  return DEV_SUCCESS;
}

int copy_bcsr(const data_d_matrix& A, data_d_matrix* B) {
  const int nbr = A.num_rows, nb = A.row[nbr], lb = A.blocksize * A.blocksize;
  copy_block_header(A, B);
  B->storage_type = Magma_BCSR;
  B->major = MagmaRowMajor;
  B->fill_mode = A.fill_mode;
  B->ldblock = lb;
  B->numblocks = nb;
  B->nnz = nb * lb;
  B->true_nnz = A.true_nnz;
  LACE_TRY(host_calloc(&B->row, nbr + 1));
  LACE_TRY(host_calloc(&B->col, std::max(nb, 1)));
  LACE_TRY(host_calloc(&B->val, (size_t)std::max(nb, 1) * lb));
  memcpy(B->row, A.row, sizeof(int) * (nbr + 1));
  memcpy(B->col, A.col, sizeof(int) * nb);
  memcpy(B->val, A.val, sizeof(double) * (size_t)nb * lb);
  return DEV_SUCCESS;
}

int copy_csr(const data_d_matrix& A, data_d_matrix* B) {
  const int n = A.num_rows, nnz = A.row[n];
  B->storage_type = Magma_CSR;
  B->major = MagmaRowMajor;
  B->fill_mode = A.fill_mode;
  B->num_rows = A.num_rows;
  B->num_cols = A.num_cols;
  B->pad_rows = A.pad_rows;
  B->pad_cols = A.pad_cols;
  B->nnz = nnz;
  B->true_nnz = A.true_nnz;
  B->max_nnz_row = A.max_nnz_row;
  B->diameter = A.diameter;
  LACE_TRY(host_calloc(&B->val, nnz));
  LACE_TRY(host_calloc(&B->row, n + 1));
  LACE_TRY(host_calloc(&B->col, nnz));
  memcpy(B->val, A.val, sizeof(double) * nnz);
  memcpy(B->col, A.col, sizeof(int) * nnz);
  memcpy(B->row, A.row, sizeof(int) * (n + 1));
  return DEV_SUCCESS;
}

int csr_to_dense(const data_d_matrix& A, data_d_matrix* B) {
  B->storage_type = Magma_DENSE;
  B->major = MagmaRowMajor;
  B->fill_mode = A.fill_mode;
  B->num_rows = A.num_rows;
  B->num_cols = A.num_cols;
  B->pad_rows = A.pad_rows;
  B->pad_cols = A.pad_cols;
  B->nnz = A.num_rows * A.num_cols;
  B->true_nnz = A.true_nnz;
  B->ld = A.num_cols;
  LACE_TRY(host_calloc(&B->val, B->nnz));
  for (int i = 0; i < A.num_rows; i++)
    for (int k = A.row[i]; k < A.row[i + 1]; k++) B->val[i * B->num_cols + A.col[k]] = A.val[k];
  return DEV_SUCCESS;
}

int dense_copy(const data_d_matrix& A, data_d_matrix* B) {
  B->storage_type = Magma_DENSE;
  B->num_rows = A.num_rows;
  B->num_cols = A.num_cols;
  B->pad_rows = A.pad_rows;
  B->pad_cols = A.pad_cols;
  B->nnz = A.nnz;
  B->true_nnz = A.true_nnz;
  B->ld = A.ld;
  B->fill_mode = A.fill_mode;
  if (B->major != MagmaRowMajor && B->major != MagmaColMajor) B->major = A.major;
  const size_t cnt = (size_t)A.num_rows * A.num_cols;
  LACE_TRY(host_calloc(&B->val, cnt > (size_t)A.nnz ? cnt : (size_t)A.nnz));
  const int ld = A.ld > 0 ? A.ld : (A.major == MagmaRowMajor ? A.num_cols : A.num_rows);
  for (int i = 0; i < A.num_rows; i++)
    for (int j = 0; j < A.num_cols; j++) {
      double v = A.major == MagmaRowMajor ? A.val[(size_t)i * ld + j] : A.val[i + (size_t)j * ld];
      if (B->major == MagmaRowMajor) B->val[(size_t)i * ld + j] = v;
      else B->val[i + (size_t)j * ld] = v;
    }
  return DEV_SUCCESS;
}

}  // namespace

extern "C" {

// control/init.cpp:5-47
int data_zvinit(data_d_matrix* x, int num_rows, int num_cols, dataType value,
                data_order_t major) {
  LACE_REQUIRE(x != nullptr, DEV_ERR_INVALID_PTR, "data_zvinit: null matrix");
  free_arrays(x);
  x->storage_type = Magma_DENSE;
  x->sym = Magma_GENERAL;
  x->diagorder_type = Magma_VALUE;
  x->fill_mode = MagmaFull;
  x->num_rows = num_rows;
  x->num_cols = num_cols;
  x->pad_rows = 0;
  x->pad_cols = 0;
  x->nnz = num_rows * num_cols;
  x->max_nnz_row = num_cols;
  x->diameter = 0;
  x->blocksize = 1;
  x->numblocks = 1;
  x->alignment = 1;
  x->major = major;
  x->ld = num_rows;
  LACE_TRY(host_calloc(&x->val, x->nnz));
  for (int i = 0; i < x->nnz; i++) x->val[i] = value;
  return DEV_SUCCESS;
}

// control/magma_zfree.cpp:37 -- releases every array and zeroes the sizes.
int data_zmfree(data_d_matrix* A) {
  if (!A) return DEV_SUCCESS;
  free_arrays(A);
  A->num_rows = 0;
  A->num_cols = 0;
  A->pad_rows = 0;
  A->pad_cols = 0;
  A->nnz = 0;
  A->true_nnz = 0;
  A->diameter = 0;
  return DEV_SUCCESS;
}

// control/magma_zmconverter.cpp:145-170
int data_rowindex(data_d_matrix* A, int** rowidx) {
  if (*rowidx) {
    free(*rowidx);
    *rowidx = nullptr;
  }
  LACE_TRY(host_calloc(rowidx, A->nnz));
  int lim = (A->pad_rows > 0 && A->pad_cols > 0) ? A->pad_rows : A->num_rows;
  for (int i = 0; i < lim; i++)
    for (int k = A->row[i]; k < A->row[i + 1]; k++) (*rowidx)[k] = i;
  return DEV_SUCCESS;
}

int data_zmtranspose(data_d_matrix A, data_d_matrix* B) {
  if (is_csr_like(A.storage_type) || is_csc_like(A.storage_type)) return transpose_csr(A, B);
  if (is_bcsr_like(A.storage_type)) return transpose_bcsr(A, B);
  if (is_dense(A.storage_type)) {
    LACE_TRY(dense_copy(A, B));
    // transpose in place of the copy
    std::vector<double> t((size_t)A.num_rows * A.num_cols);
    for (int i = 0; i < A.num_rows; i++)
      for (int j = 0; j < A.num_cols; j++)
        t[(size_t)j * A.num_rows + i] = A.val[(size_t)i * A.num_cols + j];
    memcpy(B->val, t.data(), sizeof(double) * t.size());
    std::swap(B->num_rows, B->num_cols);
    return DEV_SUCCESS;
  }
  set_error("data_zmtranspose: storage %d not supported", (int)A.storage_type);
  return DEV_ERR_NOT_SUPPORTED;
}

// control/magma_zmconverter.cpp:199-1445 (the CSR/CSC/dense branches the
// hot path uses).  The destination's arrays are released first (:221-248).
int data_zmconvert(data_d_matrix A, data_d_matrix* B, data_storage_t old_format,
                   data_storage_t new_format) {
  LACE_REQUIRE(B != nullptr, DEV_ERR_INVALID_PTR, "data_zmconvert: null destination");
  free_arrays(B);
  if (old_format == Magma_CSR || old_format == Magma_CSRL || old_format == Magma_CSRU) {
    switch (new_format) {
      case Magma_CSR: return copy_csr(A, B);
      case Magma_CUCSR: LACE_TRY(copy_csr(A, B)); B->storage_type = Magma_CUCSR; return DEV_SUCCESS;
      case Magma_CSRL: return split(A, B, false);
      case Magma_CSRU: return split(A, B, true);
      case Magma_CSRCOO: {
        LACE_TRY(copy_csr(A, B));
        B->storage_type = Magma_CSRCOO;
        return data_rowindex(B, &B->rowidx);
      }
      case Magma_DENSE: return csr_to_dense(A, B);
      case Magma_BCSR: return csr_to_bcsr(A, B);
      case Magma_CSC: {
        LACE_TRY(transpose_csr(A, B));
        B->storage_type = Magma_CSC;
        B->major = MagmaColMajor;
        return DEV_SUCCESS;
      }
      case Magma_CSCL:
      case Magma_CSCU: {
        data_d_matrix C;
        memset(&C, 0, sizeof(C));
        C.diagorder_type = B->diagorder_type;
        LACE_TRY(split(A, &C, new_format == Magma_CSCU));
        int r = transpose_csr(C, B);
        data_zmfree(&C);
        LACE_TRY(r);
        B->storage_type = new_format;
        B->major = MagmaColMajor;
        return DEV_SUCCESS;
      }
      default: break;
    }
  } else if (is_csc_like(old_format)) {
    // CSC viewed as CSR of the transpose: converting back to row storage is
    // a transpose; splits of a CSC matrix split its stored transpose
    // (control/magma_zmconverter.cpp:926-981).
    switch (new_format) {
      case Magma_CSR:
      case Magma_CSRU:
      case Magma_CSRL: {
        LACE_TRY(transpose_csr(A, B));
        B->storage_type = new_format;
        B->major = MagmaRowMajor;
        return DEV_SUCCESS;
      }
      case Magma_CSCL:
      case Magma_CSCU: {
        LACE_TRY(split(A, B, new_format == Magma_CSCU));
        B->storage_type = new_format;
        B->major = MagmaColMajor;
        return DEV_SUCCESS;
      }
      default: break;
    }
  } else if (old_format == Magma_BCSR) {
    // control/magma_zmconverter.cpp:1248-1445
    switch (new_format) {
      case Magma_BCSR: return copy_bcsr(A, B);
      case Magma_BCSRL: return bcsr_split(A, B, false);
      case Magma_BCSRU: return bcsr_split(A, B, true);
      case Magma_BCSCU: {
        data_d_matrix C;
        memset(&C, 0, sizeof(C));
        C.diagorder_type = B->diagorder_type;
        LACE_TRY(bcsr_split(A, &C, true));
        const int r = transpose_bcsr(C, B);
        data_zmfree(&C);
        LACE_TRY(r);
        B->storage_type = Magma_BCSCU;
        B->major = MagmaColMajor;
        return DEV_SUCCESS;
      }
      default: break;
    }
  } else if (is_dense(old_format) && is_dense(new_format)) {
    return dense_copy(A, B);
  }
  set_error("data_zmconvert: format %d to %d not supported", (int)old_format, (int)new_format);
  return DEV_ERR_NOT_SUPPORTED;
}

// control/magma_zmlumerge.cpp:43-109: strict-L (diagonal skipped) then U,
// row by row.
int data_zmlumerge(data_d_matrix L, data_d_matrix U, data_d_matrix* A) {
  data_d_matrix LL, UU;
  memset(&LL, 0, sizeof(LL));
  memset(&UU, 0, sizeof(UU));
  const data_d_matrix* pl = &L;
  const data_d_matrix* pu = &U;
  if (is_csc_like(L.storage_type)) {
    LACE_TRY(transpose_csr(L, &LL));
    pl = &LL;
  }
  if (is_csc_like(U.storage_type)) {
    LACE_TRY(transpose_csr(U, &UU));
    pu = &UU;
  }
  const int n = pl->num_rows;
  free_arrays(A);
  int z = 0;
  for (int i = 0; i < n; i++) {
    for (int k = pl->row[i]; k < pl->row[i + 1]; k++) z += pl->col[k] < i;
    z += pu->row[i + 1] - pu->row[i];
  }
  A->storage_type = Magma_CSR;
  A->major = MagmaRowMajor;
  A->num_rows = n;
  A->num_cols = pl->num_cols;
  A->nnz = z;
  A->true_nnz = z;
  int rc = host_calloc(&A->row, n + 1);
  if (rc == DEV_SUCCESS) rc = host_calloc(&A->col, z);
  if (rc == DEV_SUCCESS) rc = host_calloc(&A->val, z);
  if (rc == DEV_SUCCESS) {
    z = 0;
    for (int i = 0; i < n; i++) {
      A->row[i] = z;
      for (int k = pl->row[i]; k < pl->row[i + 1]; k++)
        if (pl->col[k] < i) {
          A->col[z] = pl->col[k];
          A->val[z++] = pl->val[k];
        }
      for (int k = pu->row[i]; k < pu->row[i + 1]; k++) {
        A->col[z] = pu->col[k];
        A->val[z++] = pu->val[k];
      }
    }
    A->row[n] = z;
  }
  data_zmfree(&LL);
  data_zmfree(&UU);
  return rc;
}

// control/magma_zmio.cpp:92-350 -- MatrixMarket coordinate reader:
// symmetric/hermitian entries are mirrored (:221-273), COO is counting-
// sorted into CSR and every row is sorted by column (:323-336).
int data_z_csr_mtx(data_d_matrix* A, const char* filename) {
  FILE* f = fopen(filename, "r");
  LACE_REQUIRE(f != nullptr, DEV_ERR_NOT_FOUND, "cannot open %s", filename);
  char line[1024];
  if (!fgets(line, sizeof(line), f)) {
    fclose(f);
    set_error("%s: empty file", filename);
    return DEV_ERR_NOT_SUPPORTED;
  }
  char b0[64] = {0}, b1[64] = {0}, b2[64] = {0}, b3[64] = {0}, b4[64] = {0};
  sscanf(line, "%63s %63s %63s %63s %63s", b0, b1, b2, b3, b4);
  for (char* p = b2; *p; ++p) *p = (char)tolower(*p);
  for (char* p = b3; *p; ++p) *p = (char)tolower(*p);
  for (char* p = b4; *p; ++p) *p = (char)tolower(*p);
  if (strcmp(b2, "coordinate") != 0) {
    fclose(f);
    set_error("%s: only coordinate MatrixMarket files are supported", filename);
    return DEV_ERR_NOT_SUPPORTED;
  }
  const bool pattern = strcmp(b3, "pattern") == 0;
  const bool complex_ = strcmp(b3, "complex") == 0;
  const bool symm = strcmp(b4, "symmetric") == 0 || strcmp(b4, "hermitian") == 0;
  int m = 0, n = 0, nz = 0;
  while (fgets(line, sizeof(line), f)) {
    if (line[0] == '%' || line[0] == '\n') continue;
    if (sscanf(line, "%d %d %d", &m, &n, &nz) == 3) break;
  }
  std::vector<int> r, c;
  std::vector<double> v;
  r.reserve(symm ? 2 * nz : nz);
  c.reserve(symm ? 2 * nz : nz);
  v.reserve(symm ? 2 * nz : nz);
  for (int e = 0; e < nz; e++) {
    int ii, jj;
    double x = 1.0, xi = 0.0;
    int got = pattern ? fscanf(f, " %d %d", &ii, &jj)
                      : (complex_ ? fscanf(f, " %d %d %lf %lf", &ii, &jj, &x, &xi)
                                  : fscanf(f, " %d %d %lf", &ii, &jj, &x));
    if (got < 2) {
      fclose(f);
      set_error("%s: truncated entry list", filename);
      return DEV_ERR_NOT_SUPPORTED;
    }
    r.push_back(ii - 1);
    c.push_back(jj - 1);
    v.push_back(x);
    if (symm && ii != jj) {
      r.push_back(jj - 1);
      c.push_back(ii - 1);
      v.push_back(x);
    }
  }
  fclose(f);
  const int nnz = (int)r.size();
  data_zmfree(A);
  A->storage_type = Magma_CSR;
  A->major = MagmaRowMajor;
  A->fill_mode = MagmaFull;
  A->sym = Magma_GENERAL;
  A->num_rows = m;
  A->num_cols = n;
  A->nnz = nnz;
  A->true_nnz = nnz;
  LACE_TRY(host_calloc(&A->row, m + 1));
  LACE_TRY(host_calloc(&A->col, nnz));
  LACE_TRY(host_calloc(&A->val, nnz));
  for (int e = 0; e < nnz; e++) A->row[r[e] + 1]++;
  for (int i = 0; i < m; i++) A->row[i + 1] += A->row[i];
  std::vector<int> fill(A->row, A->row + m);
  for (int e = 0; e < nnz; e++) {
    int p = fill[r[e]]++;
    A->col[p] = c[e];
    A->val[p] = v[e];
  }
  std::vector<std::pair<int, double>> rv;
  int maxrow = 0;
  for (int i = 0; i < m; i++) {
    int k0 = A->row[i], len = A->row[i + 1] - k0;
    maxrow = std::max(maxrow, len);
    rv.resize(len);
    for (int t = 0; t < len; t++) rv[t] = {A->col[k0 + t], A->val[k0 + t]};
    std::stable_sort(rv.begin(), rv.end(),
                     [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
                       return a.first < b.first;
                     });
    for (int t = 0; t < len; t++) {
      A->col[k0 + t] = rv[t].first;
      A->val[k0 + t] = rv[t].second;
    }
  }
  A->max_nnz_row = maxrow;
  return DEV_SUCCESS;
}

// control/magma_zmio.cpp:587 -- "%d %d %.16g" triplets, 1-based.
int data_zwrite_csr_mtx(data_d_matrix A, data_order_t MajorType, const char* filename) {
  FILE* f = fopen(filename, "w");
  LACE_REQUIRE(f != nullptr, DEV_ERR_NOT_FOUND, "cannot write %s", filename);
  fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
  const bool csc = is_csc_like(A.storage_type) || A.major == MagmaColMajor;
  if (MajorType == MagmaColMajor && !csc) {
    data_d_matrix B;
    memset(&B, 0, sizeof(B));
    LACE_TRY(transpose_csr(A, &B));
    fprintf(f, "%d %d %d\n", B.num_cols, B.num_rows, B.row[B.num_rows]);
    for (int i = 0; i < B.num_rows; i++)
      for (int k = B.row[i]; k < B.row[i + 1]; k++)
        fprintf(f, "%d %d %.16g\n", B.col[k] + 1, i + 1, B.val[k]);
    data_zmfree(&B);
  } else {
    fprintf(f, "%d %d %d\n", A.num_rows, A.num_cols, A.row[A.num_rows]);
    for (int i = 0; i < A.num_rows; i++)
      for (int k = A.row[i]; k < A.row[i + 1]; k++)
        fprintf(f, "%d %d %.16g\n", i + 1, A.col[k] + 1, A.val[k]);
  }
  fclose(f);
  return DEV_SUCCESS;
}

// control/magma_zmio.cpp:1056-1190 (read_z_dense_from_mtx) and :1265-1282
// (data_z_dense_mtx): a MatrixMarket "array" file read in file order (real,
// integer, pattern = 1, complex = real part); ld follows `major`.  The
// reference prints its progress as MATLAB comment lines; so does this.
int data_z_dense_mtx(data_d_matrix* A, data_order_t major, const char* filename) {
  FILE* f = fopen(filename, "r");
  if (!f) {
    printf("%% Unable to open file %s\n", filename);
    set_error("cannot open %s", filename);
    return DEV_ERR_NOT_FOUND;
  }
  printf("%% Reading dense matrix from file (%s):", filename);
  fflush(stdout);
  char line[1024];
  if (!fgets(line, sizeof(line), f)) {
    fclose(f);
    set_error("%s: empty file", filename);
    return DEV_ERR_NOT_SUPPORTED;
  }
  char b0[64] = {0}, b1[64] = {0}, b2[64] = {0}, b3[64] = {0};
  sscanf(line, "%63s %63s %63s %63s", b0, b1, b2, b3);
  for (char* p = b2; *p; ++p) *p = (char)tolower(*p);
  for (char* p = b3; *p; ++p) *p = (char)tolower(*p);
  if (strcmp(b2, "array") != 0) {
    fclose(f);
    printf("\n%% Sorry, only dense (array) MatrixMarket files are read here.\n");
    set_error("%s: not a MatrixMarket array file", filename);
    return DEV_ERR_NOT_SUPPORTED;
  }
  const bool pattern = strcmp(b3, "pattern") == 0;
  const bool complex_ = strcmp(b3, "complex") == 0;
  int m = 0, n = 0;
  while (fgets(line, sizeof(line), f)) {
    if (line[0] == '%' || line[0] == '\n') continue;
    if (sscanf(line, "%d %d", &m, &n) == 2) break;
  }
  const int nnz = m * n;
  double* val = nullptr;
  LACE_TRY(host_calloc(&val, std::max(nnz, 1)));
  for (int e = 0; e < nnz; e++) {
    double x = 1.0, xi = 0.0;
    int got = 1;
    if (pattern) {
      int dummy;
      got = fscanf(f, " %d", &dummy);
    } else if (complex_) {
      got = fscanf(f, " %lf %lf", &x, &xi);
    } else {
      got = fscanf(f, " %lf", &x);
    }
    if (got < 1) {
      fclose(f);
      free(val);
      set_error("%s: truncated value list", filename);
      return DEV_ERR_NOT_SUPPORTED;
    }
    val[e] = x;
  }
  fclose(f);
  printf(" done.  done.\n");
  free(A->val);
  A->val = val;
  A->num_rows = m;
  A->num_cols = n;
  A->nnz = nnz;
  A->true_nnz = nnz;
  A->ld = major == MagmaRowMajor ? n : m;
  A->major = major;
  return DEV_SUCCESS;
}

// control/magma_zmio.cpp:1221-1261: MatrixMarket array, one "%.16e" value
// per line in storage order.
int data_zwrite_dense_mtx(int n_row, int n_col, int nnz, data_order_t major, dataType** val,
                          const char* filename) {
  (void)nnz;
  (void)major;
  printf("%% Writing dense matrix to file (%s):", filename);
  fflush(stdout);
  FILE* f = fopen(filename, "w");
  if (!f) {
    printf("\n%% error writing matrix: file exists or missing write permission\n");
    set_error("cannot write %s", filename);
    return -1;
  }
  fprintf(f, "%%%%MatrixMarket matrix array real general\n");
  fprintf(f, "%d %d\n", n_row, n_col);
  for (long i = 0; i < (long)n_row * n_col; i++) fprintf(f, "%.16e\n", (*val)[i]);
  if (fclose(f) != 0) printf("\n%% error: writing matrix failed\n");
  else printf(" done\n");
  fflush(stdout);
  return DEV_SUCCESS;
}

// control/magma_zmio.cpp:1304-1318
int data_zwrite_dense(data_d_matrix A, const char* filename) {
  if (A.storage_type == Magma_DENSED) return data_zwrite_dense_mtx(A.nnz, 1, A.nnz, A.major, &A.val, filename);
  if (A.pad_rows > 0 && A.pad_cols > 0)
    return data_zwrite_dense_mtx(A.pad_rows, A.pad_cols, A.nnz, A.major, &A.val, filename);
  return data_zwrite_dense_mtx(A.num_rows, A.num_cols, A.nnz, A.major, &A.val, filename);
}

// blas/zaxpy.cpp:52-59: dy <- alpha dx + dy (host, strided).
void data_zaxpy(int n, dataType alpha, const dataType* dx, int incx, dataType* dy, int incy) {
  for (int i = 0; i < n; i++) dy[(size_t)i * incy] += alpha * dx[(size_t)i * incx];
}

// blas/zdot.cpp:50-65 (OpenMP reduction; one thread sums in index order)
// and :67-78 (cblas_ddot).  Both are the plain index-order sum here, so the
// two entry points return the same value.
dataType data_zdot(int n, dataType* dx, int incx, dataType* dy, int incy) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += dx[(size_t)i * incx] * dy[(size_t)i * incy];
  return s;
}

dataType data_zdot_mkl(int n, dataType* dx, int incx, dataType* dy, int incy) {
  return data_zdot(n, dx, incx, dy, incy);
}

// control/norms.cpp:7-13
dataType data_dnrm2(int n, dataType* x, int incx) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += x[(size_t)i * incx] * x[(size_t)i * incx];
  return std::sqrt(s);
}

// control/norms.cpp:69-92
int data_zfrobenius(data_d_matrix A, dataType* res) {
  double s = 0.0;
  if (is_csr_like(A.storage_type) || is_csc_like(A.storage_type)) {
    const int n = A.num_rows;
    for (int k = 0; k < A.row[n]; k++) s += A.val[k] * A.val[k];
  } else if (is_dense(A.storage_type)) {
    for (int k = 0; k < A.num_rows * A.num_cols; k++) s += A.val[k] * A.val[k];
  } else {
    set_error("data_zfrobenius: storage %d not supported", (int)A.storage_type);
    return DEV_ERR_NOT_SUPPORTED;
  }
  *res = std::sqrt(s);
  return DEV_SUCCESS;
}

// Sparse x sparse product C = alpha*A*B with sorted rows (the reference
// wraps mkl_sparse_spmm, blas/zspmm.cpp:47-167).  Dense-marker row merge.
int data_z_spmm(dataType alpha, data_d_matrix A, data_d_matrix B, data_d_matrix* C) {
  const int m = A.num_rows, n = B.num_cols;
  std::vector<int> mark(n, -1), list(n);
  std::vector<double> acc(n, 0.0);
  std::vector<int> crow(m + 1, 0), ccol;
  std::vector<double> cval;
  for (int i = 0; i < m; i++) {
    int cnt = 0;
    for (int p = A.row[i]; p < A.row[i + 1]; p++) {
      int k = A.col[p];
      for (int q = B.row[k]; q < B.row[k + 1]; q++) {
        int j = B.col[q];
        if (mark[j] != i) {
          mark[j] = i;
          list[cnt++] = j;
          acc[j] = 0.0;
        }
        acc[j] += A.val[p] * B.val[q];
      }
    }
    std::sort(list.begin(), list.begin() + cnt);
    for (int a = 0; a < cnt; a++) {
      ccol.push_back(list[a]);
      cval.push_back(alpha * acc[list[a]]);
    }
    crow[i + 1] = (int)ccol.size();
  }
  free_arrays(C);
  C->storage_type = Magma_CSR;
  C->major = MagmaRowMajor;
  C->num_rows = m;
  C->num_cols = n;
  C->nnz = crow[m];
  C->true_nnz = crow[m];
  LACE_TRY(host_calloc(&C->row, m + 1));
  LACE_TRY(host_calloc(&C->col, crow[m]));
  LACE_TRY(host_calloc(&C->val, crow[m]));
  memcpy(C->row, crow.data(), sizeof(int) * (m + 1));
  if (crow[m]) {
    memcpy(C->col, ccol.data(), sizeof(int) * crow[m]);
    memcpy(C->val, cval.data(), sizeof(double) * crow[m]);
  }
  return DEV_SUCCESS;
}

// Exact ILU(0), IKJ form, in place on A's CSR (sorted rows).  The
// reference calls MKL dcsrilu0 (blas/zcsrilu0.cpp:15-75) with ipar[30]=0,
// which stops on a pivot below 1e-16 (:41).
int data_dcsrilu0_mkl(data_d_matrix* A) {
  const int n = A->num_rows;
  std::vector<int> diag(n, -1), pos(n, -1);
  for (int i = 0; i < n; i++)
    for (int k = A->row[i]; k < A->row[i + 1]; k++)
      if (A->col[k] == i) diag[i] = k;
  for (int i = 0; i < n; i++) {
    LACE_REQUIRE(diag[i] >= 0, DEV_ERR_NAN, "data_dcsrilu0_mkl: row %d has no diagonal", i);
    for (int k = A->row[i]; k < A->row[i + 1]; k++) pos[A->col[k]] = k;
    for (int p = A->row[i]; p < A->row[i + 1] && A->col[p] < i; p++) {
      int k = A->col[p];
      A->val[p] = A->val[p] / A->val[diag[k]];
      for (int q = diag[k] + 1; q < A->row[k + 1]; q++) {
        int j = A->col[q];
        if (pos[j] >= 0) A->val[pos[j]] = A->val[pos[j]] - A->val[p] * A->val[q];
      }
    }
    for (int k = A->row[i]; k < A->row[i + 1]; k++) pos[A->col[k]] = -1;
    LACE_REQUIRE(std::fabs(A->val[diag[i]]) >= 1e-16, DEV_ERR_NAN,
                 "data_dcsrilu0_mkl: pivot %d below 1e-16", i);
  }
  return DEV_SUCCESS;
}

// src/givens.cpp:46-67
void givens_rotation(dataType a, dataType b, dataType tol, dataType* c, dataType* s) {
  if (std::fabs(b) <= tol) {
    *c = 1.0;
    *s = 0.0;
  } else if (std::fabs(b) > std::fabs(a)) {
    double t = a / b;
    *s = 1.0 / std::sqrt(1.0 + t * t);
    *c = t * (*s);
  } else {
    double t = b / a;
    *c = 1.0 / std::sqrt(1.0 + t * t);
    *s = t * (*c);
  }
}

// ---------------------------------------------------------------- levels
int lace_level_sets(int n, const int* row, const int* col, int upper, int* perm,
                    int* level_ptr) {
  std::vector<int> lev(n, 0);
  int nl = 0;
  for (int t = 0; t < n; t++) {
    int i = upper ? n - 1 - t : t;
    int l = 0;
    for (int k = row[i]; k < row[i + 1]; k++) {
      int j = col[k];
      if ((upper && j > i) || (!upper && j < i)) l = std::max(l, lev[j] + 1);
    }
    lev[i] = l;
    nl = std::max(nl, l + 1);
  }
  if (!perm || !level_ptr) return nl;
  std::vector<int> cnt(nl + 1, 0);
  for (int i = 0; i < n; i++) cnt[lev[i] + 1]++;
  for (int l = 0; l < nl; l++) cnt[l + 1] += cnt[l];
  for (int l = 0; l <= nl; l++) level_ptr[l] = cnt[l];
  for (int i = 0; i < n; i++) perm[cnt[lev[i]]++] = i;
  return nl;
}

// ---------------------------------------------------------------- RNG
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
uint32_t* lace_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
  return out;
}

// N(0,1) samples: counter i/2 of key (seed lo, seed hi) gives four 32-bit
// words; two 53-bit uniforms in (0,1] feed one Box-Muller pair, sample 2t
// takes the cosine branch and 2t+1 the sine branch.
int lace_gen_normal(uint64_t seed, int n, double* out) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  const double two_pi = 6.283185307179586476925286766559;
  for (int t = 0; 2 * t < n; t++) {
    uint32_t ctr[4] = {(uint32_t)t, 0u, 0u, 0u}, w[4];
    lace_philox4x32_10(ctr, key, w);
    uint64_t a = ((uint64_t)w[0] << 21) ^ (uint64_t)(w[1] >> 11);
    uint64_t b = ((uint64_t)w[2] << 21) ^ (uint64_t)(w[3] >> 11);
    double u1 = ((double)(a & ((1ull << 53) - 1)) + 1.0) * (1.0 / 9007199254740992.0);
    double u2 = ((double)(b & ((1ull << 53) - 1)) + 1.0) * (1.0 / 9007199254740992.0);
    double rad = std::sqrt(-2.0 * std::log(u1));
    out[2 * t] = rad * std::cos(two_pi * u2);
    if (2 * t + 1 < n) out[2 * t + 1] = rad * std::sin(two_pi * u2);
  }
  return DEV_SUCCESS;
}

// ---------------------------------------------------------------- stencils
static int stencil_csr(int n, const std::vector<int>& cnt_hint, data_d_matrix* A) {
  (void)cnt_hint;
  data_zmfree(A);
  A->storage_type = Magma_CSR;
  A->major = MagmaRowMajor;
  A->fill_mode = MagmaFull;
  A->sym = Magma_GENERAL;
  A->diagorder_type = Magma_VALUE;
  A->num_rows = n;
  A->num_cols = n;
  return DEV_SUCCESS;
}

// 5-point [4,-1,-1,-1,-1], Dirichlet boundaries, x fastest.
int lace_gen_laplace2d(int nx, int ny, data_d_matrix* A) {
  const int n = nx * ny;
  stencil_csr(n, {}, A);
  int nnz = 0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) nnz += 1 + (i > 0) + (i < nx - 1) + (j > 0) + (j < ny - 1);
  A->nnz = A->true_nnz = nnz;
  A->max_nnz_row = 5;
  LACE_TRY(host_calloc(&A->row, n + 1));
  LACE_TRY(host_calloc(&A->col, nnz));
  LACE_TRY(host_calloc(&A->val, nnz));
  int z = 0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      int p = i + nx * j;
      A->row[p] = z;
      if (j > 0) { A->col[z] = p - nx; A->val[z++] = -1.0; }
      if (i > 0) { A->col[z] = p - 1; A->val[z++] = -1.0; }
      A->col[z] = p; A->val[z++] = 4.0;
      if (i < nx - 1) { A->col[z] = p + 1; A->val[z++] = -1.0; }
      if (j < ny - 1) { A->col[z] = p + nx; A->val[z++] = -1.0; }
    }
  A->row[n] = z;
  return DEV_SUCCESS;
}

// 7-point first-order upwind convection-diffusion, h-scaled, velocity
// along +x,+y,+z with cell Peclet numbers pe_d: a_ii = 6 + sum pe_d,
// a_{i,i-e_d} = -1 - pe_d (upwind neighbour), a_{i,i+e_d} = -1.
// Dirichlet: neighbours outside the box are dropped.  x fastest.
int lace_gen_convdiff3d(int nx, int ny, int nz, double pe_x, double pe_y, double pe_z,
                        data_d_matrix* A) {
  const long nl = (long)nx * ny * nz;
  LACE_REQUIRE(nl < (1L << 31), DEV_ERR_ILLEGAL_VALUE, "grid too large for int32 indices");
  const int n = (int)nl;
  stencil_csr(n, {}, A);
  long nnz = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        nnz += 1 + (i > 0) + (i < nx - 1) + (j > 0) + (j < ny - 1) + (k > 0) + (k < nz - 1);
  LACE_REQUIRE(nnz < (1L << 31), DEV_ERR_ILLEGAL_VALUE, "nnz too large for int32 indices");
  A->nnz = A->true_nnz = (int)nnz;
  A->max_nnz_row = 7;
  LACE_TRY(host_calloc(&A->row, n + 1));
  LACE_TRY(host_calloc(&A->col, nnz));
  LACE_TRY(host_calloc(&A->val, nnz));
  const double d = 6.0 + pe_x + pe_y + pe_z;
  const int sy = nx, sz = nx * ny;
  int z = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        int p = i + sy * j + sz * k;
        A->row[p] = z;
        if (k > 0) { A->col[z] = p - sz; A->val[z++] = -1.0 - pe_z; }
        if (j > 0) { A->col[z] = p - sy; A->val[z++] = -1.0 - pe_y; }
        if (i > 0) { A->col[z] = p - 1; A->val[z++] = -1.0 - pe_x; }
        A->col[z] = p; A->val[z++] = d;
        if (i < nx - 1) { A->col[z] = p + 1; A->val[z++] = -1.0; }
        if (j < ny - 1) { A->col[z] = p + sy; A->val[z++] = -1.0; }
        if (k < nz - 1) { A->col[z] = p + sz; A->val[z++] = -1.0; }
      }
  A->row[n] = z;
  return DEV_SUCCESS;
}

// src/householder.cpp:49-84 (Stewart's housegen): u = X/||X||, u_0 += 1
// with nu = -||X|| when X_0 >= 0 (else u_0 -= 1), u scaled by
// 1/sqrt(|u_0|); ||X|| below the smallest subnormal gives u_0 = sqrt(2).
void data_housegen(int n, dataType* X, dataType* u, dataType* nu) {
  const double eps = 4.9406564584124654e-324;  // nextafter(0, 1)
  for (int i = 0; i < n; i++) u[i] = X[i];
  *nu = data_dnrm2(n, X, 1);
  if (*nu < eps) {
    u[0] = std::sqrt(2.0);
    return;
  }
  for (int i = 0; i < n; i++) u[i] = X[i] / (*nu);
  if (u[0] >= 0) {
    u[0] = u[0] + 1.0;
    *nu = -(*nu);
  } else {
    u[0] = u[0] - 1.0;
  }
  const double t = 1.0 / std::sqrt(std::fabs(u[0]));
  for (int i = 0; i < n; i++) u[i] = u[i] * t;
}

// src/householder.cpp:120-160: Householder triangularization [U, R] =
// hqrd(X) on column-major dense containers (X is overwritten); reflector k
// from X(k:n, k), applied to the columns to its right.  R's off-diagonal
// entries are addressed with X's leading dimension, as the reference does
// (:157), so U, R and X are expected to share it.
void data_hqrd(data_d_matrix* X, data_d_matrix* U, data_d_matrix* R) {
  const int n = X->num_rows, p = X->num_cols;
  const int lx = X->ld, lu = U->ld, lr = R->ld;
  std::vector<double> v(std::max(p, 1), 0.0);
  for (int k = 0; k < std::min(n, p); k++) {
    data_housegen(n - k, &X->val[k + (size_t)k * lx], &U->val[k + (size_t)k * lu],
                  &R->val[k + (size_t)k * lr]);
    for (int j = k + 1; j < p; j++) {
      double colsum = 0.0;
      for (int i = k; i < n; i++) colsum = colsum + U->val[i + (size_t)k * lu] * X->val[i + (size_t)j * lx];
      v[j] = colsum;
    }
    for (int i = k; i < n; i++)
      for (int j = k + 1; j < p; j++)
        X->val[i + (size_t)j * lx] = X->val[i + (size_t)j * lx] - U->val[i + (size_t)k * lu] * v[j];
    for (int j = k + 1; j < p; j++) R->val[k + (size_t)j * lx] = X->val[k + (size_t)j * lx];
  }
}

int lace_gen_bcsr7(int nx, int ny, int nz, int bs, uint64_t seed, double sigma, double shift,
                   data_d_matrix* A) {
  const long nl = (long)nx * ny * nz;
  LACE_REQUIRE(nx > 0 && ny > 0 && nz > 0 && bs > 0, DEV_ERR_ILLEGAL_VALUE, "lace_gen_bcsr7: sizes");
  LACE_REQUIRE(nl * bs < (1L << 31), DEV_ERR_ILLEGAL_VALUE, "grid too large for int32 indices");
  const int nbr = (int)nl, lb = bs * bs;
  long nb = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        nb += 1 + (i > 0) + (i < nx - 1) + (j > 0) + (j < ny - 1) + (k > 0) + (k < nz - 1);
  LACE_REQUIRE(nb * lb < (1L << 31), DEV_ERR_ILLEGAL_VALUE, "too many entries for int32 counts");
  data_zmfree(A);
  A->storage_type = Magma_BCSR;
  A->major = MagmaRowMajor;
  A->fill_mode = MagmaFull;
  A->sym = Magma_GENERAL;
  A->diagorder_type = Magma_VALUE;
  A->num_rows = A->num_cols = nbr;
  A->blocksize = bs;
  A->ldblock = lb;
  A->numblocks = (int)nb;
  A->nnz = A->true_nnz = (int)(nb * lb);
  A->max_nnz_row = 7;
  LACE_TRY(host_calloc(&A->row, nbr + 1));
  LACE_TRY(host_calloc(&A->col, nb));
  LACE_TRY(host_calloc(&A->val, (size_t)nb * lb));
  LACE_TRY(lace_gen_normal(seed, (int)(nb * lb), A->val));
  for (size_t e = 0; e < (size_t)nb * lb; e++) A->val[e] *= sigma;
  const int sy = nx, sz = nx * ny;
  int z = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const int p = i + sy * j + sz * k;
        A->row[p] = z;
        if (k > 0) A->col[z++] = p - sz;
        if (j > 0) A->col[z++] = p - sy;
        if (i > 0) A->col[z++] = p - 1;
        for (int a = 0; a < bs; a++) A->val[(size_t)z * lb + a * bs + a] += shift;
        A->col[z++] = p;
        if (i < nx - 1) A->col[z++] = p + 1;
        if (j < ny - 1) A->col[z++] = p + sy;
        if (k < nz - 1) A->col[z++] = p + sz;
      }
  A->row[nbr] = z;
  return DEV_SUCCESS;
}

}  // extern "C"

// The drop-in boundary depends on these layouts matching the reference's
// include/sparse_types.h; the Python binding mirrors the same offsets.
static_assert(offsetof(data_d_matrix, nnz) == 24, "layout");
static_assert(offsetof(data_d_matrix, val) == 64, "layout");
static_assert(offsetof(data_d_matrix, diag) == 128, "layout");
static_assert(offsetof(data_d_matrix, row) == 192, "layout");
static_assert(offsetof(data_d_matrix, rowidx) == 256, "layout");
static_assert(offsetof(data_d_matrix, col) == 320, "layout");
static_assert(offsetof(data_d_matrix, list) == 328, "layout");
static_assert(offsetof(data_d_matrix, blocksize) == 344, "layout");
static_assert(offsetof(data_d_matrix, major) == 360, "layout");
static_assert(offsetof(data_d_matrix, pad_cols) == 372, "layout");
static_assert(sizeof(data_d_matrix) == 384, "layout");
static_assert(offsetof(data_d_preconditioner_log, tol) == 8, "layout");
static_assert(offsetof(data_d_preconditioner_log, omp_num_threads) == 72, "layout");
static_assert(sizeof(data_d_preconditioner_log) == 128, "layout");
static_assert(offsetof(data_d_gmres_log, solve_time) == 8, "layout");
static_assert(sizeof(data_d_gmres_log) == 64, "layout");
static_assert(offsetof(data_d_gmres_param, parilu_reduction) == 40, "layout");
static_assert(offsetof(data_d_gmres_param, reorder) == 48, "layout");
static_assert(sizeof(data_d_gmres_param) == 64, "layout");
static_assert(sizeof(lace_solver_opts) == 56, "layout");
static_assert(sizeof(lace_solver_stats) == 72, "layout");
