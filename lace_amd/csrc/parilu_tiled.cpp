// parilu_tiled.cpp -- the tiled ParILU family on the host, for CPU
// comparison with the device sweeps.
//
// Reference: src/parilu_v3_0.cpp:15-229 (and v3.1, v3.2, v3.5-v3.9, which
// form the same tile products with other machinery: dense GEMM on extracted
// blocks, sparse sub-matrices, sparse products, batches or tasks).  Every
// variant pads A to whole tiles (control/magma_zmio.cpp:973-1040), splits it
// into L (CSRL, unit diagonal), U (CSRU) and the reciprocal diagonal D, and
// sweeps the tile grid in place until the scaled step is at most 1e-15
// (1000 sweeps at most in v3.0-v3.2, 10000 from v3.5 on).  For tile (ti, tj)
// the product C = L(ti rows, < span) U(< span, tj cols), span = min(ti,
// tj) + tile, is formed from the values at the start of the tile, and then
// (ti <= tj) the tile's strict U entries become Uc - C, on the diagonal tile
// also D_i = 1 / (Dc_i - C_ii) and then the tile's L entries (Lc - C) D_j;
// (ti > tj) the tile's L entries (Lc - C) D_j.  Finally U's diagonal is
// 1 / D.
//
// The reference runs the tile rows with OpenMP and races; with one thread it
// is the fixed order above, which is what this restatement computes.  C's
// entries are summed over the inner index in ascending order from zero (the
// reference leaves that order to its GEMM), only for the entries a tile
// updates, and only tiles that hold entries of L or U are visited: the other
// tiles of the reference's loop compute products nobody reads.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.hpp"

#pragma clang fp contract(off)

namespace lace {
namespace {

// A padded to whole tiles as control/magma_zmio.cpp:973-1040 does: pad_rows
// and pad_cols rounded up, one unit entry per padding row appended (at
// column num_cols + k), num_rows kept; the arrays are replaced.
int pad_csr(data_d_matrix* A, int tile) {
  const int n = A->num_rows;
  const int pr = (n + tile - 1) / tile * tile;
  const int pc = (A->num_cols + tile - 1) / tile * tile;
  printf("tile_size = %d num_rows = %d pad_rows = %d \n", tile, n, pr);
  printf(" %d additional unit diagonal terms added to pad sparse matrix in csr format\n", pr - n);
  const int nnz = A->nnz, extra = pr - n;
  double* val = nullptr;
  int *col = nullptr, *row = nullptr;
  LACE_TRY(host_calloc(&val, (size_t)nnz + extra));
  LACE_TRY(host_calloc(&col, (size_t)nnz + extra));
  LACE_TRY(host_calloc(&row, (size_t)pr + 1));
  memcpy(val, A->val, sizeof(double) * nnz);
  memcpy(col, A->col, sizeof(int) * nnz);
  memcpy(row, A->row, sizeof(int) * (n + 1));
  for (int k = 0; k < extra; k++) {
    val[nnz + k] = 1.0;
    col[nnz + k] = A->num_cols + k;
    row[n + 1 + k] = nnz + k + 1;
  }
  free(A->val);
  free(A->col);
  free(A->row);
  A->val = val;
  A->col = col;
  A->row = row;
  A->pad_rows = pr;
  A->pad_cols = pc;
  A->true_nnz = nnz;
  A->nnz = nnz + extra;
  return DEV_SUCCESS;
}

int tiled_parilu(data_d_matrix* A, data_d_matrix* L, data_d_matrix* U, int tile, int max_iter,
                 int version) {
  LACE_REQUIRE(A && L && U, DEV_ERR_INVALID_PTR, "PariLU v3.%d: null argument", version);
  LACE_REQUIRE(A->storage_type == Magma_CSR && A->num_rows == A->num_cols && A->num_rows > 0,
               DEV_ERR_ILLEGAL_VALUE, "PariLU v3.%d: A must be square CSR", version);
  LACE_REQUIRE(tile >= 1, DEV_ERR_ILLEGAL_VALUE, "PariLU v3.%d: tile %d", version, tile);
  const int n = A->num_rows;
  // control/magma_zmatrixchar.cpp:99-124
  int diam = 0;
  for (int i = 0; i < n; i++)
    for (int k = A->row[i]; k < A->row[i + 1]; k++) diam = std::max(diam, std::abs(i - A->col[k]));
  A->diameter = diam;
  LACE_TRY(pad_csr(A, tile));
  printf("A padded by %d tile:\n", tile);
  fflush(stdout);
  L->diagorder_type = Magma_UNITY;
  LACE_TRY(data_zmconvert(*A, L, Magma_CSR, Magma_CSRL));
  U->diagorder_type = Magma_VALUE;
  U->major = MagmaRowMajor;
  LACE_TRY(data_zmconvert(*A, U, Magma_CSR, Magma_CSRU));
  // A's values at the factors' positions, and the diagonal
  const std::vector<double> Lc(L->val, L->val + L->nnz), Uc(U->val, U->val + U->nnz);
  std::vector<double> Dc(n, 0.0), D(n, 0.0);
  for (int i = 0; i < n; i++)
    for (int k = A->row[i]; k < A->row[i + 1]; k++)
      if (A->col[k] == i) Dc[i] = A->val[k];
  for (int i = 0; i < n; i++) {
    LACE_REQUIRE(Dc[i] != 0.0, DEV_ERR_ILLEGAL_VALUE, "PariLU v3.%d: zero or missing diagonal in row %d",
                 version, i);
    D[i] = 1.0 / Dc[i];
  }
  // U's strict entries by column (row order), for the tile products
  std::vector<int> ucp(n + 1, 0), uck, ucq;
  for (int i = 0; i < n; i++)
    for (int q = U->row[i]; q < U->row[i + 1]; q++)
      if (U->col[q] != i) ucp[U->col[q] + 1]++;
  for (int j = 0; j < n; j++) ucp[j + 1] += ucp[j];
  uck.resize(ucp[n]);
  ucq.resize(ucp[n]);
  {
    std::vector<int> fill(ucp.begin(), ucp.end() - 1);
    for (int i = 0; i < n; i++)
      for (int q = U->row[i]; q < U->row[i + 1]; q++)
        if (U->col[q] != i) {
          const int j = U->col[q];
          uck[fill[j]] = i;
          ucq[fill[j]++] = q;
        }
  }
  // the column tiles each row tile touches (L and U entries, diagonal)
  const int row_limit = A->pad_rows > 0 ? A->pad_rows : n;
  const int col_limit = A->pad_cols > 0 ? A->pad_cols : n;
  const int ntr = row_limit / tile, ntc = col_limit / tile;
  std::vector<std::vector<int>> ctiles(ntr);
  for (int t = 0; t < ntr; t++) {
    std::vector<char> seen(ntc, 0);
    for (int i = t * tile; i < std::min(n, (t + 1) * tile); i++) {
      for (int q = L->row[i]; q < L->row[i + 1]; q++) seen[L->col[q] / tile] = 1;
      for (int q = U->row[i]; q < U->row[i + 1]; q++) seen[U->col[q] / tile] = 1;
    }
    for (int c = 0; c < ntc; c++)
      if (seen[c]) ctiles[t].push_back(c);
  }
  // sum_k L_ik U_kj over the strict entries, k ascending, from the current
  // values (the caller snapshots a tile before updating it)
  auto prod = [&](int i, int j) {
    double acc = 0.0;
    int a = L->row[i], b = ucp[j];
    const int ae = L->row[i + 1], be = ucp[j + 1];
    while (a < ae && b < be) {
      const int ka = L->col[a], kb = uck[b];
      if (ka == i) {  // L's unit diagonal is not part of the product
        a++;
        continue;
      }
      if (ka == kb) acc = acc + L->val[a] * U->val[ucq[b]];
      a += ka <= kb;
      b += ka >= kb;
    }
    return acc;
  };
  double anorm = 0.0;
  LACE_TRY(data_zfrobenius(*A, &anorm));
  printf("%% Anorm = %e\n", anorm);
  printf("%% A->diameter=%d\n", A->diameter);
  struct Upd {
    int kind, row, pos;  // kind 0: U entry, 1: diagonal, 2: L entry
    double c;
  };
  std::vector<Upd> upd;
  const auto t0 = std::chrono::steady_clock::now();
  int iter = 0;
  double step = 1.0;
  while (step > 1.0e-15 && iter < max_iter) {
    step = 0.0;
    for (int tr = 0; tr < ntr; tr++) {
      const int ti = tr * tile;
      for (const int tc : ctiles[tr]) {
        const int tj = tc * tile, maxcol = tj + tile;
        // the tile's products, all from the values before its updates
        upd.clear();
        for (int i = ti; i < std::min(n, ti + tile); i++) {
          if (ti <= tj) {
            for (int q = U->row[i]; q < U->row[i + 1]; q++) {
              const int j = U->col[q];
              if (j >= tj && j < maxcol && j != i) upd.push_back({0, i, q, prod(i, j)});
            }
          }
          if (ti == tj) upd.push_back({1, i, i, prod(i, i)});
          if (ti >= tj) {
            for (int q = L->row[i]; q < L->row[i + 1]; q++) {
              const int j = L->col[q];
              if (j >= tj && j < maxcol && j != i) upd.push_back({2, i, q, prod(i, j)});
            }
          }
        }
        // U entries, then the diagonal, then L entries (which read the new
        // D), each in row order
        for (int kind = 0; kind < 3; kind++)
          for (const Upd& u : upd) {
            if (u.kind != kind) continue;
            double tmp;
            if (kind == 0) {
              tmp = Uc[u.pos] - u.c;
              step += std::pow(U->val[u.pos] - tmp, 2);
              U->val[u.pos] = tmp;
            } else if (kind == 1) {
              tmp = 1.0 / (Dc[u.row] - u.c);
              step += std::pow(D[u.row] - tmp, 2);
              D[u.row] = tmp;
            } else {
              tmp = (Lc[u.pos] - u.c) * D[L->col[u.pos]];
              step += std::pow(L->val[u.pos] - tmp, 2);
              L->val[u.pos] = tmp;
            }
          }
      }
    }
    step /= anorm;
    iter++;
    printf("%% iteration = %d step = %e\n", iter, step);
  }
  const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int i = 0; i < n; i++)
    for (int q = U->row[i]; q < U->row[i + 1]; q++)
      if (U->col[q] == i) U->val[q] = 1.0 / D[i];
  printf("%% PariLU v3.%d used %d OpenMP threads and required %d iterations, %f wall clock seconds, and an "
         "average of %f wall clock seconds per iteration as measured by omp_get_wtime()\n",
         version, 1, iter, wall, iter > 0 ? wall / iter : 0.0);
  fflush(stdout);
  return DEV_SUCCESS;
}

}  // namespace
}  // namespace lace

using lace::tiled_parilu;

extern "C" {

// src/parilu_v3_0.cpp:15-229 (v3.1 src/parilu_v3_1.cpp, v3.2
// src/parilu_v3_2.cpp: the same with other tile products, 1000 sweeps at
// most; v3.5-v3.9 src/parilu_v3_5.cpp ... parilu_v3_9.cpp: 10000).  void in
// the reference; on failure the reason is in lace_last_error() and L, U are
// left as far as they got.
void data_PariLU_v3_0(data_d_matrix* A, data_d_matrix* L, data_d_matrix* U, int tile) {
  (void)tiled_parilu(A, L, U, tile, 1000, 0);
}
void data_PariLU_v3_1(data_d_matrix* A, data_d_matrix* L, data_d_matrix* U, int tile) {
  (void)tiled_parilu(A, L, U, tile, 1000, 1);
}
void data_PariLU_v3_2(data_d_matrix* A, data_d_matrix* L, data_d_matrix* U, int tile) {
  (void)tiled_parilu(A, L, U, tile, 1000, 2);
}
void data_PariLU_v3_5(data_d_matrix* A, data_d_matrix* L, data_d_matrix* U, int tile) {
  (void)tiled_parilu(A, L, U, tile, 10000, 5);
}
void data_PariLU_v3_6(data_d_matrix* A, data_d_matrix* L, data_d_matrix* U, int tile) {
  (void)tiled_parilu(A, L, U, tile, 10000, 6);
}
void data_PariLU_v3_7(data_d_matrix* A, data_d_matrix* L, data_d_matrix* U, int tile) {
  (void)tiled_parilu(A, L, U, tile, 10000, 7);
}
void data_PariLU_v3_8(data_d_matrix* A, data_d_matrix* L, data_d_matrix* U, int tile) {
  (void)tiled_parilu(A, L, U, tile, 10000, 8);
}
void data_PariLU_v3_9(data_d_matrix* A, data_d_matrix* L, data_d_matrix* U, int tile) {
  (void)tiled_parilu(A, L, U, tile, 10000, 9);
}

}  // extern "C"
