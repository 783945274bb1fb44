// reorder.cpp -- Cuthill-McKee orderings and the symmetric permutation of a
// CSR matrix (host set-up routines of the drop-in ABI).
//
// Reference: src/cuthill_mckee.cpp:53-265 (data_reorder_cuthill_mckee) and
// src/reorder.cpp:86-253 (data_sparse_reorder).  The ordering keeps every
// behaviour of the reference, which fixes the permutation bit for bit:
//   - neighbour lists are stored 1-offset (:100-115) but read with the
//     window [nnsn[seed-1], nnsn[seed]) (:155-157): the window holds the
//     previous row's last neighbour and all but the seed's last neighbour;
//     for seed 0 the reference reads nnsn[-1], which is 0 here;
//   - the seed is the first row whose degree exceeds row 1's degree and
//     every earlier maximum (:130-142), row 0 when none does;
//   - the queue R is sorted by descending degree of node R[i]+1 (:183-189;
//     nnsn[n] is 0), stably (glibc's qsort is a merge sort);
//   - reverse swaps nn_map[i] and nn_map[n-1-i] for i = 0..n/2, so for even
//     n the middle pair is swapped twice and stays (:241-252).
// Where the reference reads uninitialised memory (the S queue runs out
// before every row is numbered, :217) this returns DEV_ERR.
// data_sparse_reorder uses the ordering as the reference does: row i of
// the result is row P[i] of A, column c becomes Pinv[c], rows re-sorted by
// column; AMD (reorder = 3) needs SuiteSparse and is not provided.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"

using namespace lace;

extern "C" {

int data_reorder_cuthill_mckee(data_d_matrix* A, int* nn_map, lace_bool reverse) {
  LACE_REQUIRE(A && A->row && A->col && nn_map, DEV_ERR_INVALID_PTR,
               "data_reorder_cuthill_mckee: null argument");
  const int nn = A->num_rows;
  if (nn <= 0) return DEV_SUCCESS;
  const int nnz = A->row[nn];
  std::vector<int> nnsn(nn + 1, 0), nsn((size_t)std::max(nnz, 1) + 1, 0);
  for (int i = 0; i < nn; i++) {
    const int offset = i > 0 ? nnsn[i - 1] : 0;
    for (int j = A->row[i]; j < A->row[i + 1]; j++)
      if (A->col[j] != i) {
        nnsn[i]++;
        nsn[offset + nnsn[i]] = A->col[j];
      }
    if (i > 0) nnsn[i] += nnsn[i - 1];
  }
  int seed = 0;
  int maxdegree = nnsn[1] - nnsn[0];
  int degree = nnsn[0];
  for (int i = 0; i < nn; i++) {
    if (i > 0) degree = nnsn[i] - nnsn[i - 1];
    if (degree > maxdegree) {
      maxdegree = degree;
      seed = i;
    }
  }
  std::vector<char> done(nn, 0);
  std::vector<int> R, S;
  R.reserve(nn);
  S.reserve(nn);
  size_t startS = 0;
  int inode = 0;
  nn_map[seed] = inode++;
  done[seed] = 1;
  auto deg_next = [&](int v) { return nnsn[v + 1] - nnsn[v]; };
  while (inode < nn) {
    const int start_idx = seed > 0 ? nnsn[seed - 1] : 0;
    const int end_idx = nnsn[seed];
    R.clear();
    for (int i = start_idx; i < end_idx; i++) {
      const int v = nsn[i];
      if (!done[v]) {
        R.push_back(v);
        done[v] = 1;
      }
    }
    std::stable_sort(R.begin(), R.end(), [&](int a, int b) { return deg_next(a) > deg_next(b); });
    for (int v : R) {
      nn_map[v] = inode++;
      S.push_back(v);
    }
    if (inode >= nn) break;
    LACE_REQUIRE(startS < S.size(), DEV_ERR,
                 "data_reorder_cuthill_mckee: the queue ran out with %d of %d rows numbered "
                 "(the reference reads uninitialised memory here)",
                 inode, nn);
    seed = S[startS++];
  }
  if (reverse) {
    int counter = nn - 1;
    for (int i = 0; i <= nn / 2; i++) {
      std::swap(nn_map[counter], nn_map[i]);
      counter--;
    }
  }
  return DEV_SUCCESS;
}

int data_sparse_reorder(data_d_matrix* A, int* P, int* Pinv, int reorder) {
  LACE_REQUIRE(A && A->row && A->col && A->val && P && Pinv, DEV_ERR_INVALID_PTR,
               "data_sparse_reorder: null argument");
  const int n = A->num_rows;
  switch (reorder) {
    case 1: LACE_TRY(data_reorder_cuthill_mckee(A, P, 0)); break;
    case 2: LACE_TRY(data_reorder_cuthill_mckee(A, P, 1)); break;
    case 3:
      set_error("data_sparse_reorder: AMD ordering (reorder = 3) needs SuiteSparse, not provided");
      return DEV_ERR_NOT_SUPPORTED;
    default:
      for (int k = 0; k < n; k++) P[k] = Pinv[k] = k;
  }
  for (int k = 0; k < n; k++) {
    LACE_REQUIRE(P[k] >= 0 && P[k] < n, DEV_ERR, "data_sparse_reorder: ordering is not a permutation");
    Pinv[P[k]] = k;
  }
  const int nnz = A->row[n];
  data_int_t *row2 = nullptr, *col2 = nullptr;
  dataType* val2 = nullptr;
  LACE_TRY(host_calloc(&row2, n + 1));
  LACE_TRY(host_calloc(&col2, std::max(nnz, 1)));
  LACE_TRY(host_calloc(&val2, std::max(nnz, 1)));
  std::vector<std::pair<int, double>> ent;
  for (int i = 0; i < n; i++) {
    const int o = P[i];
    ent.clear();
    for (int j = A->row[o]; j < A->row[o + 1]; j++) ent.emplace_back(Pinv[A->col[j]], A->val[j]);
    std::sort(ent.begin(), ent.end(),
              [](const std::pair<int, double>& a, const std::pair<int, double>& b) { return a.first < b.first; });
    row2[i + 1] = row2[i] + (int)ent.size();
    for (size_t t = 0; t < ent.size(); t++) {
      col2[row2[i] + t] = ent[t].first;
      val2[row2[i] + t] = ent[t].second;
    }
  }
  free(A->row);
  free(A->col);
  free(A->val);
  A->row = row2;
  A->col = col2;
  A->val = val2;
  return DEV_SUCCESS;
}

}  // extern "C"
