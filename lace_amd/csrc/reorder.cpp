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

namespace {

// Breadth-first search from r over the nodes not yet numbered (mark -1 is
// numbered; other marks are search stamps): the number of levels, and in
// far the last level's node of least (degree, number).
int bfs_depth(int r, const std::vector<int>& ptr, const std::vector<int>& adj, const std::vector<int>& deg,
              std::vector<int>& mark, int& stamp, std::vector<int>& q, int& far) {
  ++stamp;
  q.clear();
  q.push_back(r);
  mark[r] = stamp;
  size_t b = 0, last_b = 0, last_e = 1;
  int levels = 0;
  while (b < q.size()) {
    const size_t e = q.size();
    last_b = b;
    last_e = e;
    for (size_t t = b; t < e; t++)
      for (int k = ptr[q[t]]; k < ptr[q[t] + 1]; k++) {
        const int u = adj[k];
        if (mark[u] != stamp && mark[u] != -1) {
          mark[u] = stamp;
          q.push_back(u);
        }
      }
    b = e;
    levels++;
  }
  far = q[last_b];
  for (size_t t = last_b + 1; t < last_e; t++) {
    const int u = q[t];
    if (deg[u] < deg[far] || (deg[u] == deg[far] && u < far)) far = u;
  }
  return levels;
}

}  // namespace

extern "C" {

// Reverse Cuthill-McKee in its textbook form (George & Liu, Computer
// Solution of Large Sparse Positive Definite Systems, 1981, ch. 4), on the
// symmetrised graph of A without self loops:
//   - each component starts from a pseudo-peripheral node: from its node of
//     least (degree, number), move to the last BFS level's node of least
//     (degree, number) while the number of levels grows;
//   - its nodes are numbered breadth-first, each node's unnumbered
//     neighbours in increasing (degree, number);
//   - the whole order is reversed.
// P[k] is the old number of the row placed at k.  This is not the
// reference's ordering: the reference's quirks above give the 7-point grid
// a bandwidth several times its natural one, on which synchronous ParILU
// sweeps overflow.
int lace_reorder_rcm(const data_d_matrix* A, int* P) {
  LACE_REQUIRE(A && A->row && A->col && P, DEV_ERR_INVALID_PTR, "lace_reorder_rcm: null argument");
  LACE_REQUIRE(A->num_rows == A->num_cols, DEV_ERR_ILLEGAL_VALUE, "lace_reorder_rcm: matrix is not square");
  const int n = A->num_rows;
  std::vector<std::vector<int>> nb(n);
  for (int i = 0; i < n; i++)
    for (int k = A->row[i]; k < A->row[i + 1]; k++) {
      const int j = A->col[k];
      LACE_REQUIRE(j >= 0 && j < n, DEV_ERR_ILLEGAL_VALUE, "lace_reorder_rcm: column %d out of range", j);
      if (j != i) {
        nb[i].push_back(j);
        nb[j].push_back(i);
      }
    }
  std::vector<int> ptr(n + 1, 0), adj, deg(n);
  for (int i = 0; i < n; i++) {
    std::sort(nb[i].begin(), nb[i].end());
    nb[i].erase(std::unique(nb[i].begin(), nb[i].end()), nb[i].end());
    deg[i] = (int)nb[i].size();
    ptr[i + 1] = ptr[i] + deg[i];
  }
  adj.reserve(ptr[n]);
  for (int i = 0; i < n; i++) adj.insert(adj.end(), nb[i].begin(), nb[i].end());
  nb.clear();
  auto least = [&](int a, int b) { return deg[a] < deg[b] || (deg[a] == deg[b] && a < b); };
  std::vector<int> by_deg(n);
  for (int i = 0; i < n; i++) by_deg[i] = i;
  std::sort(by_deg.begin(), by_deg.end(), least);
  std::vector<int> mark(n, 0), order, q, cand;
  order.reserve(n);
  int stamp = 0;
  size_t next = 0;
  while ((int)order.size() < n) {
    while (mark[by_deg[next]] == -1) next++;
    int root = by_deg[next], far = root, far2 = root;
    int depth = bfs_depth(root, ptr, adj, deg, mark, stamp, q, far);
    for (;;) {
      const int d2 = bfs_depth(far, ptr, adj, deg, mark, stamp, q, far2);
      root = far;
      if (d2 <= depth) break;
      depth = d2;
      far = far2;
    }
    const size_t base = order.size();
    order.push_back(root);
    mark[root] = -1;
    for (size_t t = base; t < order.size(); t++) {
      const int v = order[t];
      cand.clear();
      for (int k = ptr[v]; k < ptr[v + 1]; k++)
        if (mark[adj[k]] != -1) cand.push_back(adj[k]);
      std::sort(cand.begin(), cand.end(), least);
      for (int u : cand) {
        mark[u] = -1;
        order.push_back(u);
      }
    }
  }
  for (int k = 0; k < n; k++) P[k] = order[n - 1 - k];
  return DEV_SUCCESS;
}

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
    case 4: LACE_TRY(lace_reorder_rcm(A, P)); break;  // addition: textbook RCM
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
