// parilu.hpp -- device handles shared by the ParILU, triangular-solve and
// Krylov translation units.
#pragma once

#include <vector>

#include "common.hpp"
#include "dia.hpp"

namespace lace {

// Marker of an unsolved row in the exact solves' solution vectors: a
// signalling-NaN bit pattern, which arithmetic never produces (it yields
// quiet NaNs).
template <typename T>
struct Sentinel;
template <>
struct Sentinel<double> {
  using U = unsigned long long;
  static constexpr U bits = 0x7FF4DEADBEEF5A5AULL;
};
template <>
struct Sentinel<float> {
  using U = unsigned int;
  static constexpr U bits = 0x7FA5A5A5U;
};

// Pattern of a triangular factor in CSR with its level schedule.
//   upper = false: diagonal last in each row (CSRL layout)
//   upper = true : diagonal first in each row (CSRU layout)
// Level sets: lev(i) = 1 + max lev(k) over the off-diagonal entries of row
// i.  `sched` lists rows level by level; each level starts on a multiple of
// 64 slots (padding slots hold -1) so that one wavefront never holds two
// levels -- the sync-free kernel relies on it.
struct TriPlan {
  int n = 0, nnz = 0;
  bool upper = false;
  int nlev = 0;
  int nslots = 0;
  bool unit_diag = false;  // every diagonal value is exactly 1 (ParILU's L, L^T)
  int maxoff = 0;          // most off-diagonal entries in one row
  DevBuf row, col;
  DevBuf perm, lptr;        // rows by level, level pointers (unpadded)
  DevBuf sched;             // padded schedule for the sync-free kernel
  std::vector<int> h_lptr;  // level pointers on the host
  // Off-diagonal form for the Jacobi sweeps: CSR of the off-diagonal
  // entries in row order (srow, scol), each one's position in the full CSR
  // (smap) and each row's diagonal position (dpos); sval / dval hold the
  // off-diagonal and diagonal values gathered from the value array
  // strict_src (the sweeps use them when called with that array).
  // refresh_strict records strict_src and marks every form below stale;
  // ensure_forms gathers the stale ones a solve is about to read, so the
  // forms no solve reads (the transposed factors' in a forward solve, the
  // ELLPACK and CSR copies where the diagonal form runs) cost nothing.
  int snnz = 0;
  DevBuf srow, scol, smap, dpos, sval, dval;
  const void* strict_src = nullptr;
  // The same off-diagonal entries in ELLPACK form when every row has at most
  // 8 of them and padding stays under 25 %: ew slots per row, slot-major
  // (slot q of row i at q*n + i), column -1 marking padding; emap gives each
  // slot's position in the full CSR, eval its gathered value.
  int ew = 0;
  DevBuf ecol, emap, eval;
  // The same off-diagonal entries in diagonal form when they lie on at most
  // 8 diagonals (stencil factors); gathered with the other forms.
  DiaForm dia;
  // For factors stored in 16 bits (BF16 or F16): the diagonal form's values
  // and the diagonal as 16-bit words relative to each row's power-of-two
  // scale, and that scale's exponent per row; refreshed from the FP32 forms
  // by refresh_dia16 (the Jacobi sweeps use them when called with dia16_src).
  lace_prec_t dia16_fmt = LACE_F32;
  DevBuf dia16, dexp, dval16;
  const void* dia16_src = nullptr;
  // Pipelined Jacobi on the diagonal form (several sweeps per launch): the
  // per-tile progress flags, the tile ticket counter, the intermediate
  // sweep vectors, and the next launch's ticket and flag bases.
  mutable DevBuf pipe_flags, pipe_ctr, pipe_buf;
  mutable unsigned pipe_tickets = 0, pipe_epoch = 0;
  // Exact solves in box order.  When the rows are the points of an
  // nx x ny x nz grid (row = x + nx (y + ny z)) and every off-diagonal entry
  // lies at non-negative coordinate steps from its row (towards lower rows
  // for a lower factor, higher ones for an upper factor: the factors of a
  // 7- or 5-point stencil), the persistent exact solve runs on the rows
  // renumbered box by box -- boxes of tile_dims rows per side taken in
  // order of their x+y+z, rows inside a box in natural order.  That order
  // is a topological one, so the permuted factor stays triangular, and a
  // block of the persistent solve is one box: its in-block dependency
  // chains are as long as the box's diagonal instead of a grid line, and
  // the blocks' wavefront crosses the grid diagonally.  tperm[p] is the row
  // at position p; trow / tcol the CSR in that order (columns renumbered,
  // each row's entries in their original order, so every row sum is the
  // same); tmap[k] is entry k's position in the original CSR.  tval holds
  // the values gathered from the array tval_src (cleared by refresh_strict,
  // which every change of the factor values passes through); tsol is the
  // solution in box order.
  int tile_dims[3] = {0, 0, 0};  // box sides; 0: no box order
  DevBuf tperm, trow, tcol, tmap;
  mutable DevBuf tval, tsol;
  mutable const void* tval_src = nullptr;
  static constexpr unsigned kFormStrict = 1, kFormDiag = 2, kFormEll = 4, kFormDia = 8, kFormAll = 15;
  unsigned stale = 0;  // forms not yet gathered from strict_src
  size_t src_es = 0;   // element size of strict_src
  int build(int n, const int* row, const int* col, bool upper);
  int build_tiles(const int* row, const int* col);
  int refresh_strict(const void* val, size_t elem_size);
  int ensure_forms(unsigned forms);
  // 16-bit diagonal form (fmt BF16 or F16) from the current FP32 forms,
  // whose values are already 16-bit values times the row scales rs.
  int refresh_dia16(lace_prec_t fmt, const float* rs);
};

// y = T^{-1} r, exact.  Level-scheduled: one launch per level.
template <typename T>
int tri_solve_levels(const TriPlan& P, const T* val, const T* r, T* y);
// y = T^{-1} r, exact, one launch: rows run in level order and wait on the
// rows they depend on (value slots are pre-filled with a NaN sentinel).
template <typename T>
int tri_solve_syncfree(const TriPlan& P, const T* val, const T* r, T* y, int* err);
// y = T^{-1} r exact on raw device CSR arrays (upper = false: diagonal last
// in each row, CSRL; true: diagonal first, CSRU), one persistent launch:
// blocks of rows in ticket order, dependencies inside a block through LDS,
// across blocks through y itself (a NaN sentinel marks unsolved rows).
// ticket and err are device words; err is set if a wait gave up.  init:
// fill y with the sentinel and zero the ticket first (false: the caller
// has done so).  r and y must be distinct.  wgs_per_cu: resident
// workgroups per CU (0: as many as fit); poll_sleep: sleeps of 64 cycles
// between two polls of the solution vector; rows: rows per block (512 or
// 256).
template <typename T>
int tri_solve_persist_raw(int n, bool upper, const int* row, const int* col, const T* val, const T* r, T* y,
                          unsigned* ticket, int* err, bool init = true, int wgs_per_cu = 0, int poll_sleep = 0,
                          int rows = 512);
// `sweeps` synchronous Jacobi sweeps from y = 0; tmp is a scratch vector.
// epi_y / epi_d: the last sweep may also write epi_y = y / epi_d (the first
// sweep of a following solve with diagonal epi_d) and then sets *epi_done;
// it does so only on the diagonal form, and epi_y must not be r, y or tmp.
// first_done: the first sweep's result r / diag is already in the buffer
// the first sweep would write (y for odd sweeps, tmp for even).
template <typename T>
int tri_solve_jacobi(const TriPlan& P, const T* val, const T* r, T* y, T* tmp, int sweeps,
                     T* epi_y = nullptr, const T* epi_d = nullptr, bool* epi_done = nullptr,
                     bool first_done = false);

// Device ParILU factors (see parilu.hip for the layout).
struct ParILU {
  int n = 0, nnz = 0, nnzL = 0, nnzU = 0;
  lace_prec_t prec = LACE_F64;
  double anorm = 0.0;
  DevBuf arow, acol, arowidx, aval, dst;
  DevBuf lrow, lcol, lval[2];
  DevBuf ucp, uri, uval[2];
  int cur = 0;
  DevBuf partial, scalar;
  std::vector<int> h_lrow, h_lcol, h_ucp, h_uri;
  // Structurally symmetric patterns: L (CSRL) and U (CSCU) then share one
  // pattern (row i's strict-lower columns, then i), so the sweep runs over
  // L positions and updates l_ij and u_ji from one merge.  prow: row of
  // each position; AL/AU: a_ij and a_ji in position order (iteration
  // precision); maxl: longest strict-lower row.
  bool sym = false;
  // sweep kernel: 0 auto, 1 general CSR-L/CSC-U, 2 symmetric per-position,
  // 3 symmetric per-row, 4 symmetric streaming (precomputed merge data)
  int variant = 0;
  int maxl = 0;
  int last_nb = 0;  // per-block partials written by the last sweep
  DevBuf prow, AL, AU;
  // streaming kernel: pmeta[p] = index of u_jj for a strict position,
  // -(row start + 1) for a diagonal one; strict positions' own matches
  // (k < j) as index pairs, present only when some position has one
  DevBuf pmeta, mptr, mpairs;
  bool strict_matches = false;

  // Triangular solves.  lower: L (CSRL arrays, values lval[cur]).
  // upper: U in CSR (diagonal first), values gathered from the CSC buffer
  // by u_csr_from_csc after every factor update.  Transposed solves:
  // lowerT = U^T, whose CSR is U's CSC (diagonal last, values uval[cur]);
  // upperT = L^T (diagonal first), values gathered through l_t_from_csr.
  TriPlan lower, upper, lowerT, upperT;
  DevBuf u_csr_from_csc, uval_csr;
  DevBuf l_t_from_csr, lval_t;
  bool csr_dirty = true;
  DevBuf work, work2;  // scratch vectors for the solves (n elements each)
  // [0] an exact solve gave up waiting, [1] a pipelined Jacobi launch gave
  // up waiting, [2] the persistent exact solve's block ticket
  DevBuf syncfree_err;

  // 16-bit factor storage (prec BF16 or F16): the sweeps iterate in FP32
  // on the FP32 value arrays, and at reset and after every sweep each
  // factor value is rounded to 16 bits relative to the power-of-two scale
  // of its row (rs, from A's row maxima); L's unit diagonal stays exact.
  // The triangular solves read the rounded values (the Jacobi sweeps of L
  // and U as 16-bit words); lrowof is the row of each L position.
  DevBuf rs, lrowof;
  bool s16() const { return prec == LACE_BF16 || prec == LACE_F16; }
  int round16(bool both_buffers, bool with_step);
  size_t elem_size() const { return prec == LACE_F64 ? 8 : 4; }
  int reset();
  // New values of A on the pattern given at create (host array, nnz
  // entries in A's CSR order): refreshes ||A||_F and every device copy of A
  // the sweeps read, so the next reset() starts from them.
  int set_values(const double* vals);
  int sweep(bool with_step);
  int read_step(double* step);
  int nonlinres(double* out);
  double sweep_bytes() const;
  int get(data_d_matrix* L, data_d_matrix* U);
  int values_f64(double* lv, double* uv);
  int build_trsv();
  int refresh_csr();
  // Everything a solve of this kind and direction reads (block: the
  // 32-column solves), current for the present factor values: L^T's values
  // (lval_t) on a first transposed solve, and the off-diagonal forms the
  // Jacobi sweeps will read.
  bool lval_t_dirty = true;
  int prepare(lace_trsv_kind_t kind, int trans, bool block);
  // z = M^{-1} r (trans = 0) or M^{-T} r (trans = 1); r and z in the
  // factor precision (double for F64, float otherwise).
  int apply(lace_trsv_kind_t kind, int sweeps, int trans, const void* r, void* z);
};

int parilu_create(const data_d_matrix* A, lace_prec_t prec, ParILU** out);

}  // namespace lace
