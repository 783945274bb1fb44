// bcsr.hpp -- block-sparse (BCSR: bs x bs row-major blocks) operator and
// block ParILU(0) factors on the device.
#pragma once

#include <vector>

#include "common.hpp"

namespace lace {

constexpr int kMaxBlock = 8;  // largest block size with a kernel instance

// Device BCSR matrix: nbr block rows, block row pointers, block columns,
// values (numblocks * bs * bs, row-major inside each block).
struct DevBCSR {
  int nbr = 0, bs = 0, nblocks = 0;
  lace_prec_t prec = LACE_F64;
  DevBuf row, col, val;
  int create(const data_d_matrix* A, lace_prec_t p);
  int n() const { return nbr * bs; }
  // Algorithmic bytes of one product (values, block columns, row pointers,
  // x, y; b for the residual form).
  double bytes(int sx, int sy, bool residual) const;
};

// y = A x, or r = b - A x (residual form).  Vector precisions follow the
// template arguments; FP64 accumulation when A or y is FP64.
template <typename TX, typename TY>
int bcsr_spmv(const DevBCSR& A, const TX* x, TY* y);
template <typename TX, typename TY>
int bcsr_residual(const DevBCSR& A, const TX* x, const TY* b, TY* r);

// Block ParILU(0).  Factor layouts are the reference's: L in BCSRL (block
// row I: strict-lower blocks by ascending column, then the identity
// diagonal block, which is implicit here and never read), U in BCSCU
// (block column J: blocks of rows K <= J ascending, the diagonal last;
// interiors as they appear in U, not transposed).
//
// When A's block pattern is structurally symmetric (and every diagonal
// block is present, which is always required), U's column J has exactly
// the pattern of L's row J, so a factor "position" q carries both L_IJ and
// U_JI and one lane group updates the pair from the merged rows I and J.
// Otherwise one lane group per block of A merges L row I with U column J.
//
// Storage precision: F64 (FP64 iteration), F32, BF16 or F16 (FP32
// iteration).  16-bit storage keeps a power-of-two scale per scalar row:
// a stored entry is value / 2^e(row) with 2^e(row) <= max |A row| < 2^(e+1),
// so stored magnitudes stay near one whatever the row's size.
struct BILU {
  int nbr = 0, bs = 0, nnzL = 0, nnzU = 0, nstrict = 0, nblocksA = 0;
  bool sym = false;
  int variant = 0;  // 0: symmetric-pattern kernel when possible, 1: general;
                    // 2-4: symmetric kernel forms (merge, tables, tables + LDS)
  lace_prec_t store = LACE_F32;
  double anorm = 0.0;
  // L (BCSRL) and U (BCSCU) patterns; lrow = block row of each L entry
  std::vector<int> h_lptr, h_lcol, h_lrow, h_ucp, h_uri;
  // U by block rows for the solves: entries (U index, block column K > I)
  std::vector<int> h_urp, h_uidx, h_ucolk;
  DevBuf lptr, lcol, lrow, ucp, uri, urp, uidx, ucolk;
  std::vector<int> h_arow, h_acol;  // A (host copy, for the general kernel)
  std::vector<double> h_aval;
  bool have_general = false;
  DevBuf arowidx, acol;  // A's block coordinates (general kernel)
  DevBuf AL, AU;         // A_IJ and A_JI at each L position (symmetric kernel)
  DevBuf smeta, spend, spairs;  // symmetric kernel's per-position tables
  bool wave_local = false;      // every product operand within a wave run or its halo
  DevBuf swhalo;                // per wave run: 1 if a product reaches before the run
  DevBuf AV;             // A's blocks in its own order (general kernel)
  DevBuf lfrom, ufrom;   // A block feeding each L / U entry (-1: none)
  DevBuf Lv[2], Uv[2];
  DevBuf rs;             // row scales (16-bit storage)
  int cur = 0;
  DevBuf partial, scalar;
  int last_nb = 0;
  // triangular solves
  std::vector<int> h_lperm, h_llev, h_uperm, h_ulev;  // level schedules
  DevBuf lperm, uperm;
  DevBuf Dinv;  // inverses of U's diagonal blocks (iteration precision)
  bool dirty = true;
  DevBuf work;
  bool c64() const { return store == LACE_F64; }
  bool is16s() const { return store == LACE_BF16 || store == LACE_F16; }
  int selem() const { return store == LACE_F64 ? 8 : (store == LACE_F32 ? 4 : 2); }
  int celem() const { return c64() ? 8 : 4; }
  int n() const { return nbr * bs; }
  bool use_sym() const { return sym && variant != 1; }
  int create(const data_d_matrix* A, lace_prec_t storage);
  int upload_general();
  int set_variant(int v);
  int reset();
  int sweep(bool with_step);
  int read_step(double* step);  // sum of squared changes of the last sweep
  int refresh();                // diagonal-block inverses after sweeps
  double sweep_bytes() const;
  int get(data_d_matrix* L, data_d_matrix* U);
  // z = U^{-1} L^{-1} r; r and z in the iteration precision (double for F64
  // storage, float otherwise).
  int apply(lace_trsv_kind_t kind, int sweeps, const void* r, void* z);
};

}  // namespace lace
