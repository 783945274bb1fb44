// dia.hpp -- diagonal (DIA) storage for sparse matrices whose entries lie
// on a few diagonals, as the structured-grid stencils do.  Each diagonal d
// has an offset off[d] = column - row (ascending); the value of row i on it
// sits at val[d * n + i], zero where row i has no entry there, and bit d of
// mask[i] says whether it has one.  Compared with ELLPACK the column index
// of every entry disappears (only the offsets remain), and the gathers of
// the vector operand become contiguous shifted reads.
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace lace {

constexpr int kDiaMax = 8;

struct DiaOffsets {
  int o[kDiaMax];
};

struct DiaForm {
  int n = 0, nd = 0;  // nd = 0: the form does not apply
  DiaOffsets off{};
  DevBuf mask;  // uint8_t per row
  DevBuf map;   // int per (diagonal, row): position in the source value array, or -1
  DevBuf val;   // gathered values, diagonal-major

  // Entries (row i, CSR position k) for which keep(i, k) holds.  The form
  // applies when they lie on at most kDiaMax diagonals, padding stays under
  // 25 %, and each row lists them by strictly ascending column (so the DIA
  // order of a row's entries is its CSR order).
  template <typename Keep>
  int build(int n_, const int* row, const int* col, Keep keep) {
    n = n_;
    nd = 0;
    if (n <= 0) return DEV_SUCCESS;
    std::vector<int> offs;
    long kept = 0;
    for (int i = 0; i < n; i++) {
      int last = -1;
      for (int k = row[i]; k < row[i + 1]; k++) {
        if (!keep(i, k)) continue;
        if (col[k] <= last) return DEV_SUCCESS;  // unsorted or repeated columns
        last = col[k];
        kept++;
        const int o = col[k] - i;
        if (std::find(offs.begin(), offs.end(), o) == offs.end()) {
          if ((int)offs.size() == kDiaMax) return DEV_SUCCESS;
          offs.push_back(o);
        }
      }
    }
    if (offs.empty() || (double)offs.size() * n > 1.25 * (double)kept) return DEV_SUCCESS;
    std::sort(offs.begin(), offs.end());
    const int D = (int)offs.size();
    std::vector<uint8_t> mk(n, 0);
    std::vector<int> mp((size_t)D * n, -1);
    for (int i = 0; i < n; i++)
      for (int k = row[i]; k < row[i + 1]; k++) {
        if (!keep(i, k)) continue;
        const int d = (int)(std::lower_bound(offs.begin(), offs.end(), col[k] - i) - offs.begin());
        mk[i] |= (uint8_t)(1u << d);
        mp[(size_t)d * n + i] = k;
      }
    for (int d = 0; d < kDiaMax; d++) off.o[d] = d < D ? offs[d] : 0;
    LACE_TRY(upload(mask, mk.data(), mk.size()));
    LACE_TRY(upload(map, mp.data(), sizeof(int) * mp.size()));
    nd = D;
    return DEV_SUCCESS;
  }

  // val <- the source values (es = 4 or 8 bytes each) through map.
  int gather(const void* src, size_t es);

  // Bytes one pass over the form reads: values and masks.
  double bytes(size_t es) const { return (double)nd * n * (double)es + (double)n; }
};

}  // namespace lace
