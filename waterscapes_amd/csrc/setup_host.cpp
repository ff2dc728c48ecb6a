// Host-side setup helpers (C++/OpenMP): node-pair sparsity patterns of the
// assembled operators from cell -> node maps.
#include <stdint.h>
#include <string.h>
#include <vector>
#include <algorithm>
#include <omp.h>
#include "../../include/mpet_hip.h"

extern "C" void mpet_set_error(const char* msg);

// CSR pattern of all (r, c) pairs with r in rmap[e, :] and c in cmap[e, :]
// for some cell e.  Two passes (count, fill); rowptr has nrow + 1 entries,
// cols must hold rowptr[nrow] entries (call with cols == NULL to size it).
extern "C" int mpet_build_pattern(int ncell, int nr, const int32_t* rmap, int nc,
                                  const int32_t* cmap, int nrow, int64_t* rowptr,
                                  int32_t* cols) {
  if (ncell < 0 || nr <= 0 || nc <= 0 || nrow < 0 || !rmap || !cmap || !rowptr) {
    mpet_set_error("mpet_build_pattern: invalid argument");
    return MPET_ERR_INVALID;
  }
  // node -> cells (CSR) for the row map
  std::vector<int64_t> nptr(nrow + 1, 0);
  for (int64_t k = 0; k < (int64_t)ncell * nr; ++k) nptr[rmap[k] + 1]++;
  for (int i = 0; i < nrow; ++i) nptr[i + 1] += nptr[i];
  std::vector<int32_t> ncell_of(nptr[nrow]);
  {
    std::vector<int64_t> pos(nptr.begin(), nptr.end() - 1);
    for (int e = 0; e < ncell; ++e)
      for (int a = 0; a < nr; ++a) ncell_of[pos[rmap[(int64_t)e * nr + a]]++] = e;
  }
  const bool fill = cols != nullptr;
  if (!fill) rowptr[0] = 0;
#pragma omp parallel
  {
    std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 512)
    for (int r = 0; r < nrow; ++r) {
      buf.clear();
      for (int64_t k = nptr[r]; k < nptr[r + 1]; ++k) {
        const int32_t* cm = cmap + (int64_t)ncell_of[k] * nc;
        buf.insert(buf.end(), cm, cm + nc);
      }
      std::sort(buf.begin(), buf.end());
      int64_t m = std::unique(buf.begin(), buf.end()) - buf.begin();
      if (fill)
        memcpy(cols + rowptr[r], buf.data(), sizeof(int32_t) * m);
      else
        rowptr[r + 1] = m;
    }
  }
  if (!fill)
    for (int r = 0; r < nrow; ++r) rowptr[r + 1] += rowptr[r];
  return MPET_OK;
}
