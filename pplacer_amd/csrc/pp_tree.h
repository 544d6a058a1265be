#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "pp_internal.h"

namespace pp {

struct TreeDesc {
  int n_nodes, n_sites;
  const int *child_start, *children;
  const double *bl;
};

// pplacer layout: [G][R][N][S] doubles, [G][N] int64 exponents
struct TreeGlvs {
  std::vector<double> darr, parr, snodes;
  std::vector<int64_t> darr_e, parr_e, snodes_e;
};


}  // namespace pp
