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

// Gtree.tree_length (gtree.ml:102-114): the non-root branch lengths, folded
// subtree by subtree in the reference's order.
double tree_length(const TreeDesc &t);
// Core.prior_mean_map with top_segment_f = (+ lower) (core.ml:34-52):
// means[i] = lower + bl[i] + the average distance from node i to the leaves
// below it, summed in leaf order; the root entry is left at 0.
void prior_mean_map(const TreeDesc &t, double lower, double *means);

}  // namespace pp
