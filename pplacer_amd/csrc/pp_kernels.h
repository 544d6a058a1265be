// Kernel argument blocks and kernel declarations (pp_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pp_internal.h"

namespace pp {

struct Res {  // one optimised (query, edge): Three_tax.get_results + location
  double ll, pend, dist;
  int loc;
  int flags;
  double marg;  // -p: the marginal likelihood (Three_tax.calc_marg_prob)
};

struct CodeLogArgs {
  const double *sn;     // [N][Epad][R*4] FP64 supernodes
  int N, Epad, n_inf;
  const double *ttab;   // [n_tsl][n_codes][R*S] character LVs evolved by start_pend
  int n_codes;
  const int *site_cat;  // Price-CAT: table of each site (nullptr: table 0)
  double *lt64;         // [N][n_inf+1][Epad], or
  float *lt32;          // the same in FP32 (when non-null)
};

struct Stage1Args {
  const void *lt;         // [N][C1][Epad] Real: log util per (site, character, edge)
  int N, E, Epad, C1;     // C1 = n_inf + 1
  const int *sn_e;        // [N][Epad] supernode exponents
  const uint8_t *tcodes;  // [n_tiles][N][TQ], gap = n_inf
  const int *tile_lo, *tile_hi, *tile_q;
  const int *qM;          // masked length per (batch-local) query
  double log_R;
  double *scores;         // [Q][E]
  int n_tiles;            // query tiles; the grid covers n_tiles x Epad/256 workgroups
  int per_xcd;            // workgroups per XCD: ceil(n_tiles * Epad/256 / 8)
};

struct PitchArgs {
  const DevModel *model;
  const uint8_t *codes;  // [Q][N]
  const double *scores;  // [Q][E]
  const int *qlist;
  int nq;
  int *counter;
  const void *darr, *parr;
  const int *darr_e, *parr_e;
  const double *bl;
  int N, E;
  double start_pend, max_pend, initial_tol, strike_box;
  long long t_max_pitches, t_max_strikes;
  int cap;
  Res *res;  // [Q][cap]
  int *n_res, *status;
  // n_calls[0]: likelihood calls (Three_tax.n_like_calls); [2], [3]: sum of M
  // over the pendant-only / distal probes actually run
  // (all stage-2 kernels); [4], [5]: the same for this kernel alone
  unsigned long long *n_calls, *n_pitched;
  // 20-state engine: per-workgroup pair scratch (ProtLayout::scratch_reals(cap_pairs) each);
  // 4-state streaming form: per-workgroup products [R*4][cap_sites]
  void *scratch;
  int cap_pairs, cap_sites;
  const int *site_cat;  // Price-CAT: category of each pre-masked site (else nullptr)
};

struct KeepArgs {
  int Q, cap, K;
  int keep;  // keep_at_most (<= K, the stride of kept)
  const Res *res;
  const int *n_res;
  const int *status;
  double log_keep_factor;
  Res *kept;  // [Q][K]
  int *n_keep;
};

struct OptArgs {
  const DevModel *model;
  const uint8_t *codes;
  int N, mode, n_jobs, K;
  const int *qlist;  // mode 1: bucket queries; mode 0: optional job indirection
  int *counter;
  const void *darr, *parr;
  const int *darr_e, *parr_e;
  const double *bl;
  double tol, max_pend, start_pend;
  // mode 1
  Res *kept;
  const int *n_keep;
  int *status;
  // mode 0
  const int *job_q, *job_loc;
  const double *job_pend, *job_dist;
  double *out_ll, *out_pend, *out_dist;
  int *out_calls, *out_status;
  unsigned long long *n_calls;  // [0]: likelihood calls; [2], [3]: as in PitchArgs
  void *scratch;                // as in PitchArgs
  int cap_pairs, cap_sites;
  const int *site_cat;          // as in PitchArgs
  // marginal likelihoods (-p): mode 1 writes kept[].marg, mode 0 out_ll
  double pp_rel_err;
  int prior_kind;               // PRIOR_EXP or PRIOR_UNIFORM
  const double *prior_mean;     // [G] exponential prior mean of each edge
};

struct FinalArgs {
  int Q, K;
  int calc_pp;  // post_prob / marginal_prob from kept[].marg
  const Res *kept;
  const int *n_keep;
  const int *status;
  pp_placement *out;
  int *n_out;
};

}  // namespace pp
