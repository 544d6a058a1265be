// Kernels of the per-site log sums shared by the linear_c.c twins: the
// util vector of ten_log_like3_c / ten_*_logdot_c / mat_* and its
// sequential log sum.  Internal linkage, so that every translation unit that
// launches them holds its own copy of the same code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pp_crmath.h"

namespace pp {

// per-site util of ten_log_like3_c (statd != null) or ten_*_logdot_c;
// sites [first, first+n_used), optional mask
static __global__ void util_kernel(double *util, const double *statd, const double *x, const double *y,
                                   const double *z, const uint16_t *mask, int first, int n_used, int R,
                                   int N, int S) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_used) return;
  const int site = first + t;
  double u = 0.0;
  if (!mask || mask[site]) {
    for (int r = 0; r < R; r++) {
      const long o = ((long)r * N + site) * S;
      for (int s = 0; s < S; s++) {
        if (statd) u = u + ((statd[s] * x[o + s]) * y[o + s]) * z[o + s];
        else u = u + x[o + s] * y[o + s];
      }
    }
  }
  util[t] = u;
}

// sequential site-order log sum (linear_c.c:416-421 / 527-536) of correctly
// rounded logs (pp_crmath.h), as the stubs' libm returns them; the mat_*
// stubs (linear_c.c:170-340) have no n log R term (rate_term = 0)
static __global__ void logsum_kernel(const double *util, const uint16_t *mask, int first, int n_used,
                                     int R, int rate_term, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double ll = 0;
  int used = 0;
  for (int t = 0; t < n_used; t++)
    if (!mask || mask[first + t]) { ll += cr::log(util[t]); used++; }
  if (rate_term) ll -= ((double)(float)used) * cr::log((double)(float)R);
  *out = ll;
}

}  // namespace pp
