// Device-side building blocks shared by the placement kernels: wavefront and
// block reductions, the stage-1 ranking key, and the three-taxon engine that
// runs Three_tax.optimize (three_tax.ml:70-83) with the GSL Brent minimiser
// restated on the GPU.
//
// Execution model of the engine: one workgroup owns one (query, edge) at a
// time; lane j holds masked site j of the query (pend GLV code, distal and
// proximal GLV rows, the cached (pi*prox_ev)*dist_ev products).  Every lane
// runs the same scalar Brent / baseball logic on block-uniform values, so the
// control flow is uniform and each likelihood probe is one collective call.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pp_internal.h"

namespace pp {

#define PP_LOG2 0.69314718055994530942  // OCaml `log 2.` (gstar_support.ml:7)

enum {
  ST_OK = 0,
  ST_GSL = 1,          // Gsl_exn: EINVAL from the minimiser's set, or EBADTOL
  ST_GSL_BADFUNC = 2,  // Gsl_exn: EBADFUNC, a non-finite objective value
  ST_MAXITER = 3,      // Minimization.ExceededMaxIter
  ST_FIND_START = 4,   // Minimization.FindStartFailure
  ST_INVALID_START = 5,// Minimization.InvalidStartValues
  ST_ASSERT = 6        // assert(prox_bl >= 0.)
};

__device__ __forceinline__ bool is_gsl(int st) { return st == ST_GSL || st == ST_GSL_BADFUNC; }

// Stable descending order of stage-1 scores as a 64-bit ascending key
// (fig.ml:143-147 with OCaml float compare: nan below everything, -0 == 0).
__device__ __forceinline__ unsigned long long desc_key(double s) {
  if (s != s) return ~0ull;
  if (s == 0.0) s = 0.0;
  unsigned long long u = (unsigned long long)__double_as_longlong(s);
  u = (u >> 63) ? ~u : (u | 0x8000000000000000ull);
  return ~u;
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// v from the lane the DPP control CTRL selects (every lane of the wave active)
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over a full wave, the same value in every lane: pairwise steps within
// rows of 16 lanes through DPP (lane ^ 1, lane ^ 2, then the other quad and
// the other half-row, each already uniform), then the four row sums.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f64<0x141>(v);  // row_half_mirror
  v += dpp_f64<0x140>(v);  // row_mirror
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}
__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Sum over the block; identical bit pattern in every lane (butterfly adds are
// commutative pairwise, then every lane adds the wave partials in order).
// Needs sm[] of >= 16 doubles; contains one barrier.
__device__ __forceinline__ double block_sum(double v, double *sm) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) sm[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < nw; i++) t += sm[i];
  return t;
}
__device__ __forceinline__ long long block_sum_ll(long long v, long long *sm) {
  v = wave_sum_ll(v);
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) sm[w] = v;
  __syncthreads();
  long long t = 0;
  for (int i = 0; i < nw; i++) t += sm[i];
  return t;
}

// (key, edge) minimum across the block; ties resolved by smaller edge id.
__device__ __forceinline__ void block_argmin(unsigned long long &key, int &idx,
                                             unsigned long long *skey, int *sidx) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    unsigned long long k2 = __shfl_xor(key, m, 64);
    int i2 = __shfl_xor(idx, m, 64);
    if (k2 < key || (k2 == key && i2 < idx)) { key = k2; idx = i2; }
  }
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) { skey[w] = key; sidx[w] = idx; }
  __syncthreads();
  key = skey[0];
  idx = sidx[0];
  for (int i = 1; i < nw; i++)
    if (skey[i] < key || (skey[i] == key && sidx[i] < idx)) { key = skey[i]; idx = sidx[i]; }
}

}  // namespace pp
