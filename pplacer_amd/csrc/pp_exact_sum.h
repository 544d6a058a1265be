// The stub's sequential log sum (linear_c.c:416-419), 0 + v[0] + v[1] + ...
// + v[n-1] with one rounding per term in site order, bit for bit, computed by
// a whole wave instead of one lane.
//
// Why a wave can do it: let s_k be the sequential partial sums and u_E =
// 2^(E-53) the spacing of doubles in the binade [2^(E-1), 2^E) (frexp
// exponent E).  If s_{k-1} and s_k lie in the same binade E, s_{k-1} is a
// multiple of u_E, so fl(s_{k-1} + v_k) = s_{k-1} + rint(v_k / u_E) u_E unless
// v_k / u_E is exactly halfway between integers (then ties-to-even depends on
// s_{k-1}).  Such a step is an exact translation by an integer number of u_E
// that does not depend on s_{k-1}.  A run of them adds up exactly, in any
// order, when all its steps have one sign and the run stays in the binade
// (the integers then sum to less than 2^52).
//
// So the wave
//  1. splits v into 64 contiguous blocks, takes approximate prefix sums (a
//     DPP scan of block sums) and Q = sum |v|.  Every approximate partial sum
//     p_k is within delta = (4n + 256) 2^-53 Q of s_k (recursive summation
//     error bounds, doubled);
//  2. walks each block: the binade of s_k is certain when p_k is farther than
//     delta from both binade boundaries (otherwise the wave falls back).  A
//     step is an "event" when it is the first, changes the binade, is a tie,
//     or moves toward zero; every other step is a translation r_k u_E, and
//     the translations between two events are summed per block (exactly);
//  3. numbers the events by a scan of per-lane counts, lists their term
//     indices in event order, and adds the blocks' leading translations into
//     the event before them (LDS atomics; exact integer multiples of one u_E,
//     so order does not matter);
//  4. replays the events in order on one wave-uniform value:
//     s = (s + v_event) + translation.  The first add is the real rounding
//     step of the sequential sum, the second is exact.
// Sums with non-finite terms, more than kSeqSumEvents events, an uncertain
// binade or more than 4096 terms fall back to the one-lane chain.  Each step
// of the wave's path is a dependent FP64 operation of ~12 cycles, so below
// kSeqSumWaveMin terms the one-lane chain is faster and is used directly.
#pragma once

#include <hip/hip_runtime.h>

#include "pp_device.h"

namespace pp {

constexpr int kSeqSumEvents = 64;
// Below this many terms the one-lane chain is the faster of the two (measured
// on MI355X: 2,500 against 6,800 cycles at 100 terms, 8,600 against 7,400 at
// 500, 25,000 against 19,300 at 1,600).
constexpr int kSeqSumWaveMin = 400;

// 0 + v[0] + v[1] + ... + v[n-1], one lane.  A dependent chain, so it is fed
// from LDS two blocks of eight ahead: the loads of one block are in flight
// while the previous block is added.
__device__ __forceinline__ double ordered_sum(const double *v, int n) {
  double s = 0.0;
  double a[8], b[8];
#pragma unroll
  for (int j = 0; j < 8; j++) a[j] = j < n ? v[j] : 0.0;
#pragma unroll
  for (int j = 0; j < 8; j++) b[j] = 8 + j < n ? v[8 + j] : 0.0;
  int i = 0;
  for (; i + 32 <= n; i += 16) {  // a = v[i..i+8), b = v[i+8..i+16), both in range
#pragma unroll
    for (int j = 0; j < 8; j++) s += a[j];
#pragma unroll
    for (int j = 0; j < 8; j++) a[j] = v[i + 16 + j];
#pragma unroll
    for (int j = 0; j < 8; j++) s += b[j];
#pragma unroll
    for (int j = 0; j < 8; j++) b[j] = v[i + 24 + j];
  }
  // fewer than 32 left: a and b hold v[i .. i+16) where in range
#pragma unroll
  for (int j = 0; j < 8; j++)
    if (i + j < n) s += a[j];
#pragma unroll
  for (int j = 0; j < 8; j++)
    if (i + 8 + j < n) s += b[j];
  for (i += 16; i < n; i++) s += v[i];
  return s;
}

// v from the lane CTRL selects; lanes whose row ROWMASK leaves out, or whose
// source lane falls outside the row, get 0.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_i32_or0(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWMASK, 0xF, false);
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64_or0(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)dpp_i32_or0<CTRL, ROWMASK>((int)(unsigned)b);
  const unsigned hi = (unsigned)dpp_i32_or0<CTRL, ROWMASK>((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Inclusive prefix sums over the wave (all 64 lanes active): shifts by 1, 2,
// 4 and 8 within each row of 16, then row 0's total into row 1 and row 2's
// into row 3 (row_bcast:15), then rows 0-1's total into rows 2 and 3
// (row_bcast:31).
__device__ __forceinline__ int wave_scan_i32(int v) {
  v += dpp_i32_or0<0x111, 0xF>(v);
  v += dpp_i32_or0<0x112, 0xF>(v);
  v += dpp_i32_or0<0x114, 0xF>(v);
  v += dpp_i32_or0<0x118, 0xF>(v);
  v += dpp_i32_or0<0x142, 0xA>(v);
  v += dpp_i32_or0<0x143, 0xC>(v);
  return v;
}
__device__ __forceinline__ double wave_scan_f64(double v) {
  v += dpp_f64_or0<0x111, 0xF>(v);
  v += dpp_f64_or0<0x112, 0xF>(v);
  v += dpp_f64_or0<0x114, 0xF>(v);
  v += dpp_f64_or0<0x118, 0xF>(v);
  v += dpp_f64_or0<0x142, 0xA>(v);
  v += dpp_f64_or0<0x143, 0xC>(v);
  return v;
}

// The binade of a partial sum s known only as p with |p - s| <= delta: its
// frexp exponent, if s cannot lie in another binade (and is not 0).
__device__ __forceinline__ bool certain_binade(double p, double delta, int &e) {
  const double m = fabs(__builtin_amdgcn_frexp_mant(p));
  e = __builtin_amdgcn_frexp_exp(p);
  const double sc = __builtin_amdgcn_ldexp(1.0, e);
  return ((m - 0.5) * sc > delta) & ((1.0 - m) * sc > delta) & (e > -960);
}

// Sequential sum of v[0..n) (LDS, n wave-uniform), the same value in every
// lane; call with the whole wave active.  acc: n doubles of LDS the wave may
// overwrite; ev: 2 * kSeqSumEvents doubles of LDS.  v is left unchanged.
// *n_events (when given) receives the number of events, or -1 where the sum
// fell back to the one-lane chain.  Sums shorter than min_wave terms take the
// one-lane chain directly.
__device__ __forceinline__ double wave_sequential_sum(const double *v, int n, double *acc, double *ev,
                                                      int *n_events = nullptr, int min_wave = 0) {
  constexpr int C = 8;  // terms per LDS batch
  const int lane = threadIdx.x & 63;
  if (n <= 0) {
    if (n_events) *n_events = 0;
    return 0.0;
  }
  const int B = (n + 63) >> 6;
  if (B <= 64 && n >= min_wave) {
    const int lo = min(lane * B, n), hi = min(lo + B, n);
    double a = 0.0, q = 0.0;
    for (int c = lo; c < hi; c += C) {
      double xs[C];
#pragma unroll
      for (int j = 0; j < C; j++) xs[j] = c + j < hi ? v[c + j] : 0.0;
#pragma unroll
      for (int j = 0; j < C; j++) {
        a += xs[j];
        q += fabs(xs[j]);
      }
    }
    double p = wave_scan_f64(a) - a;  // approximate partial sum before the block
    const double Q = wave_sum(q);
    if (Q <= 0x1p+996 && Q >= 0x1p-900) {  // every term finite, not all zero
      const double delta = (4.0 * n + 256.0) * 0x1p-53 * Q;
      bool bad = false;
      int prevE = 0;
      bool prevNeg = false;
      if (lane > 0 && lo < hi) {
        bad = !certain_binade(p, delta, prevE);
        prevNeg = p < 0.0;
      }
      unsigned long long mask = 0;  // events of the block, bit k - lo
      double cur = 0.0, head = 0.0;
      int curk = -1;
      for (int c = lo; c < hi; c += C) {
        double xs[C];
#pragma unroll
        for (int j = 0; j < C; j++) xs[j] = c + j < hi ? v[c + j] : 0.0;
#pragma unroll
        for (int j = 0; j < C; j++) {
          const int k = c + j;
          const bool in = k < hi;
          const double x = xs[j];
          p += x;
          int E;
          const bool cert = certain_binade(p, delta, E);
          const double y = __builtin_amdgcn_ldexp(x, 53 - E), r = __builtin_rint(y);
          const bool tie = fabs(y - r) == 0.5;
          const bool toward0 = (x != 0.0) & ((x < 0.0) != prevNeg);
          const bool ev_k = in & ((k == 0) | (E != prevE) | tie | toward0);
          bad |= in & !cert;
          // an event closes the translation run before it
          const double seg = __builtin_amdgcn_ldexp(cur, prevE - 53);
          if (ev_k & (curk >= 0)) acc[curk] = seg;
          head = (ev_k & (curk < 0)) ? seg : head;
          mask |= ev_k ? 1ull << (k - lo) : 0ull;
          cur = ev_k ? 0.0 : (in ? cur + r : cur);
          curk = ev_k ? k : curk;
          prevE = in ? E : prevE;
          prevNeg = in ? p < 0.0 : prevNeg;
        }
      }
      if (lo < hi) {
        const double seg = __builtin_amdgcn_ldexp(cur, prevE - 53);
        if (curk >= 0) acc[curk] = seg;
        else head = seg;
      }
      const int cnt = __popcll(mask);
      const int incl = wave_scan_i32(cnt);
      const int base = incl - cnt;  // this block's first event number
      const int n_ev = __builtin_amdgcn_readlane(incl, 63);
      if (__ballot(bad) == 0 && n_ev <= kSeqSumEvents) {
        // pos[t] = the term index of event t
        int *pos = reinterpret_cast<int *>(ev);
        {
          int j = base;
          for (unsigned long long mm = mask; mm; mm &= mm - 1) pos[j++] = lo + __builtin_ctzll(mm);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // a block's leading translation belongs to the last event before it
        const bool has_head = (lo < hi) & (base > 0) & (head != 0.0);
        const int kprev = has_head ? pos[base - 1] : 0;
        const int kj = lane < n_ev ? pos[lane] : 0;
        if (has_head) __hip_atomic_fetch_add(&acc[kprev], head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const double xj = lane < n_ev ? v[kj] : 0.0;
        const double rj = lane < n_ev ? acc[kj] : 0.0;
        if (n_events) *n_events = n_ev;
        double s = 0.0;
        for (int t = 0; t < n_ev; t++) {
          s = s + readlane_f64(xj, t);
          s = s + readlane_f64(rj, t);
        }
        return s;
      }
    }
  }
  if (n_events) *n_events = -1;
  double s = 0.0;
  if (lane == 0) s = ordered_sum(v, n);
  return readlane_f64(s, 0);
}

}  // namespace pp
