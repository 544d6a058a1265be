// Master/worker three-taxon engine for 20-state (protein) models.
//
// Same command protocol and control as the nucleotide engine
// (pp_engine_mw.h, pp_engine_ctl.h), with the evolves taken in the
// eigenbasis of Q.  P_r(t)[i][j] = sum_k x[i][k] e_r[k] xit[j][k] with
// e_r[k] = exp(rate_r * t * l_k) (Diagd.multi_exp, diagd.ml:79-100), so
//
//   evolved row  y_i = sum_k x[i][k] * e_r[k] * a_k,   a_k = sum_j xit[j][k] * src_j
//   pendant row  z_i = sum_k x[i][k] * e_r[k] * L[c][k], L[c][k] = sum_j xit[j][k] * lv[c][j]
//   util = sum_r sum_i ((pi_i x_i) y_i) z_i = sum_r sum_k e_pend,r[k] * G_r[k],
//          G_r[k] = L[c][k] * sum_i ((pi_i x_i) y_i) x[i][k].
//
// A probe therefore needs only the 20 R exponentials per changed branch
// length, which the master computes; no P matrix is formed.  Work item = one
// (masked site, rate) pair; the R pairs of a site sit in consecutive lanes
// and their utils are summed across those lanes.  Per-pair state -- the
// eigen coordinates a of the distal and proximal rows (set at each gather)
// and G (set at each distal probe) -- lives in a per-workgroup region of a
// global scratch buffer laid out [k][pair], so the worker's registers hold
// only the probe's temporaries.
//
// This is the same likelihood as ten_log_like3_c (linear_c.c:367-424) on
// the same GLVs; the sums are grouped differently from the stubs' explicit
// P(t) * GLV products, so results agree to rounding (~1e-14 relative).
#pragma once

#include <climits>
#include <type_traits>

#include "pp_engine.h"
#include "pp_engine_ctl.h"
#include "pp_engine_mw.h"

namespace pp {

constexpr int kProtS = 20;

struct ProtShared {
  double x[kProtS * kProtS], xit[kProtS * kProtS];  // [i][k], [j][k]
  double lt[kMaxCodes * kProtS];                    // L[c][k]
  double l[kProtS], statd[kProtS], rates[kMaxRates];
  // exp(rate_r * bl * l_k) at [r][k] for the proximal, distal, pendant lengths
  double ep[kMaxRates * kProtS], ed[kMaxRates * kProtS], en[kMaxRates * kProtS];
  int op, M, q, loc;
  // FP32 form: which transition tables a probe rebuilds (bit 0: proximal and
  // distal, bit 1: pendant), and the number of character codes
  int tflags, n_codes;
  int tile_next;  // FP32 form: the next 64-pair tile a worker wave takes
  double part[16];      // per-worker-wave sums of site logs
  long long ipart[16];  // per-worker-wave sums of site exponents
};

struct ProtLayout {
  static constexpr size_t head = (sizeof(ProtShared) + 15) & ~size_t(15);
  // shared header, masked-site list
  static __host__ __device__ size_t bytes(int N) { return head + sizeof(int) * (size_t)N; }
  static __device__ int *sidx(char *s) { return reinterpret_cast<int *>(s + head); }
  static __host__ __device__ size_t align(size_t v) { return (v + 15) & ~size_t(15); }
  // Reals of per-pair scratch per workgroup: a(dist), a(prox), G, each [20][cap]
  static __host__ __device__ size_t scratch_reals(int cap_pairs) { return (size_t)3 * kProtS * cap_pairs; }
};

__device__ __forceinline__ void prot_copy_model(ProtShared *sh, const DevModel *dm) {
  constexpr int S = kProtS;
  for (int t = threadIdx.x; t < S * S; t += blockDim.x) { sh->x[t] = dm->x[t]; sh->xit[t] = dm->xit[t]; }
  for (int t = threadIdx.x; t < S; t += blockDim.x) { sh->l[t] = dm->l[t]; sh->statd[t] = dm->statd[t]; }
  for (int t = threadIdx.x; t < kMaxRates; t += blockDim.x) sh->rates[t] = dm->rates[t];
  for (int t = threadIdx.x; t < dm->n_codes * S; t += blockDim.x) {
    const int c = t / S, k = t % S;
    double acc = dm->xit[k] * dm->lv[c * S];
    for (int j = 1; j < S; j++) acc = acc + dm->xit[j * S + k] * dm->lv[c * S + j];
    sh->lt[t] = acc;
  }
}

// The eigenbasis contractions use fused multiply-adds in both precisions:
// this grouping of the likelihood already differs from the stubs' rounding.
__device__ __forceinline__ double emadd(double acc, double a, double b) { return fma(a, b, acc); }
__device__ __forceinline__ float emadd(float acc, float a, float b) { return fmaf(a, b, acc); }

// ---------------------------------------------------------------------------
// worker waves
// ---------------------------------------------------------------------------
// site_cat: Price-CAT (R == 1), the exponential tables of site s are those of
// its category site_cat[s]; nullptr: the tables of rate r.
template <typename Real, int R>
__device__ __forceinline__ void prot_worker(char *smem, Real *scr, int cap, int N, const uint8_t *codes,
                                            const Real *darr, const Real *parr, const int *darr_e,
                                            const int *parr_e, const int *site_cat) {
  constexpr int S = kProtS, RS = R * S;
  ProtShared *sh = reinterpret_cast<ProtShared *>(smem);
  const int *sidx = ProtLayout::sidx(smem);
  const int lane = threadIdx.x & 63;
  const int wv = wave_id() - 1;
  const int stride = (int)blockDim.x - 64;  // worker lanes, a multiple of 64
  const int my = wv * 64 + lane;
  Real *Ad = scr, *Ap = scr + (size_t)S * cap, *Gs = scr + (size_t)2 * S * cap;
  // The command stream is QUERY (GATHER PROBE*)* ... EXIT (see pp_engine_mw.h).
  __syncthreads();  // A
  int op = sh->op;
  while (op == CMD_QUERY) {
    const int n_pairs = sh->M * R;
    const int npl = (n_pairs + stride - 1) / stride;  // pairs per lane, uniform
    const uint8_t *qc = codes + (size_t)sh->q * N;
    __syncthreads();  // B
    __syncthreads();  // A
    op = sh->op;
    while (op == CMD_GATHER) {
      const int loc = sh->loc;
      long long es = 0;
      for (int k = 0; k < npl; k++) {
        const int p = my + k * stride;
        if (p < n_pairs) {
          const int m = p / R, r = p % R;
          const size_t row = (size_t)loc * N + sidx[m];
          const Real *dp = darr + row * RS + r * S, *pp = parr + row * RS + r * S;
          Real v[S];
#pragma unroll
          for (int j = 0; j < S; j++) v[j] = dp[j];
#pragma unroll 2
          for (int kk = 0; kk < S; kk++) {
            Real acc = (Real)sh->xit[kk] * v[0];
#pragma unroll
            for (int j = 1; j < S; j++) acc = emadd(acc, (Real)sh->xit[j * S + kk], v[j]);
            Ad[(size_t)kk * cap + p] = acc;
          }
#pragma unroll
          for (int j = 0; j < S; j++) v[j] = pp[j];
#pragma unroll 2
          for (int kk = 0; kk < S; kk++) {
            Real acc = (Real)sh->xit[kk] * v[0];
#pragma unroll
            for (int j = 1; j < S; j++) acc = emadd(acc, (Real)sh->xit[j * S + kk], v[j]);
            Ap[(size_t)kk * cap + p] = acc;
          }
          if (r == 0) es += (long long)darr_e[row] + (long long)parr_e[row];
        }
      }
      es = wave_sum_ll(es);
      if (lane == 0) sh->ipart[wv] = es;
      __syncthreads();  // B
      __syncthreads();  // A
      op = sh->op;
      while (op == CMD_PROBE || op == CMD_PROBE_DIST) {
        const bool nd = op == CMD_PROBE_DIST;
        double lsum = 0.0;
        for (int k = 0; k < npl; k++) {
          const int p = my + k * stride;
          // the R pairs of a site are all active or all inactive
          const bool act = p < n_pairs;
          const int pc = act ? p : 0;
          const int m = pc / R, r = pc % R;
          const int er = site_cat ? site_cat[sidx[m]] : r;  // exponential table
          Real util;
          if (nd) {
            Real bd[S], bp[S], h[S];
#pragma unroll
            for (int kk = 0; kk < S; kk++) {
              bd[kk] = (Real)sh->ed[er * S + kk] * Ad[(size_t)kk * cap + pc];
              bp[kk] = (Real)sh->ep[er * S + kk] * Ap[(size_t)kk * cap + pc];
            }
#pragma unroll 1
            for (int i = 0; i < S; i++) {
              const double *xr = sh->x + i * S;
              Real xi = (Real)xr[0] * bp[0], yi = (Real)xr[0] * bd[0];
#pragma unroll
              for (int kk = 1; kk < S; kk++) {
                xi = emadd(xi, (Real)xr[kk], bp[kk]);
                yi = emadd(yi, (Real)xr[kk], bd[kk]);
              }
              const Real wi = ((Real)sh->statd[i] * xi) * yi;
#pragma unroll
              for (int kk = 0; kk < S; kk++) h[kk] = i == 0 ? wi * (Real)xr[kk] : emadd(h[kk], wi, (Real)xr[kk]);
            }
            const int c = qc[sidx[m]];
            util = 0;
#pragma unroll
            for (int kk = 0; kk < S; kk++) {
              const Real g = (Real)sh->lt[c * S + kk] * h[kk];
              if (act) Gs[(size_t)kk * cap + pc] = g;
              util = kk == 0 ? (Real)sh->en[er * S] * g : emadd(util, (Real)sh->en[er * S + kk], g);
            }
          } else {
            util = (Real)sh->en[er * S] * Gs[pc];
#pragma unroll
            for (int kk = 1; kk < S; kk++) util = emadd(util, (Real)sh->en[er * S + kk], Gs[(size_t)kk * cap + pc]);
          }
          double u = act ? (double)util : 0.0;
          if (R == 4) {  // the site's four rates sit in one quad of lanes
            u += dpp_f64<0xB1>(u);  // lane ^ 1
            u += dpp_f64<0x4E>(u);  // lane ^ 2
          }
          if (act && r == 0) lsum += log(u);
        }
        lsum = wave_sum(lsum);
        if (lane == 0) sh->part[wv] = lsum;
        __syncthreads();  // B
        __syncthreads();  // A
        op = sh->op;
      }
    }
  }
  // op == CMD_EXIT
}

// ---------------------------------------------------------------------------
// master wave
// ---------------------------------------------------------------------------
template <typename Real, int R>
struct ProtMaster : ThreeTaxControl<ProtMaster<Real, R>> {
  using Ctl = ThreeTaxControl<ProtMaster<Real, R>>;
  using Ctl::M;
  using Ctl::pend_bl;
  using Ctl::dist_bl;
  using Ctl::prox_bl;
  using Ctl::c_pend;
  using Ctl::c_dist;
  using Ctl::c_prox;
  using Ctl::esum;
  using Ctl::last_ll;
  using Ctl::have_last;
  using Ctl::n_calls;
  using Ctl::n_probes;
  using Ctl::n_dist_probes;
  static constexpr int S = kProtS;
  // FP32 form: state-basis MFMA workers (pp_engine_prot32.h)
  static constexpr bool kF32 = std::is_same<Real, float>::value;
  ProtShared *sh;
  int *sidx;
  int lane, n_workers, cap_sites, n_inf, n_sl;
  double log_R;

  __device__ __forceinline__ void init(char *smem, const DevModel *dm, int cap_pairs) {
    n_sl = dm->n_sl;
    sh = reinterpret_cast<ProtShared *>(smem);
    sidx = ProtLayout::sidx(smem);
    lane = threadIdx.x & 63;
    n_workers = (blockDim.x >> 6) - 1;
    cap_sites = cap_pairs / R;
    n_inf = dm->n_inf;
    log_R = dm->log_R;
    c_pend = c_dist = c_prox = __builtin_nan("");
  }

  __device__ __forceinline__ void command(int op) {
    if (lane == 0) sh->op = op;
    __syncthreads();  // A
    if (op == CMD_EXIT) return;
    if constexpr (kF32) {
      if (op == CMD_PROBE || op == CMD_PROBE_DIST) __syncthreads();  // C: the workers built the FP32 tables
    }
    __syncthreads();  // B
  }

  // mask_reindex of the query's informative sites (core.ml:25-28,146-156),
  // then CMD_QUERY.  Returns M.
  __device__ __forceinline__ int load_query(const uint8_t *codes, int N, int q) {
    const int chunk = (N + 63) / 64;
    const int lo = lane * chunk, hi = min(N, lo + chunk);
    int cnt = 0;
    for (int k = lo; k < hi; k++) cnt += codes[k] < n_inf;
    int incl = cnt;
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
      const int o = __shfl_up(incl, m, 64);
      if (lane >= m) incl += o;
    }
    const int total = __builtin_amdgcn_readfirstlane(__shfl(incl, 63, 64));
    int pos = incl - cnt;
    for (int k = lo; k < hi; k++)
      if (codes[k] < n_inf) sidx[pos++] = k;
    if (lane == 0) {
      sh->M = total;
      sh->q = q;
    }
    M = total;
    if (total > 0 && total <= cap_sites) command(CMD_QUERY);
    return total;
  }

  // Glv.mask_into of darr[loc] / parr[loc] and their exponent totals.
  __device__ __forceinline__ void gather(int loc) {
    if (lane == 0) sh->loc = loc;
    command(CMD_GATHER);
    long long iv[16];
#pragma unroll
    for (int w = 0; w < 16; w++) iv[w] = w < n_workers ? sh->ipart[w] : 0;
    long long es = 0;
#pragma unroll
    for (int w = 0; w < 16; w++) es += iv[w];
    esum = (double)es;
    c_dist = c_prox = __builtin_nan("");
    have_last = false;
  }

  // The exponentials of Diagd.multi_exp for the changed branch lengths.
  __device__ __forceinline__ void tables(bool new_pend, bool new_dist) {
    // one table per transition-matrix slice: rate (gamma) or category (CAT)
    if (new_dist) {
      for (int t = lane; t < 2 * n_sl * S; t += 64) {
        const int which = t / (n_sl * S), rem = t % (n_sl * S), r = rem / S, k = rem % S;
        const double e = exp((sh->rates[r] * (which ? dist_bl : prox_bl)) * sh->l[k]);
        (which ? sh->ed : sh->ep)[rem] = e;
      }
    }
    if (new_pend) {
      for (int t = lane; t < n_sl * S; t += 64) {
        const int r = t / S, k = t % S;
        sh->en[t] = exp((sh->rates[r] * pend_bl) * sh->l[k]);
      }
    }
  }

  // One likelihood probe (Model.log_like3 at the current three-taxon state).
  __device__ __forceinline__ double log_like() {
    const bool new_pend = !(pend_bl == c_pend);
    const bool new_dist = !(dist_bl == c_dist && prox_bl == c_prox);
    tables(new_pend, new_dist);
    if (lane == 0) {
      sh->tflags = (new_dist ? 1 : 0) | (new_pend ? 2 : 0);
      sh->tile_next = 0;
    }
    command(new_dist ? CMD_PROBE_DIST : CMD_PROBE);
    double pv[16];
#pragma unroll
    for (int w = 0; w < 16; w++) pv[w] = w < n_workers ? sh->part[w] : 0.0;
    double tot = pv[0];
#pragma unroll
    for (int w = 1; w < 16; w++) tot += pv[w];
    c_pend = pend_bl;
    c_dist = dist_bl;
    c_prox = prox_bl;
    n_calls++;
    n_probes++;
    n_dist_probes += new_dist;
    const double ll = (tot - (double)M * log_R) + PP_LOG2 * (esum + 0.0);
    last_ll = ll;
    have_last = true;
    return ll;
  }
};

}  // namespace pp
