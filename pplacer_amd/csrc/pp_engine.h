// Three-taxon likelihood engine for nucleotide models (S = 4), FP64 or FP32.
//
// Restates, per workgroup and on block-uniform values:
//   Three_tax.log_like / set_pend_bl / set_dist_bl   three_tax.ml:30-45
//   Model.log_like3 -> ten_log_like3_c                gmix_model.ml:298-304, linear_c.c:367-424
//   Model.evolve_into -> gemmish_c                    gmix_model.ml:308-315, linear_c.c:29-90
//   Diagd.multi_exp -> dediagonalize                  diagd.ml:79-100, linear_c.c:92-146
//   Glv.mask_into                                     gmix_model.ml:121-141
//   Three_tax.optimize, optimize_something            three_tax.ml:48-83
//   Minimization.brent + bisection_start_finder       minimization.ml:11-41,83-100
//   GSL 2.x Brent minimiser (set / init / iterate)    third-party, restated
//   safe_ml_optimize_location                         core.ml:255-281
//
// Arithmetic order follows the C stubs: each evolved entry is a sequential
// sum over j of P[i][j]*src[j]; each site's util is the sequential (rate,
// state) sum of ((pi*prox)*dist)*pend.  The FP64 form rounds every product
// and sum separately (the library is built with -ffp-contract=off); the FP32
// form uses fused multiply-adds.  Per-site logs are summed with a block
// reduction in FP64 (the stubs sum them sequentially).
#pragma once

#include "pp_device.h"

namespace pp {

template <typename Real> struct RealOps;
template <> struct RealOps<double> {
  static __device__ __forceinline__ double madd(double acc, double a, double b) { return acc + a * b; }
  static __device__ __forceinline__ double logd(double u) { return log(u); }
};
template <> struct RealOps<float> {
  static __device__ __forceinline__ float madd(float acc, float a, float b) { return fmaf(a, b, acc); }
  static __device__ __forceinline__ double logd(float u) { return (double)__logf(u); }
};

struct EngineShared {
  // model (LDS copy)
  double x[16], l[4], xit[16], statd[4], rates[kMaxRates];
  double lv[kMaxCodes * 4];
  double red[16];
  long long ired[16];
  unsigned long long skey[16];
  int sidx[16];
  int bcast_i[4];
};

template <typename Real, int R>
struct Engine4 {
  static constexpr int S = 4, RS = R * 4;
  using Ops = RealOps<Real>;

  EngineShared *sh;
  Real *sT;   // [kMaxCodes][RS] pendant GLV rows evolved by pend_bl
  Real *sPp;  // [R][4][4]  P(prox_bl)
  Real *sPd;  // [R][4][4]  P(dist_bl)
  int *sidx;  // [N] masked site indices
  int n_codes, n_inf;
  double log_R;

  // lane data (lane j = masked site j)
  Real dO[RS], pO[RS], w[RS];
  int qc;
  bool active;

  // block-uniform state
  int M;
  double pend_bl, dist_bl, prox_bl;  // Three_tax state (last probe)
  double c_pend, c_dist, c_prox;     // values the caches were built for
  double esum;                       // log2 exponent terms of prox and dist
  double last_ll;
  bool have_last;
  long long n_calls;
  long long n_dist_probes;  // probes that re-evolved the proximal/distal GLVs

  __device__ void init(char *smem, int N, const DevModel *dm) {
    sh = reinterpret_cast<EngineShared *>(smem);
    char *p = smem + ((sizeof(EngineShared) + 15) & ~15);
    sT = reinterpret_cast<Real *>(p);
    p += sizeof(Real) * kMaxCodes * RS;
    sPp = reinterpret_cast<Real *>(p);
    p += sizeof(Real) * R * 16;
    sPd = reinterpret_cast<Real *>(p);
    p += sizeof(Real) * R * 16;
    sidx = reinterpret_cast<int *>(p);
    for (int t = threadIdx.x; t < 16; t += blockDim.x) { sh->x[t] = dm->x[t]; sh->xit[t] = dm->xit[t]; }
    for (int t = threadIdx.x; t < 4; t += blockDim.x) { sh->l[t] = dm->l[t]; sh->statd[t] = dm->statd[t]; }
    for (int t = threadIdx.x; t < R; t += blockDim.x) sh->rates[t] = dm->rates[t];
    for (int t = threadIdx.x; t < dm->n_codes * 4; t += blockDim.x) sh->lv[t] = dm->lv[t];
    n_codes = dm->n_codes;
    n_inf = dm->n_inf;
    log_R = dm->log_R;
    c_pend = c_dist = c_prox = __builtin_nan("");
    __syncthreads();
  }

  static __host__ __device__ __forceinline__ size_t engine_lds_bytes(int N) {
    return ((sizeof(EngineShared) + 15) & ~15) + sizeof(Real) * (kMaxCodes * RS + 2 * R * 16) +
           sizeof(int) * N;
  }

  // Masked-site list of the query (core.ml:146-156, mask_reindex core.ml:25-28)
  // and this lane's pendant code.  Returns M.
  __device__ int load_query(const uint8_t *codes, int N) {
    const int bd = blockDim.x, tid = threadIdx.x;
    const int chunk = (N + bd - 1) / bd;
    const int lo = tid * chunk, hi = min(N, lo + chunk);
    int cnt = 0;
    for (int k = lo; k < hi; k++) cnt += codes[k] < n_inf;
    // exclusive scan of cnt over the block
    long long incl = cnt;
    const int lane = tid & 63, wv = tid >> 6, nw = (bd + 63) >> 6;
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
      long long o = __shfl_up(incl, m, 64);
      if (lane >= m) incl += o;
    }
    if (lane == 63 || tid == bd - 1) sh->ired[wv] = incl;
    __syncthreads();
    long long base = 0, total = 0;
    for (int i = 0; i < nw; i++) {
      if (i < wv) base += sh->ired[i];
      total += sh->ired[i];
    }
    int pos = (int)(base + incl - cnt);
    for (int k = lo; k < hi; k++)
      if (codes[k] < n_inf) sidx[pos++] = k;
    __syncthreads();
    M = (int)total;
    active = tid < M;
    qc = active ? codes[sidx[tid]] : 0;
    return M;
  }

  // Glv.mask_into of darr[loc] / parr[loc] (gmix_model.ml:121-141): lane j
  // takes site sidx[j] of both GLVs ([G][N][R*S] rows) and their exponents.
  __device__ void gather(const Real *darr, const Real *parr, const int *darr_e,
                         const int *parr_e, int N, int loc) {
    long long es = 0;
    if (active) {
      const size_t row = (size_t)loc * N + sidx[threadIdx.x];
      const Real *d = darr + row * RS;
      const Real *p = parr + row * RS;
#pragma unroll
      for (int t = 0; t < RS; t++) { dO[t] = d[t]; pO[t] = p[t]; }
      es = (long long)darr_e[row] + (long long)parr_e[row];
    }
    esum = (double)block_sum_ll(es, sh->ired);
    c_dist = c_prox = __builtin_nan("");
    have_last = false;
    __syncthreads();
  }

  // One likelihood probe: Model.log_like3 of (prox evolved by prox_bl, dist
  // evolved by dist_bl, pend evolved by pend_bl).  Collective.
  __device__ double log_like() {
    const bool new_pend = !(pend_bl == c_pend);
    const bool new_dist = !(dist_bl == c_dist && prox_bl == c_prox);
    const int bd = blockDim.x, tid = threadIdx.x;
    // phase 1: transition matrices (Diagd.multi_exp) and the evolved pendant rows
    const int nP = new_dist ? 2 * R * 16 : 0;
    const int nT = new_pend ? n_codes * RS : 0;
    for (int t = tid; t < nP + nT; t += bd) {
      if (t < nP) {
        const int which = t / (R * 16), rem = t % (R * 16);
        const int r = rem / 16, i = (rem / 4) % 4, j = rem % 4;
        const double rb = sh->rates[r] * (which ? dist_bl : prox_bl);
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++) acc = acc + (exp(rb * sh->l[k]) * sh->x[i * 4 + k]) * sh->xit[j * 4 + k];
        (which ? sPd : sPp)[rem] = (Real)acc;
      } else {
        const int u = t - nP;
        const int c = u / RS, r = (u / 4) % R, i = u % 4;
        const double rb = sh->rates[r] * pend_bl;
        double ex[4];
#pragma unroll
        for (int k = 0; k < 4; k++) ex[k] = exp(rb * sh->l[k]);
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          double pij = 0.0;
#pragma unroll
          for (int k = 0; k < 4; k++) pij = pij + (ex[k] * sh->x[i * 4 + k]) * sh->xit[j * 4 + k];
          acc = acc + pij * sh->lv[c * 4 + j];
        }
        sT[u] = (Real)acc;
      }
    }
    __syncthreads();
    // phase 2: per-site util and log
    double lsum = 0.0;
    if (active) {
      if (new_dist) {
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
          for (int i = 0; i < 4; i++) {
            Real xv = 0, yv = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
              xv = Ops::madd(xv, sPp[r * 16 + i * 4 + j], pO[r * 4 + j]);
              yv = Ops::madd(yv, sPd[r * 16 + i * 4 + j], dO[r * 4 + j]);
            }
            w[r * 4 + i] = ((Real)sh->statd[i] * xv) * yv;
          }
      }
      const Real *z = sT + qc * RS;
      Real util = 0;
#pragma unroll
      for (int t = 0; t < RS; t++) util = Ops::madd(util, w[t], z[t]);
      lsum = Ops::logd(util);
    }
    const double tot = block_sum(lsum, sh->red);
    c_pend = pend_bl;
    c_dist = dist_bl;
    c_prox = prox_bl;
    n_calls++;
    n_dist_probes += new_dist;
    const double ll = (tot - (double)M * log_R) + PP_LOG2 * (esum + 0.0);
    last_ll = ll;
    have_last = true;
    return ll;
  }

  // ---- Minimization.brent with the GSL Brent minimiser ---------------------
  // kind 0: pendant (three_tax.ml:38-39), 1: distal (three_tax.ml:41-45)
  __device__ double fprobe(int kind, double v, int &err) {
    if (kind == 0) {
      pend_bl = v;
    } else {
      const double prox = (dist_bl + prox_bl) - v;
      if (!(prox >= 0.)) err = ST_ASSERT;
      dist_bl = v;
      prox_bl = prox;
    }
    return -log_like();
  }

  __device__ int brent(int kind, double raw_start, double left, double right, double tol,
                       int max_iters, double *result) {
    int err = 0;
    if (left >= raw_start || raw_start >= right) return ST_INVALID_START;
    const double righty = fprobe(kind, right, err);  // (f left, f right) evaluates right first
    if (err) return err;
    const double lefty = fprobe(kind, left, err);
    if (err) return err;
    const double smaller = lefty < righty ? left : right;
    const double miny = lefty <= righty ? lefty : righty;
    double prev = raw_start, pv;
    for (int iter = 1;; iter++) {
      pv = fprobe(kind, prev, err);
      if (err) return err;
      if (pv < miny) break;
      if (fabs(prev - smaller) < tol) { *result = smaller; return ST_OK; }  // FoundMin
      if (iter > max_iters) return ST_FIND_START;
      prev = (prev + smaller) / 2.;
    }
    // gsl_min_fminimizer_set: f(lower), f(upper), f(start) are the values just
    // computed (the three-taxon state is already at start, as after the set).
    double x_lower = left, x_upper = right, x_min = prev;
    double f_lower = lefty, f_upper = righty, f_min = pv;
    if (!isfinite(f_lower) || !isfinite(f_upper) || !isfinite(f_min)) return ST_GSL;
    if (x_lower > x_upper || x_min >= x_upper || x_min <= x_lower) return ST_GSL;
    if (f_min >= f_lower || f_min >= f_upper) return ST_GSL;
    const double golden = 0.3819660;
    double v = x_lower + golden * (x_upper - x_lower), w_ = v;
    double sd = 0., se = 0.;
    double f_v = fprobe(kind, v, err);
    if (err) return err;
    if (!isfinite(f_v)) return ST_GSL;
    double f_w = f_v;
    for (int step = 1;; step++) {
      if (step > max_iters) return ST_MAXITER;
      const double x_left = x_lower, x_right = x_upper, z = x_min, f_z = f_min;
      double d = se, e = sd, u;
      const double w_lower = z - x_left, w_upper = x_right - z;
      const double tl = 1.4901161193847656e-08 * fabs(z);
      double p = 0, q = 0, r = 0;
      const double midpoint = 0.5 * (x_left + x_right);
      if (fabs(e) > tl) {
        r = (z - w_) * (f_z - f_v);
        q = (z - v) * (f_z - f_w);
        p = (z - v) * q - (z - w_) * r;
        q = 2 * (q - r);
        if (q > 0) p = -p; else q = -q;
        r = e;
        e = d;
      }
      if (fabs(p) < fabs(0.5 * q * r) && p < q * w_lower && p < q * w_upper) {
        const double t2 = 2 * tl;
        d = p / q;
        u = z + d;
        if ((u - x_left) < t2 || (x_right - u) < t2) d = (z < midpoint) ? tl : -tl;
      } else {
        e = (z < midpoint) ? x_right - z : -(z - x_left);
        d = golden * e;
      }
      if (fabs(d) >= tl) u = z + d;
      else u = z + ((d > 0) ? tl : -tl);
      se = e;
      sd = d;
      const double f_u = fprobe(kind, u, err);
      if (err) return err;
      if (!isfinite(f_u)) return ST_GSL;
      if (f_u <= f_z) {
        if (u < z) { x_upper = z; f_upper = f_z; }
        else { x_lower = z; f_lower = f_z; }
        v = w_; f_v = f_w; w_ = z; f_w = f_z;
        x_min = u; f_min = f_u;
      } else {
        if (u < z) { x_lower = u; f_lower = f_u; }
        else { x_upper = u; f_upper = f_u; }
        if (f_u <= f_w || w_ == z) { v = w_; f_v = f_w; w_ = u; f_w = f_u; }
        else if (f_u <= f_v || v == z || v == w_) { v = u; f_v = f_u; }
      }
      if (x_upper - x_lower < tol) { *result = x_min; return ST_OK; }
    }
  }

  // three_tax.ml:48-58
  __device__ int optimize_something(int kind, double tol, double start_v, double min_v,
                                    double max_v, double *res) {
    if (min_v == max_v) { *res = start_v; return ST_OK; }
    return brent(kind, start_v, min_v, max_v, tol, 100, res);
  }

  // three_tax.ml:70-83
  __device__ int optimize(double tol, double max_pend, int max_iter) {
    double prev_pend = pend_bl, prev_dist = dist_bl;
    for (int step = 1;; step++) {
      if (step > max_iter) return ST_MAXITER;
      double cp, cd;
      int st = optimize_something(0, tol, pend_bl, 1e-8, max_pend, &cp);
      if (st) return st;
      st = optimize_something(1, tol, dist_bl, 0., dist_bl + prox_bl, &cd);
      if (st) return st;
      if (fabs(prev_pend - cp) > tol || fabs(prev_dist - cd) > tol) {
        prev_pend = cp;
        prev_dist = cd;
      } else {
        return ST_OK;
      }
    }
  }

  // core.ml:225-239 (set_tt_edges; the GLVs are already gathered)
  __device__ void set_state(double cut_bl, double pendant, double distal) {
    pend_bl = pendant;
    dist_bl = distal;
    prox_bl = cut_bl - distal;
    have_last = false;
  }

  // core.ml:255-281: optimize, on a GSL error retry once from the default
  // branch lengths; ExceededMaxIter keeps the current state.  Result is the
  // three-taxon state as Three_tax.get_results reports it (three_tax.ml:85).
  __device__ int safe_optimize(double tol, double max_pend, double start_pend, double cut_bl,
                               double *ll, double *pend, double *dist) {
    for (int attempt = 0; attempt < 2; attempt++) {
      int st = optimize(tol, max_pend, 200);
      if (st == ST_OK || st == ST_MAXITER) {
        *ll = have_last ? last_ll : log_like();
        *pend = pend_bl;
        *dist = dist_bl;
        return st;
      }
      if (st != ST_GSL || attempt == 1) return st;
      set_state(cut_bl, start_pend, cut_bl / 2.);
    }
    return ST_OK;
  }
};

}  // namespace pp
