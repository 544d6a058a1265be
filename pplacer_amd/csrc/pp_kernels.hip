// HIP kernels of the placement path (gfx950).  Launch code is in pp_capi.hip.
#include <climits>

#include "pp_engine.h"
#include "pp_kernels.h"

namespace pp {

// ---------------------------------------------------------------------------
// Stage 1: Glv.masked_logdot of every (query, edge) pair (core.ml:187-194,
// gmix_model.ml:194-198, linear_c.c:465-539).
//
// One lane per edge, TQ queries per workgroup.  The lane streams its edge's
// supernode rows site by site from the site-major array sn[N][Epad][R*4]
// (64 lanes read 64 consecutive rows: one contiguous 8 KB (FP64) burst) and
// reuses each row for the TQ queries of the tile, whose characters sit in
// tcodes[tile][site][TQ] (wave-uniform loads).  Queries are tiled in order
// of their first informative column, so a tile's columns overlap.
// The per-site util is the stub's sequential (rate, state) sum; its log is
// taken through frexp mantissas folded over kLogGroup sites (below).
// ---------------------------------------------------------------------------
template <typename Real> struct Frexp;
template <> struct Frexp<double> {
  static __device__ __forceinline__ double mant(double x) { return __builtin_amdgcn_frexp_mant(x); }
  static __device__ __forceinline__ int expo(double x) { return __builtin_amdgcn_frexp_exp(x); }
  static __device__ __forceinline__ double logd(double x) { return log(x); }
};
template <> struct Frexp<float> {
  static __device__ __forceinline__ float mant(float x) { return __builtin_amdgcn_frexp_mantf(x); }
  static __device__ __forceinline__ int expo(float x) { return __builtin_amdgcn_frexp_expf(x); }
  static __device__ __forceinline__ double logd(float x) { return (double)__logf(x); }
};

// Sites whose utils are folded into one log: the product of kLogGroup
// frexp mantissas stays >= 2^-kLogGroup, and the binary exponents are summed
// as integers, so log(util_1 ... util_k) = log(prod mantissas) + (sum exps) ln 2
// has no underflow and needs one log per kLogGroup sites.
constexpr int kLogGroup = 8;

template <typename Real, int R, int TQ>
__global__ __launch_bounds__(256) void stage1_kernel(Stage1Args a) {
  constexpr int RS = R * 4;
  using Ops = RealOps<Real>;
  using F = Frexp<Real>;
  __shared__ Real tt[kMaxCodes * RS];
  for (int t = threadIdx.x; t < a.n_codes * RS; t += blockDim.x) tt[t] = (Real)a.ttab[t];
  __syncthreads();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;  // < Epad by grid construction
  const int tile = blockIdx.y;
  const int lo = a.tile_lo[tile], hi = a.tile_hi[tile];
  const int n_inf = a.n_inf;
  double ll[TQ];
  Real pm[TQ];   // running product of util mantissas
  int pe[TQ];    // util binary exponents + supernode exponents
  int cnt[TQ];   // sites folded into pm
#pragma unroll
  for (int j = 0; j < TQ; j++) { ll[j] = 0.0; pm[j] = 1; pe[j] = 0; cnt[j] = 0; }
  const Real *snp = reinterpret_cast<const Real *>(a.sn) + ((size_t)lo * a.Epad + e) * RS;
  const int *sep = a.sn_e + (size_t)lo * a.Epad + e;
  const uint8_t *cp = a.tcodes + ((size_t)tile * a.N + lo) * TQ;
  const size_t sstride = (size_t)a.Epad * RS;
  for (int site = lo; site < hi; ++site) {
    Real s[RS];
#pragma unroll
    for (int t = 0; t < RS; t++) s[t] = snp[t];
    const int ex = *sep;
#pragma unroll
    for (int j = 0; j < TQ; j++) {
      const int c = cp[j];
      if (c < n_inf) {
        const Real *z = tt + c * RS;
        Real util = 0;
#pragma unroll
        for (int t = 0; t < RS; t++) util = Ops::madd(util, z[t], s[t]);
        pm[j] *= F::mant(util);
        pe[j] += F::expo(util) + ex;
        if (++cnt[j] == kLogGroup) {
          ll[j] += F::logd(pm[j]);
          pm[j] = 1;
          cnt[j] = 0;
        }
      }
    }
    snp += sstride;
    sep += a.Epad;
    cp += TQ;
  }
  if (e < a.E) {
#pragma unroll
    for (int j = 0; j < TQ; j++) {
      const int q = a.tile_q[tile * TQ + j];
      if (q >= 0) {
        const double l = ll[j] + F::logd(pm[j]);
        a.scores[(size_t)q * a.E + e] = (l - (double)a.qM[q] * a.log_R) + PP_LOG2 * (double)pe[j];
      }
    }
  }
}

// tcodes[tile][site][j] = codes[tile_q[tile*TQ+j]][site] (gap code for empty slots)
__global__ void tile_codes_kernel(const uint8_t *codes, const int *tile_q, int n_tiles, int N,
                                  int TQ, uint8_t gap, uint8_t *tcodes) {
  const size_t total = (size_t)n_tiles * N * TQ;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(i % TQ);
    const size_t ts = i / TQ;
    const int site = (int)(ts % N);
    const int tile = (int)(ts / N);
    const int q = tile_q[tile * TQ + j];
    tcodes[i] = q >= 0 ? codes[(size_t)q * N + site] : gap;
  }
}

// ---------------------------------------------------------------------------
// Stage 2a: baseball over the stage-1 ranking, one workgroup per query from
// a work queue (core.ml:283-320).  The next pitch is the best remaining edge
// of the stable descending sort of the scores (fig.ml:143-147), selected on
// demand; each pitch is Three_tax.optimize from the default branch lengths
// at initial_tolerance.
// ---------------------------------------------------------------------------
template <typename Real, int R, int MAXBD>
__global__ __launch_bounds__(MAXBD) void pitch_kernel(PitchArgs a) {
  extern __shared__ __align__(16) char smem[];
  __shared__ int s_job;
  Engine4<Real, R> eng;
  eng.init(smem, a.N, a.model);
  unsigned *taken = reinterpret_cast<unsigned *>(smem + Engine4<Real, R>::engine_lds_bytes(a.N));
  const Real *darr = reinterpret_cast<const Real *>(a.darr);
  const Real *parr = reinterpret_cast<const Real *>(a.parr);
  const int bd = blockDim.x, tid = threadIdx.x;
  const int words = (a.E + 31) >> 5;
  for (;;) {
    if (tid == 0) s_job = atomicAdd(a.counter, 1);
    __syncthreads();
    const int qi = s_job;
    __syncthreads();
    if (qi >= a.nq) break;
    const int q = a.qlist[qi];
    const int M = eng.load_query(a.codes + (size_t)q * a.N, a.N);
    if (M == 0 || M > bd) {
      if (tid == 0) { a.n_res[q] = 0; a.status[q] = M == 0 ? PP_Q_NO_INFORMATIVE : PP_Q_RANK_EXHAUSTED; }
      continue;
    }
    for (int t = tid; t < words; t += bd) taken[t] = 0u;
    __syncthreads();
    const double *sc = a.scores + (size_t)q * a.E;
    double like_record = -INFINITY;
    long long n_strikes = 0, calls = 0, dprobes = 0;
    int n_res = 0, status = PP_Q_OK;
    for (;;) {
      unsigned long long bk = ~0ull;
      int bi = INT_MAX;
      for (int e = tid; e < a.E; e += bd) {
        if (taken[e >> 5] & (1u << (e & 31))) continue;
        const unsigned long long k = desc_key(sc[e]);
        if (k < bk || (k == bk && e < bi)) { bk = k; bi = e; }
      }
      block_argmin(bk, bi, eng.sh->skey, eng.sh->sidx);
      if (bi == INT_MAX) break;
      if (tid == 0) taken[bi >> 5] |= 1u << (bi & 31);
      const int loc = bi;
      eng.gather(darr, parr, a.darr_e, a.parr_e, a.N, loc);
      const double cut = a.bl[loc];
      eng.set_state(cut, a.start_pend, cut / 2.);
      eng.n_calls = 0;
      eng.n_dist_probes = 0;
      double ll, pend, dist;
      const int st = eng.safe_optimize(a.initial_tol, a.max_pend, a.start_pend, cut, &ll, &pend, &dist);
      calls += eng.n_calls;
      dprobes += eng.n_dist_probes;
      if (st == ST_GSL) continue;  // "GSL problem with location": skipped (core.ml:286-297)
      if (st != ST_OK && st != ST_MAXITER) {
        status = st == ST_FIND_START ? PP_Q_FIND_START
                 : st == ST_INVALID_START ? PP_Q_INVALID_START : PP_Q_ASSERT;
        break;
      }
      if (tid == 0) {
        Res r;
        r.ll = ll; r.pend = pend; r.dist = dist; r.loc = loc; r.flags = st == ST_MAXITER;
        a.res[(size_t)q * a.cap + n_res] = r;
      }
      const int prev_len = n_res;
      n_res++;
      if (prev_len >= a.t_max_pitches) break;
      else if (ll > like_record) like_record = ll;
      else if (ll < like_record - a.strike_box) {
        if (n_strikes + 1 >= a.t_max_strikes) break;
        n_strikes++;
      }
      if (n_res >= a.cap) break;
    }
    if (tid == 0) {
      a.n_res[q] = n_res;
      a.status[q] = (status == PP_Q_OK && n_res == 0) ? PP_Q_NO_RESULTS : status;
      atomicAdd(a.n_calls, (unsigned long long)calls);
      atomicAdd(a.n_pitched, (unsigned long long)n_res);
      atomicAdd(a.n_calls + 2, (unsigned long long)((calls - dprobes) * M));
      atomicAdd(a.n_calls + 3, (unsigned long long)(dprobes * M));
    }
  }
}

// ---------------------------------------------------------------------------
// Keep (core.ml:336-351): stable sort of the pitched results by likelihood,
// descending, keeping i < keep_at_most with ll >= best + log keep_factor.
// One thread per query; only the first keep_at_most ranks are needed.
// ---------------------------------------------------------------------------
__global__ void keep_kernel(KeepArgs a) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.Q) return;
  a.n_keep[q] = 0;
  if (a.status[q] != PP_Q_OK) return;
  const int n = a.n_res[q];
  const Res *r = a.res + (size_t)q * a.cap;
  Res *kept = a.kept + (size_t)q * a.K;
  unsigned long long used_mask[4] = {0, 0, 0, 0};  // up to 256 results tracked exactly
  double best = 0;
  int nk = 0;
  for (int i = 0; i < a.K && i < n; i++) {
    unsigned long long bk = ~0ull;
    int bi = -1;
    for (int j = 0; j < n; j++) {
      if (j < 256 && (used_mask[j >> 6] >> (j & 63) & 1ull)) continue;
      if (j >= 256) {  // rare: more results than the bitmask holds; linear check
        bool used = false;
        for (int t = 0; t < nk; t++) used |= (kept[t].flags >> 8) == j;
        if (used) continue;
      }
      const unsigned long long k = desc_key(r[j].ll);
      if (bi < 0 || k < bk) { bk = k; bi = j; }
    }
    if (bi < 0) break;
    if (bi < 256) used_mask[bi >> 6] |= 1ull << (bi & 63);
    const double ll = r[bi].ll;
    if (i == 0) best = ll;
    if (!(ll >= a.log_keep_factor + best)) break;
    Res k = r[bi];
    k.flags = (bi << 8) | (k.flags & 0xff);
    kept[nk++] = k;
  }
  a.n_keep[q] = nk;
}

// ---------------------------------------------------------------------------
// Three_tax.optimize jobs: one workgroup per (query, edge, start) from a work
// queue.  mode 0: explicit job arrays (pp_optimize_batch); mode 1: the
// final refinement of every kept placement at final_tolerance
// (core.ml:353-369), whose results replace the kept entries.
// ---------------------------------------------------------------------------
template <typename Real, int R, int MAXBD>
__global__ __launch_bounds__(MAXBD) void optimize_kernel(OptArgs a) {
  extern __shared__ __align__(16) char smem[];
  __shared__ int s_job;
  Engine4<Real, R> eng;
  eng.init(smem, a.N, a.model);
  const Real *darr = reinterpret_cast<const Real *>(a.darr);
  const Real *parr = reinterpret_cast<const Real *>(a.parr);
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) s_job = atomicAdd(a.counter, 1);
    __syncthreads();
    const int job = s_job;
    __syncthreads();
    if (job >= a.n_jobs) break;
    int q, loc;
    double pend0, dist0;
    Res *slot = nullptr;
    if (a.mode == 1) {
      const int qi = job / a.K, k = job % a.K;
      q = a.qlist[qi];
      if (k >= a.n_keep[q]) continue;
      slot = a.kept + (size_t)q * a.K + k;
      loc = slot->loc;
      pend0 = slot->pend;
      dist0 = slot->dist;
    } else {
      const int jj = a.qlist ? a.qlist[job] : job;
      q = a.job_q[jj];
      loc = a.job_loc[jj];
      pend0 = a.job_pend[jj];
      dist0 = a.job_dist[jj];
    }
    const int M = eng.load_query(a.codes + (size_t)q * a.N, a.N);
    if (M == 0 || M > (int)blockDim.x) {
      if (tid == 0 && a.mode == 0) {
        const int jj = a.qlist ? a.qlist[job] : job;
        a.out_status[jj] = M == 0 ? PP_Q_NO_INFORMATIVE : PP_Q_RANK_EXHAUSTED;
      }
      continue;
    }
    eng.gather(darr, parr, a.darr_e, a.parr_e, a.N, loc);
    const double cut = a.bl[loc];
    eng.set_state(cut, pend0, dist0);
    eng.n_calls = 0;
    eng.n_dist_probes = 0;
    double ll, pend, dist;
    const int st = eng.safe_optimize(a.tol, a.max_pend, a.start_pend, cut, &ll, &pend, &dist);
    if (tid == 0) {
      atomicAdd(a.n_calls, (unsigned long long)eng.n_calls);
      atomicAdd(a.n_calls + 2, (unsigned long long)((eng.n_calls - eng.n_dist_probes) * M));
      atomicAdd(a.n_calls + 3, (unsigned long long)(eng.n_dist_probes * M));
      if (a.mode == 1) {
        if (st == ST_OK || st == ST_MAXITER) {
          slot->ll = ll;
          slot->pend = pend;
          slot->dist = dist;
        } else if (st != ST_GSL) {
          a.status[q] = st == ST_FIND_START ? PP_Q_FIND_START
                        : st == ST_INVALID_START ? PP_Q_INVALID_START : PP_Q_ASSERT;
        }  // GSL problem in the final optimisation keeps the initial result
      } else {
        const int jj = a.qlist ? a.qlist[job] : job;
        a.out_ll[jj] = ll;
        a.out_pend[jj] = pend;
        a.out_dist[jj] = dist;
        a.out_calls[jj] = (int)eng.n_calls;
        a.out_status[jj] = st;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Like-weight ratios and the final order (core.ml:370-377,405-408;
// ppatteries.ml:175-184; pquery.ml:86-92).  One thread per query.
// ---------------------------------------------------------------------------
__global__ void finalize_kernel(FinalArgs a) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.Q) return;
  const int n = a.status[q] == PP_Q_OK ? a.n_keep[q] : 0;
  const Res *k = a.kept + (size_t)q * a.K;
  double ratio[64];
  int order[64];
  for (int i = 0; i < n; i++) {
    double tot = 0.;
    for (int j = 0; j < n; j++) tot = tot + exp(k[j].ll - k[i].ll);
    ratio[i] = 1. / tot;
    order[i] = i;
  }
  for (int i = 1; i < n; i++) {  // stable insertion sort, ratio descending
    const int v = order[i];
    int j = i - 1;
    while (j >= 0 && desc_key(ratio[order[j]]) > desc_key(ratio[v])) { order[j + 1] = order[j]; j--; }
    order[j + 1] = v;
  }
  pp_placement *out = a.out + (size_t)q * a.K;
  for (int i = 0; i < n; i++) {
    const Res &r = k[order[i]];
    out[i].loc = r.loc;
    out[i].ml_ratio = ratio[order[i]];
    out[i].log_like = r.ll;
    out[i].dist_bl = r.dist;
    out[i].pend_bl = r.pend;
  }
  a.n_out[q] = n;
}

// Top-k of each query's ranking (pp_rank_topk), by the same selection the
// pitch kernel uses.
__global__ __launch_bounds__(256) void topk_kernel(const double *scores, int Q, int E, int k,
                                                   int *ranked) {
  __shared__ unsigned long long skey[16];
  __shared__ int sidx[16];
  extern __shared__ unsigned taken_words[];
  const int q = blockIdx.x;
  if (q >= Q) return;
  const double *sc = scores + (size_t)q * E;
  const int words = (E + 31) >> 5;
  for (int t = threadIdx.x; t < words; t += blockDim.x) taken_words[t] = 0u;
  __syncthreads();
  for (int i = 0; i < k; i++) {
    unsigned long long bk = ~0ull;
    int bi = INT_MAX;
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
      if (taken_words[e >> 5] & (1u << (e & 31))) continue;
      const unsigned long long kk = desc_key(sc[e]);
      if (kk < bk || (kk == bk && e < bi)) { bk = kk; bi = e; }
    }
    block_argmin(bk, bi, skey, sidx);
    if (threadIdx.x == 0) {
      ranked[(size_t)q * k + i] = bi == INT_MAX ? -1 : bi;
      if (bi != INT_MAX) taken_words[bi >> 5] |= 1u << (bi & 31);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// linear_c.c twins (per-call stubs)
// ---------------------------------------------------------------------------
__global__ void gemmish_kernel(double *dst, const double *a, const double *b, int n_sites, int S) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)n_sites * S) return;
  const long site = t / S;
  const int i = (int)(t % S);
  double acc = 0;
  for (int j = 0; j < S; j++) acc = acc + a[i * S + j] * b[site * S + j];
  dst[t] = acc;
}

__global__ void dediag_kernel(double *dst, const double *u, const double *lambda,
                              const double *uit, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * n) return;
  const int i = t / n, j = t % n;
  double acc = 0;
  for (int k = 0; k < n; k++) acc = acc + (lambda[k] * u[i * n + k]) * uit[j * n + k];
  dst[t] = acc;
}

__global__ void pairwise_kernel(double *dst, const double *x, const double *y, long n,
                                const double *statd, int S) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    dst[t] = statd ? (statd[t % S] * x[t]) * y[t] : x[t] * y[t];
}

// per-site util of ten_log_like3_c (statd != null) or ten_*_logdot_c;
// sites [first, first+n_used), optional mask
__global__ void util_kernel(double *util, const double *statd, const double *x, const double *y,
                            const double *z, const uint16_t *mask, int first, int n_used,
                            int R, int N, int S) {
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

// sequential site-order log sum (linear_c.c:416-421 / 527-536)
__global__ void logsum_kernel(const double *util, const uint16_t *mask, int first, int n_used,
                              int R, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double ll = 0;
  int used = 0;
  for (int t = 0; t < n_used; t++)
    if (!mask || mask[first + t]) { ll += log(util[t]); used++; }
  ll -= ((double)(float)used) * log((double)(float)R);
  *out = ll;
}

}  // namespace pp

// ---------------------------------------------------------------------------
// Reference-tree precompute (Like_stree.calc_distal_and_proximal,
// like_stree.ml:43-110; perhaps_pull_exponent gmix_model.ml:163-187;
// Glv_arr.prep_supernodes glv_arr.ml:46-56), level-synchronous: one launch
// per post-order height (distal) and per pre-order depth (proximal), one
// lane per (node, site).  GLV rows are [node][site][R*S] FP64; P tables are
// FP64 [node][R][S][S] from the host's Diagd.multi_exp.  The arithmetic is
// the stubs': products in list order, sequential gemmish sums, exact
// power-of-two exponent pulls.
// ---------------------------------------------------------------------------
namespace pre {

template <int RS, int S>
__device__ __forceinline__ void pull(double *v, int &e, int min_twoexp) {
  constexpr int R = RS / S;
  int mx_te = INT_MIN;
#pragma unroll
  for (int r = 0; r < R; r++) {
    double mx = v[r * S];
#pragma unroll
    for (int k = 1; k < S; k++)
      if (v[r * S + k] > mx) mx = v[r * S + k];
    int te;
    (void)frexp(mx, &te);
    if (te > mx_te) mx_te = te;
  }
  if (mx_te < min_twoexp) {
    const double f = ldexp(0.5, -mx_te + 1);
#pragma unroll
    for (int t = 0; t < RS; t++) v[t] *= f;
    e += mx_te;
  }
}

template <int RS, int S>
__device__ __forceinline__ void evolve_row(const double *P, const double *src, double *dst) {
  constexpr int R = RS / S;
#pragma unroll
  for (int r = 0; r < R; r++)
#pragma unroll
    for (int i = 0; i < S; i++) {
      double acc = 0;
#pragma unroll
      for (int j = 0; j < S; j++) acc = acc + P[(r * S + i) * S + j] * src[r * S + j];
      dst[r * S + i] = acc;
    }
}

struct TreeArgs {
  int N, n_nodes;
  const int *level_nodes;   // node ids of this launch
  int n_level;
  const int *child_start, *children, *parent;
  const double *P, *Phalf;  // [G][R][S][S]
  const uint8_t *leaf_codes;  // [G][N]
  const double *lv;           // [n_codes][S]
  double *darr, *parr, *evd;  // [G][N][RS]
  int *darr_e, *parr_e, *evd_e;
};

// post-order: distal = product of the children's evolved distals (or the
// leaf's LV), pulled at 2^-50, then evolv_dist = P(bl) * distal
template <int RS, int S>
__global__ void distal_level_kernel(TreeArgs a) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)a.n_level * a.N) return;
  const int site = (int)(t % a.N);
  const int node = a.level_nodes[t / a.N];
  const size_t row = (size_t)node * a.N + site;
  double v[RS];
  int e = 0;
  const int c0 = a.child_start[node], c1 = a.child_start[node + 1];
  if (c1 == c0) {
    const int code = a.leaf_codes[row];
#pragma unroll
    for (int r = 0; r < RS / S; r++)
#pragma unroll
      for (int k = 0; k < S; k++) v[r * S + k] = a.lv[code * S + k];
  } else {
    const size_t r0 = (size_t)a.children[c0] * a.N + site;
#pragma unroll
    for (int k = 0; k < RS; k++) v[k] = a.evd[r0 * RS + k];
    e = a.evd_e[r0];
    for (int c = c0 + 1; c < c1; c++) {
      const size_t rc = (size_t)a.children[c] * a.N + site;
#pragma unroll
      for (int k = 0; k < RS; k++) v[k] = v[k] * a.evd[rc * RS + k];
      e += a.evd_e[rc];
    }
    pull<RS, S>(v, e, -50);
  }
#pragma unroll
  for (int k = 0; k < RS; k++) a.darr[row * RS + k] = v[k];
  a.darr_e[row] = e;
  double w[RS];
  evolve_row<RS, S>(a.P + (size_t)node * RS * S, v, w);
#pragma unroll
  for (int k = 0; k < RS; k++) a.evd[row * RS + k] = w[k];
  a.evd_e[row] = e;
}

// pre-order: proximal of a child = P(bl_parent) * proximal(parent) times the
// evolved distals of its siblings in order, pulled at 2^-50
template <int RS, int S>
__global__ void proximal_level_kernel(TreeArgs a) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)a.n_level * a.N) return;
  const int site = (int)(t % a.N);
  const int child = a.level_nodes[t / a.N];
  const int par = a.parent[child];
  const size_t prow = (size_t)par * a.N + site;
  double src[RS], v[RS];
#pragma unroll
  for (int k = 0; k < RS; k++) src[k] = a.parr[prow * RS + k];
  int e = a.parr_e[prow];
  evolve_row<RS, S>(a.P + (size_t)par * RS * S, src, v);
  for (int c = a.child_start[par]; c < a.child_start[par + 1]; c++) {
    const int sib = a.children[c];
    if (sib == child) continue;
    const size_t rs = (size_t)sib * a.N + site;
#pragma unroll
    for (int k = 0; k < RS; k++) v[k] = v[k] * a.evd[rs * RS + k];
    e += a.evd_e[rs];
  }
  pull<RS, S>(v, e, -50);
  const size_t row = (size_t)child * a.N + site;
#pragma unroll
  for (int k = 0; k < RS; k++) a.parr[row * RS + k] = v[k];
  a.parr_e[row] = e;
}

__global__ void fill_unit_kernel(double *row, int *e, int N, int RS) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < (long)N * RS; t += (long)gridDim.x * blockDim.x) {
    row[t] = 1.0;
    if (t < N) e[t] = 0;
  }
}

// pull both arrays at 2^-10 (pplacer_run.ml:331-336), then the supernodes
// pi * P(bl/2) darr * P(bl/2) parr (glv_arr.ml:46-56), written site-major
// for stage 1, with FP32 copies of everything
template <int RS, int S>
__global__ void finish_kernel(TreeArgs a, const double *statd, int E, int Epad, double *sn64,
                              int *sn_e) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)a.n_nodes * a.N) return;
  const int site = (int)(t % a.N);
  const int node = (int)(t / a.N);
  const size_t row = (size_t)node * a.N + site;
  double d[RS], p[RS];
#pragma unroll
  for (int k = 0; k < RS; k++) { d[k] = a.darr[row * RS + k]; p[k] = a.parr[row * RS + k]; }
  int ed = a.darr_e[row], ep = a.parr_e[row];
  pull<RS, S>(d, ed, -10);
  pull<RS, S>(p, ep, -10);
#pragma unroll
  for (int k = 0; k < RS; k++) {
    a.darr[row * RS + k] = d[k];
    a.parr[row * RS + k] = p[k];
  }
  a.darr_e[row] = ed;
  a.parr_e[row] = ep;
  if (node < E) {
    double ud[RS], up[RS];
    const double *Ph = a.Phalf + (size_t)node * RS * S;
    evolve_row<RS, S>(Ph, d, ud);
    evolve_row<RS, S>(Ph, p, up);
    const size_t srow = (size_t)site * Epad + node;
#pragma unroll
    for (int k = 0; k < RS; k++) sn64[srow * RS + k] = (statd[k % S] * ud[k]) * up[k];
    sn_e[srow] = ed + ep;
  }
}

}  // namespace pre

// FP32 copy of a GLV array with its own per-site power-of-two rescaling: each
// row of R*S values is scaled by 2^-k, k the frexp exponent of the row's
// largest value, and k is added to the row's exponent.  The FP64 arrays keep
// the reference's pulls (gmix_model.ml:163-187), which stop pulling a site
// once one rate's vector is exactly zero (frexp 0 = exponent 0); their values
// can then sit far below FP32's range, hence the separate scaling.
__global__ void fp32_rescale_kernel(const double *src, const int *e, float *dst, int *e32,
                                    long rows, int RS) {
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < rows;
       r += (long)gridDim.x * blockDim.x) {
    const double *v = src + r * RS;
    double mx = 0.0;
    for (int k = 0; k < RS; k++) mx = v[k] > mx ? v[k] : mx;
    int te = 0;
    if (mx > 0.0) (void)frexp(mx, &te);
    const double sc = ldexp(1.0, -te);
    for (int k = 0; k < RS; k++) dst[r * RS + k] = (float)(v[k] * sc);
    e32[r] = e[r] + te;
  }
}
