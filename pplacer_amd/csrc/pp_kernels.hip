// HIP kernels of the placement path (gfx950).  Launch code is in pp_capi.hip.
#include <climits>

#include "pp_engine.h"
#include "pp_engine_mw.h"
#include "pp_engine_prot.h"
#include "pp_engine_prot32.h"
#include "pp_kernels.h"
#include "pp_twin_kernels.h"

namespace pp {

// ---------------------------------------------------------------------------
// Stage 1: Glv.masked_logdot of every (query, edge) pair (core.ml:187-194,
// gmix_model.ml:194-198, linear_c.c:465-539).
//
// A query's stage-1 GLV at a site is its character's LV evolved by
// start_pend, so the per-site util of (query, edge) depends only on
// (site, character, edge).  code_log_kernel evaluates log(util) once for
// every informative character, in the stub's order (rates outer, states
// inner, FP64), into lt[N][C1][Epad] (C1 = n_inf + 1; row n_inf holds 0.0 for
// non-informative characters).  stage1_kernel then sums, per (query, edge),
// the rows its characters select, site by site in column order: the same
// sequential FP64 sum of the same logs as ten_masked_logdot_c.  The supernode
// exponents are summed as integers alongside.
// ---------------------------------------------------------------------------
template <int RS>
__global__ __launch_bounds__(256) void code_log_kernel(CodeLogArgs a) {
  __shared__ double tt[kMaxCodes * RS];
  const int site = blockIdx.y;
  const double *tab = a.ttab + (a.site_cat ? (size_t)a.site_cat[site] * a.n_codes * RS : 0);
  for (int t = threadIdx.x; t < a.n_inf * RS; t += blockDim.x) tt[t] = tab[t];
  __syncthreads();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;  // < Epad by grid construction
  double s[RS];
  const double *sp = a.sn + ((size_t)site * a.Epad + e) * RS;
#pragma unroll
  for (int t = 0; t < RS; t++) s[t] = sp[t];
  const size_t base = (size_t)site * (a.n_inf + 1) * a.Epad + e;
  for (int c = 0; c < a.n_inf; c++) {
    const double *z = tt + c * RS;
    double util = 0.0;
#pragma unroll
    for (int t = 0; t < RS; t++) util = util + z[t] * s[t];
    // the FP64 table's logs correctly rounded, as the stubs' libm's
    const double l = a.lt32 ? log(util) : cr::log(util);
    if (a.lt32) a.lt32[base + (size_t)c * a.Epad] = (float)l;
    else a.lt64[base + (size_t)c * a.Epad] = l;
  }
  if (a.lt32) a.lt32[base + (size_t)a.n_inf * a.Epad] = 0.0f;
  else a.lt64[base + (size_t)a.n_inf * a.Epad] = 0.0;
}

// Character rows are read with buffer loads: the lane's byte offset is fixed,
// the row offset is a scalar, and the resource is rebased once per site.
template <typename Real> struct RowLoad;
template <> struct RowLoad<double> {
  static __device__ __forceinline__ double get(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  }
};
template <> struct RowLoad<float> {
  static __device__ __forceinline__ float get(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  }
};
constexpr int kBufferDword3 = 0x00020000;  // gfx9 raw buffer resource word 3

// The grid is one-dimensional, 8 * per_xcd workgroups.  The device deals
// workgroups to its 8 XCDs round robin, so workgroup b runs on XCD b % 8;
// XCD k takes the work items [k * per_xcd, (k + 1) * per_xcd) of the order
// (edge tile, query tile), query tile fastest.  The workgroups resident on
// one XCD at a time then share an edge tile and neighbouring query tiles,
// whose spans overlap (the host plan groups queries by span), so they read
// the same table rows through that XCD's L2.  Items past the end exit.
template <typename Real, int TQ>
__global__ __launch_bounds__(256) void stage1_kernel(Stage1Args a) {
  static_assert(TQ % 4 == 0, "codes are read four per dword");
  const int item = (int)(blockIdx.x % 8u) * a.per_xcd + (int)(blockIdx.x / 8u);
  if (item >= a.n_tiles * (a.Epad / 256)) return;
  const int tile = item % a.n_tiles;
  const int e = (item / a.n_tiles) * 256 + threadIdx.x;  // < Epad
  const int lo = a.tile_lo[tile], hi = a.tile_hi[tile];
  const unsigned gap = (unsigned)(a.C1 - 1);
  double ll[TQ];
  int pe[TQ];
#pragma unroll
  for (int j = 0; j < TQ; j++) { ll[j] = 0.0; pe[j] = 0; }
  const unsigned rowb = (unsigned)a.Epad * sizeof(Real);  // bytes between character rows
  const size_t siteb = (size_t)a.C1 * rowb;               // bytes per site
  const char *ltp = reinterpret_cast<const char *>(a.lt) + (size_t)lo * siteb;
  const unsigned voff = (unsigned)e * sizeof(Real);
  const int *sep = a.sn_e + (size_t)lo * a.Epad + e;
  const unsigned *cp = reinterpret_cast<const unsigned *>(a.tcodes + ((size_t)tile * a.N + lo) * TQ);
  for (int site = lo; site < hi; ++site) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(ltp), (short)0, (int)siteb, kBufferDword3);
    unsigned cj[TQ];
#pragma unroll
    for (int w = 0; w < TQ / 4; w++) {
      const unsigned word = __builtin_amdgcn_readfirstlane(cp[w]);
#pragma unroll
      for (int b = 0; b < 4; b++) cj[4 * w + b] = (word >> (8 * b)) & 0xffu;
    }
    Real v[TQ];
#pragma unroll
    for (int j = 0; j < TQ; j++) {
      v[j] = 0;
      if (cj[j] != gap) v[j] = RowLoad<Real>::get(rs, voff, cj[j] * rowb);
    }
    const int ex = *sep;
#pragma unroll
    for (int j = 0; j < TQ; j++) {
      ll[j] += (double)v[j];
      pe[j] += cj[j] != gap ? ex : 0;
    }
    ltp += siteb;
    sep += a.Epad;
    cp += TQ / 4;
  }
  if (e < a.E) {
#pragma unroll
    for (int j = 0; j < TQ; j++) {
      const int q = a.tile_q[tile * TQ + j];
      if (q >= 0)
        a.scores[(size_t)q * a.E + e] = (ll[j] - (double)a.qM[q] * a.log_R) + PP_LOG2 * (double)pe[j];
    }
  }
}

// tcodes[tile][site][j] = codes[tile_q[tile*TQ+j]][site]; empty slots and
// non-informative characters read the gap code (= n_inf)
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
    const uint8_t c = q >= 0 ? codes[(size_t)q * N + site] : gap;
    tcodes[i] = c < gap ? c : gap;
  }
}

// ---------------------------------------------------------------------------
// Stage 2 (pp_engine_mw.h): wave 0 of a workgroup runs the control, waves
// 1.. hold SPL masked sites per lane.
//
// pitch_mw_kernel: baseball over the stage-1 ranking, one workgroup per query
// from a work queue (core.ml:283-320).  The next pitch is the best remaining
// edge of the stable descending sort of the scores (fig.ml:143-147), selected
// on demand; each pitch is Three_tax.optimize from the default branch lengths
// at initial_tolerance.
//
// optimize_mw_kernel: Three_tax.optimize jobs, one workgroup per (query,
// edge, start) from a work queue.  mode 0: explicit job arrays
// (pp_optimize_batch); mode 1: the final refinement of every kept placement
// at final_tolerance (core.ml:353-369), whose results replace the kept entries.
// ---------------------------------------------------------------------------

// wave-level (key, edge) minimum, ties by smaller edge id; uniform result
__device__ __forceinline__ void wave_argmin(unsigned long long &key, int &idx) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const unsigned long long k2 = __shfl_xor(key, m, 64);
    const int i2 = __shfl_xor(idx, m, 64);
    if (k2 < key || (k2 == key && i2 < idx)) { key = k2; idx = i2; }
  }
}

// Per-edge "already pitched" flags of the baseball loop, one bit per edge:
// lane l owns edges e = l + 64 j, bit j % 32 of word taken[(j / 32) * 64 + l]
// (a lane reads and writes only its own words).
__host__ __device__ __forceinline__ size_t taken_bytes(int E) { return (size_t)256 * ((E + 2047) / 2048); }

// Baseball over one work queue of queries, for either engine's master.
template <class Master>
__device__ __forceinline__ void pitch_loop(const PitchArgs &a, Master &m, unsigned *taken) {
  const int lane = m.lane;
  for (;;) {
    int qi = 0;
    if (lane == 0) qi = atomicAdd(a.counter, 1);
    qi = __builtin_amdgcn_readfirstlane(qi);  // wave-uniform: branches below hold barriers
    if (qi >= a.nq) break;
    const int q = a.qlist[qi];
    const int M = m.load_query(a.codes + (size_t)q * a.N, a.N, q);
    if (M == 0 || M > m.cap_sites) {
      if (lane == 0) { a.n_res[q] = 0; a.status[q] = M == 0 ? PP_Q_NO_INFORMATIVE : PP_Q_RANK_EXHAUSTED; }
      continue;
    }
    const int n_words = (int)(taken_bytes(a.E) / 256);
    for (int w = 0; w < n_words; w++) taken[w * 64 + lane] = 0u;
    const double *sc = a.scores + (size_t)q * a.E;
    double like_record = -INFINITY;
    long long n_strikes = 0, calls = 0, probes = 0, dprobes = 0;
    int n_res = 0, status = PP_Q_OK;
    for (;;) {
      // next edge of the stable descending sort of the scores (fig.ml:143-147;
      // a nan score sorts last but is still pitched): bi == INT_MAX is the
      // only "none left" marker
      unsigned long long bk = ~0ull;
      int bi = INT_MAX;
      for (int e = lane, j = 0; e < a.E; e += 64, j++) {
        if ((taken[(j >> 5) * 64 + lane] >> (j & 31)) & 1u) continue;
        const unsigned long long k = desc_key(sc[e]);
        if (k < bk || (k == bk && e < bi)) { bk = k; bi = e; }
      }
      wave_argmin(bk, bi);
      bi = __builtin_amdgcn_readfirstlane(bi);
      if (bi == INT_MAX) break;
      if (lane == (bi & 63)) {
        const int j = bi >> 6;
        taken[(j >> 5) * 64 + lane] |= 1u << (j & 31);
      }
      const int loc = bi;
      m.gather(loc);
      const double cut = a.bl[loc];
      m.set_state(cut, a.start_pend, cut / 2.);
      m.n_calls = 0;
      m.n_probes = 0;
      m.n_dist_probes = 0;
      double ll, pend, dist;
      const int st = m.safe_optimize(a.initial_tol, a.max_pend, a.start_pend, cut, &ll, &pend, &dist);
      calls += m.n_calls;
      probes += m.n_probes;
      dprobes += m.n_dist_probes;
      if (is_gsl(st)) continue;  // "GSL problem with location": skipped (core.ml:286-297)
      if (st != ST_OK && st != ST_MAXITER) {
        status = st == ST_FIND_START ? PP_Q_FIND_START
                 : st == ST_INVALID_START ? PP_Q_INVALID_START : PP_Q_ASSERT;
        break;
      }
      if (lane == 0) {
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
    if (lane == 0) {
      a.n_res[q] = n_res;
      a.status[q] = (status == PP_Q_OK && n_res == 0) ? PP_Q_NO_RESULTS : status;
      atomicAdd(a.n_calls, (unsigned long long)calls);
      atomicAdd(a.n_pitched, (unsigned long long)n_res);
      atomicAdd(a.n_calls + 2, (unsigned long long)((probes - dprobes) * M));
      atomicAdd(a.n_calls + 3, (unsigned long long)(dprobes * M));
      atomicAdd(a.n_calls + 4, (unsigned long long)((probes - dprobes) * M));
      atomicAdd(a.n_calls + 5, (unsigned long long)(dprobes * M));
    }
  }
  m.command(CMD_EXIT);
}

// Waves per SIMD the 4-state stage-2 kernels are built for.  Latency, not
// issue, bounds a probe, so occupancy pays: with one FP64 site or two FP32
// sites per worker lane the rows fit the 128 VGPRs of 4 waves per SIMD, and
// the master wave's control state spills only outside the per-probe loop.
// Two FP64 sites per lane (the longest queries) keep the launch bound's budget.
// SPL = 0 is the streaming form (any masked length, 1024-lane blocks).
template <typename Real, int SPL>
struct MWOcc {
  static constexpr int waves = (sizeof(Real) == 8 && SPL == 2) ? 1 : 4;
};

// The longest masked length a workgroup takes: its worker lanes times SPL,
// or in the streaming form the capacity of its scratch slice.
template <int SPL>
__device__ __forceinline__ int mw_cap(int stream_cap) {
  return SPL == 0 ? stream_cap : 64 * SPL * ((int)(blockDim.x >> 6) - 1);
}

// The worker waves of a 4-state stage-2 workgroup: register-resident rows
// (SPL >= 1, cached products in LDS at wl) or the streaming form (SPL == 0,
// products in the workgroup's slice of the global scratch).
template <typename Real, int R, int SPL, bool CAT, bool STRICT>
__device__ __forceinline__ void mw_workers(char *smem, const MWLayout<Real, R, STRICT> &lay, char *wl, const DevModel *,
                                           int N, const uint8_t *codes, const void *darr, const void *parr,
                                           const int *darr_e, const int *parr_e, const int *site_cat,
                                           void *scratch, int cap) {
  const Real *d = reinterpret_cast<const Real *>(darr);
  const Real *p = reinterpret_cast<const Real *>(parr);
  if constexpr (SPL == 0) {
    Real *prod = reinterpret_cast<Real *>(scratch) + (size_t)blockIdx.x * cap * (R * 4);
    mw_worker_stream<Real, R, CAT, STRICT>(smem, lay, prod, cap, N, codes, d, p, darr_e, parr_e, site_cat);
  } else {
    mw_worker<Real, R, SPL, CAT, STRICT>(smem, lay, wl, N, codes, d, p, darr_e, parr_e, site_cat);
  }
}

template <typename Real, int R, int MAXBD, int SPL, bool CAT, bool STRICT>
__global__ __launch_bounds__(MAXBD) __attribute__((amdgpu_waves_per_eu(MWOcc<Real, SPL>::waves)))
void pitch_mw_kernel(PitchArgs a) {
  extern __shared__ __align__(16) char smem[];
  const int cap = mw_cap<SPL>(a.cap_sites);
  const MWLayout<Real, R, STRICT> lay = mw_layout<Real, R, STRICT>(a.model, cap, SPL == 0);
  MWShared *sh = reinterpret_cast<MWShared *>(smem);
  mw_copy_model(sh, a.model);
  __syncthreads();
  unsigned *taken = reinterpret_cast<unsigned *>(smem + lay.align(lay.bytes()));
  if (wave_id() != 0) {
    char *wl = smem + lay.align(lay.align(lay.bytes()) + taken_bytes(a.E));
    mw_workers<Real, R, SPL, CAT, STRICT>(smem, lay, wl, a.model, a.N, a.codes, a.darr, a.parr, a.darr_e, a.parr_e,
                                  a.site_cat, a.scratch, a.cap_sites);
    return;
  }
  MWMaster<Real, R, SPL, CAT, STRICT> m;
  m.init(smem, a.model, cap);
  pitch_loop(a, m, taken);
}

// The 20-state workers of either form: FP64 in the eigenbasis
// (pp_engine_prot.h), FP32 in the state basis on the matrix cores
// (pp_engine_prot32.h).  Returns the LDS byte past the engine's layout.
template <typename Real, int R>
__device__ __forceinline__ size_t prot_setup(char *smem, const DevModel *dm, int N) {
  if constexpr (std::is_same<Real, float>::value) {
    prot32_copy_model<R>(smem, dm, N);
    __syncthreads();
    return ProtLayout::align(Prot32Layout<R>::bytes(N));
  } else {
    prot_copy_model(reinterpret_cast<ProtShared *>(smem), dm);
    __syncthreads();
    return ProtLayout::align(ProtLayout::bytes(N));
  }
}

template <typename Real, int R, class Args>
__device__ __forceinline__ void prot_workers(char *smem, const Args &a) {
  if constexpr (std::is_same<Real, float>::value) {
    float *scr = reinterpret_cast<float *>(a.scratch) +
                 Prot32Layout<R>::scratch_floats(a.cap_pairs / 64) * blockIdx.x;
    prot32_worker<R>(smem, scr, a.N, a.codes, reinterpret_cast<const float *>(a.darr),
                     reinterpret_cast<const float *>(a.parr), a.darr_e, a.parr_e);
  } else {
    double *scr = reinterpret_cast<double *>(a.scratch) + ProtLayout::scratch_reals(a.cap_pairs) * blockIdx.x;
    prot_worker<double, R>(smem, scr, a.cap_pairs, a.N, a.codes, reinterpret_cast<const double *>(a.darr),
                           reinterpret_cast<const double *>(a.parr), a.darr_e, a.parr_e, a.site_cat);
  }
}

// Waves per SIMD the 20-state stage-2 kernels are built for: the FP64
// engine's pair temporaries take the launch bound's budget (one workgroup per
// CU); the FP32 form fits 128 VGPRs, so two workgroups share a CU and one's
// table build and control overlap the other's matrix-core work.
template <typename Real>
struct ProtOcc {
  static constexpr int waves = std::is_same<Real, float>::value ? 4 : 2;
};

template <typename Real, int R>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(ProtOcc<Real>::waves)))
void pitch_prot_kernel(PitchArgs a) {
  extern __shared__ __align__(16) char smem[];
  const size_t end = prot_setup<Real, R>(smem, a.model, a.N);
  unsigned *taken = reinterpret_cast<unsigned *>(smem + end);
  if (wave_id() != 0) {
    prot_workers<Real, R>(smem, a);
    return;
  }
  ProtMaster<Real, R> m;
  m.init(smem, a.model, a.cap_pairs);
  pitch_loop(a, m, taken);
}

// Three_tax.optimize jobs of one work queue, for either engine's master.
template <class Master>
__device__ __forceinline__ void optimize_loop(const OptArgs &a, Master &m) {
  const int lane = m.lane;
  for (;;) {
    int job = 0;
    if (lane == 0) job = atomicAdd(a.counter, 1);
    job = __builtin_amdgcn_readfirstlane(job);
    if (job >= a.n_jobs) break;
    int q, loc;
    double pend0, dist0;
    Res *slot = nullptr;
    int jj = 0;
    if (a.mode == 1) {
      const int qi = job / a.K, k = job % a.K;
      q = a.qlist[qi];
      if (k >= a.n_keep[q]) continue;
      slot = a.kept + (size_t)q * a.K + k;
      loc = slot->loc;
      pend0 = slot->pend;
      dist0 = slot->dist;
    } else {
      jj = a.qlist ? a.qlist[job] : job;
      q = a.job_q[jj];
      loc = a.job_loc[jj];
      pend0 = a.job_pend[jj];
      dist0 = a.job_dist[jj];
    }
    const int M = m.load_query(a.codes + (size_t)q * a.N, a.N, q);
    if (M == 0 || M > m.cap_sites) {
      if (lane == 0 && a.mode == 0) a.out_status[jj] = M == 0 ? PP_Q_NO_INFORMATIVE : PP_Q_RANK_EXHAUSTED;
      continue;
    }
    m.gather(loc);
    const double cut = a.bl[loc];
    m.set_state(cut, pend0, dist0);
    m.n_calls = 0;
    m.n_probes = 0;
    m.n_dist_probes = 0;
    double ll, pend, dist;
    const int st = m.safe_optimize(a.tol, a.max_pend, a.start_pend, cut, &ll, &pend, &dist);
    if (lane == 0) {
      atomicAdd(a.n_calls, (unsigned long long)m.n_calls);
      atomicAdd(a.n_calls + 2, (unsigned long long)((m.n_probes - m.n_dist_probes) * M));
      atomicAdd(a.n_calls + 3, (unsigned long long)(m.n_dist_probes * M));
      if (a.mode == 1) {
        if (st == ST_OK || st == ST_MAXITER) {
          slot->ll = ll;
          slot->pend = pend;
          slot->dist = dist;
        } else if (!is_gsl(st)) {
          a.status[q] = st == ST_FIND_START ? PP_Q_FIND_START
                        : st == ST_INVALID_START ? PP_Q_INVALID_START : PP_Q_ASSERT;
        }  // a GSL problem in the final optimisation keeps the initial result
      } else {
        a.out_ll[jj] = ll;
        a.out_pend[jj] = pend;
        a.out_dist[jj] = dist;
        a.out_calls[jj] = (int)m.n_calls;
        a.out_status[jj] = st;
      }
    }
  }
  m.command(CMD_EXIT);
}

// Marginal likelihoods for -p (core.ml:378-392): per kept placement (mode 1)
// or per job (mode 0), Three_tax.calc_marg_prob from the placement's branch
// lengths with the prior of its edge.  FP64 only, as the final likelihoods.
template <class Master>
__device__ __forceinline__ void marginal_loop(const OptArgs &a, Master &m) {
  const int lane = m.lane;
#pragma unroll 1
  for (;;) {
    int job = 0;
    if (lane == 0) job = atomicAdd(a.counter, 1);
    job = __builtin_amdgcn_readfirstlane(job);
    if (job >= a.n_jobs) break;
    int q, loc;
    double pend0, dist0;
    Res *slot = nullptr;
    int jj = 0;
    if (a.mode == 1) {
      const int qi = job / a.K, k = job % a.K;
      q = a.qlist[qi];
      if (a.status[q] != PP_Q_OK || k >= a.n_keep[q]) continue;
      slot = a.kept + (size_t)q * a.K + k;
      loc = slot->loc;
      pend0 = slot->pend;
      dist0 = slot->dist;
    } else {
      jj = a.qlist ? a.qlist[job] : job;
      q = a.job_q[jj];
      loc = a.job_loc[jj];
      pend0 = a.job_pend[jj];
      dist0 = a.job_dist[jj];
    }
    const int M = m.load_query(a.codes + (size_t)q * a.N, a.N, q);
    if (M == 0 || M > m.cap_sites) {
      if (lane == 0 && a.mode == 0) a.out_status[jj] = M == 0 ? PP_Q_NO_INFORMATIVE : PP_Q_RANK_EXHAUSTED;
      continue;
    }
    m.gather(loc);
    m.set_state(a.bl[loc], pend0, dist0);  // tt_edges_from_placement (core.ml:241-250)
    m.pr_kind = a.prior_kind;
    m.pr_mu = a.prior_mean[loc];
    m.n_calls = 0;
    m.n_probes = 0;
    m.n_dist_probes = 0;
    double marg = 0.;
    const int st = m.calc_marg_prob(a.pp_rel_err, a.max_pend, &marg);
    if (lane == 0) {
      atomicAdd(a.n_calls, (unsigned long long)m.n_calls);
      atomicAdd(a.n_calls + 2, (unsigned long long)((m.n_probes - m.n_dist_probes) * M));
      atomicAdd(a.n_calls + 3, (unsigned long long)(m.n_dist_probes * M));
      if (a.mode == 1) {
        if (st == ST_OK) slot->marg = marg;
        else  // uncaught in the reference: the query fails
          a.status[q] = st == ST_MAXITER ? PP_Q_MAXITER
                        : st == ST_FIND_START ? PP_Q_FIND_START
                        : st == ST_INVALID_START ? PP_Q_INVALID_START
                        : st == ST_ASSERT ? PP_Q_ASSERT : PP_Q_GSL;
      } else {
        a.out_ll[jj] = marg;
        a.out_calls[jj] = (int)m.n_calls;
        a.out_status[jj] = st;
      }
    }
  }
  m.command(CMD_EXIT);
}

template <typename Real, int R, int MAXBD, int SPL, bool CAT, bool STRICT>
__global__ __launch_bounds__(MAXBD) __attribute__((amdgpu_waves_per_eu(MWOcc<Real, SPL>::waves)))
void optimize_mw_kernel(OptArgs a) {
  extern __shared__ __align__(16) char smem[];
  const int cap = mw_cap<SPL>(a.cap_sites);
  const MWLayout<Real, R, STRICT> lay = mw_layout<Real, R, STRICT>(a.model, cap, SPL == 0);
  MWShared *sh = reinterpret_cast<MWShared *>(smem);
  mw_copy_model(sh, a.model);
  __syncthreads();
  if (wave_id() != 0) {
    char *wl = smem + lay.align(lay.bytes());
    mw_workers<Real, R, SPL, CAT, STRICT>(smem, lay, wl, a.model, a.N, a.codes, a.darr, a.parr, a.darr_e, a.parr_e,
                                  a.site_cat, a.scratch, a.cap_sites);
    return;
  }
  MWMaster<Real, R, SPL, CAT, STRICT> m;
  m.init(smem, a.model, cap);
  optimize_loop(a, m);
}

template <typename Real, int R>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(ProtOcc<Real>::waves)))
void optimize_prot_kernel(OptArgs a) {
  extern __shared__ __align__(16) char smem[];
  (void)prot_setup<Real, R>(smem, a.model, a.N);
  if (wave_id() != 0) {
    prot_workers<Real, R>(smem, a);
    return;
  }
  ProtMaster<Real, R> m;
  m.init(smem, a.model, a.cap_pairs);
  optimize_loop(a, m);
}

template <int R, int MAXBD, int SPL, bool CAT, bool STRICT>
__global__ __launch_bounds__(MAXBD) __attribute__((amdgpu_waves_per_eu(MWOcc<double, SPL>::waves)))
void marginal_mw_kernel(OptArgs a) {
  extern __shared__ __align__(16) char smem[];
  const int cap = mw_cap<SPL>(a.cap_sites);
  const MWLayout<double, R, STRICT> lay = mw_layout<double, R, STRICT>(a.model, cap, SPL == 0);
  mw_copy_model(reinterpret_cast<MWShared *>(smem), a.model);
  __syncthreads();
  if (wave_id() != 0) {
    char *wl = smem + lay.align(lay.bytes());
    mw_workers<double, R, SPL, CAT, STRICT>(smem, lay, wl, a.model, a.N, a.codes, a.darr, a.parr, a.darr_e, a.parr_e,
                                    a.site_cat, a.scratch, a.cap_sites);
    return;
  }
  MWMaster<double, R, SPL, CAT, STRICT> m;
  m.init(smem, a.model, cap);
  marginal_loop(a, m);
}

template <int R>
__global__ __launch_bounds__(512) void marginal_prot_kernel(OptArgs a) {
  extern __shared__ __align__(16) char smem[];
  prot_copy_model(reinterpret_cast<ProtShared *>(smem), a.model);
  __syncthreads();
  if (wave_id() != 0) {
    double *scr = reinterpret_cast<double *>(a.scratch) + ProtLayout::scratch_reals(a.cap_pairs) * blockIdx.x;
    prot_worker<double, R>(smem, scr, a.cap_pairs, a.N, a.codes, reinterpret_cast<const double *>(a.darr),
                           reinterpret_cast<const double *>(a.parr), a.darr_e, a.parr_e, a.site_cat);
    return;
  }
  ProtMaster<double, R> m;
  m.init(smem, a.model, a.cap_pairs);
  marginal_loop(a, m);
}

// ---------------------------------------------------------------------------
// Keep (core.ml:336-351): stable sort of the pitched results by likelihood,
// descending, keeping i < keep_at_most with ll >= best + log keep_factor.
// One thread per query; only the first keep_at_most ranks are needed.
// ---------------------------------------------------------------------------
// n ints <- 0: work-queue counters, cleared by a kernel so that a captured
// graph holds no memset node
__global__ void zero_ints_kernel(int *p, int n) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) p[t] = 0;
}

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
  for (int i = 0; i < a.keep && i < n; i++) {
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
  // -p: posterior probabilities = ll_normalized_prob of the marginal
  // likelihoods in kept order (core.ml:393-401)
  double post[64];
  for (int i = 0; i < n; i++) {
    double tot = 0.;
    if (a.calc_pp)
      for (int j = 0; j < n; j++) tot = tot + exp(k[j].marg - k[i].marg);
    post[i] = a.calc_pp ? 1. / tot : __builtin_nan("");
  }
  pp_placement *out = a.out + (size_t)q * a.K;
  for (int i = 0; i < n; i++) {
    const Res &r = k[order[i]];
    out[i].loc = r.loc;
    out[i].ml_ratio = ratio[order[i]];
    out[i].log_like = r.ll;
    out[i].dist_bl = r.dist;
    out[i].pend_bl = r.pend;
    out[i].post_prob = post[order[i]];
    out[i].marginal_prob = a.calc_pp ? r.marg : __builtin_nan("");
  }
  a.n_out[q] = n;
}

// The character codes of queries idx[0 .. n) of a resident set, as a new
// set of n queries (dst[n][N]).
__global__ void gather_codes_kernel(const uint8_t *codes, const int *idx, int n, int N, uint8_t *dst) {
  const size_t total = (size_t)n * N;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x)
    dst[t] = codes[(size_t)idx[t / N] * N + t % N];
}

// The results of a pass over such a set back into the outputs of the pass
// over the full set: query j of the set is query idx[j] there.
__global__ void scatter_results_kernel(const pp_placement *src, const int *src_n, const int *src_st,
                                       const int *idx, int n, int K, pp_placement *out, int *n_out,
                                       int *status) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int q = idx[j];
  for (int k = 0; k < K; k++) out[(size_t)q * K + k] = src[(size_t)j * K + k];
  n_out[q] = src_n[j];
  status[q] = src_st[j];
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

// One lane per (row, rate): a GLV row [R][S] is held by R consecutive lanes,
// S values each.  R divides 64, so a row never straddles a wave, and every
// launch covers whole rows.

// Glv.perhaps_pull_exponent (gmix_model.ml:163-187) on one row: the frexp
// exponent of each rate's largest state, maximised over the row's R lanes.
template <int R, int S>
__device__ __forceinline__ void pull(double *v, int &e, int min_twoexp) {
  double mx = v[0];
#pragma unroll
  for (int k = 1; k < S; k++)
    if (v[k] > mx) mx = v[k];
  int te;
  (void)frexp(mx, &te);
#pragma unroll
  for (int o = 1; o < R; o <<= 1) te = max(te, __shfl_xor(te, o, 64));
  if (te < min_twoexp) {
    const double f = ldexp(0.5, -te + 1);
#pragma unroll
    for (int k = 0; k < S; k++) v[k] *= f;
    e += te;
  }
}

// dst = P_r * src for one rate (gemmish_c order: sequential sum over j)
template <int S>
__device__ __forceinline__ void evolve_rate(const double *P, const double *src, double *dst) {
#pragma unroll 4
  for (int i = 0; i < S; i++) {
    double acc = 0;
#pragma unroll
    for (int j = 0; j < S; j++) acc = acc + P[i * S + j] * src[j];
    dst[i] = acc;
  }
}

struct TreeArgs {
  int N, n_nodes;
  const int *level_nodes;   // node ids of this launch
  int n_level;
  const int *child_start, *children, *parent;
  const double *P, *Phalf;  // [G][n_sl][S][S]
  const int *site_cat;      // Price-CAT: slice of each site (nullptr: slice = rate)
  int n_sl;
  const uint8_t *leaf_codes;  // [G][N]
  const double *lv;           // [n_codes][S]
  double *darr, *parr, *evd;  // [G][N][RS]
  int *darr_e, *parr_e, *evd_e;
};

// post-order: distal = product of the children's evolved distals (or the
// leaf's LV), pulled at 2^-50, then evolv_dist = P(bl) * distal
template <int R, int S>
__global__ void distal_level_kernel(TreeArgs a) {
  constexpr int RS = R * S;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)a.n_level * a.N * R) return;
  const int r = (int)(t % R);
  const long ts = t / R;
  const int site = (int)(ts % a.N);
  const int node = a.level_nodes[ts / a.N];
  const size_t row = (size_t)node * a.N + site;
  double v[S];
  int e = 0;
  const int c0 = a.child_start[node], c1 = a.child_start[node + 1];
  if (c1 == c0) {
    const int code = a.leaf_codes[row];
#pragma unroll
    for (int k = 0; k < S; k++) v[k] = a.lv[code * S + k];
  } else {
    const size_t r0 = (size_t)a.children[c0] * a.N + site;
#pragma unroll
    for (int k = 0; k < S; k++) v[k] = a.evd[r0 * RS + r * S + k];
    e = a.evd_e[r0];
    for (int c = c0 + 1; c < c1; c++) {
      const size_t rc = (size_t)a.children[c] * a.N + site;
#pragma unroll
      for (int k = 0; k < S; k++) v[k] = v[k] * a.evd[rc * RS + r * S + k];
      e += a.evd_e[rc];
    }
    pull<R, S>(v, e, -50);
  }
#pragma unroll
  for (int k = 0; k < S; k++) a.darr[row * RS + r * S + k] = v[k];
  if (r == 0) a.darr_e[row] = e;
  double w[S];
  const int sl = a.site_cat ? a.site_cat[site] : r;
  evolve_rate<S>(a.P + ((size_t)node * a.n_sl + sl) * S * S, v, w);
#pragma unroll
  for (int k = 0; k < S; k++) a.evd[row * RS + r * S + k] = w[k];
  if (r == 0) a.evd_e[row] = e;
}

// pre-order: proximal of a child = P(bl_parent) * proximal(parent) times the
// evolved distals of its siblings in order, pulled at 2^-50
template <int R, int S>
__global__ void proximal_level_kernel(TreeArgs a) {
  constexpr int RS = R * S;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)a.n_level * a.N * R) return;
  const int r = (int)(t % R);
  const long ts = t / R;
  const int site = (int)(ts % a.N);
  const int child = a.level_nodes[ts / a.N];
  const int par = a.parent[child];
  const size_t prow = (size_t)par * a.N + site;
  double src[S], v[S];
#pragma unroll
  for (int k = 0; k < S; k++) src[k] = a.parr[prow * RS + r * S + k];
  int e = a.parr_e[prow];
  const int sl = a.site_cat ? a.site_cat[site] : r;
  evolve_rate<S>(a.P + ((size_t)par * a.n_sl + sl) * S * S, src, v);
  for (int c = a.child_start[par]; c < a.child_start[par + 1]; c++) {
    const int sib = a.children[c];
    if (sib == child) continue;
    const size_t rs = (size_t)sib * a.N + site;
#pragma unroll
    for (int k = 0; k < S; k++) v[k] = v[k] * a.evd[rs * RS + r * S + k];
    e += a.evd_e[rs];
  }
  pull<R, S>(v, e, -50);
  const size_t row = (size_t)child * a.N + site;
#pragma unroll
  for (int k = 0; k < S; k++) a.parr[row * RS + r * S + k] = v[k];
  if (r == 0) a.parr_e[row] = e;
}

__global__ void fill_unit_kernel(double *row, int *e, int N, int RS) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < (long)N * RS; t += (long)gridDim.x * blockDim.x) {
    row[t] = 1.0;
    if (t < N) e[t] = 0;
  }
}

// pull both arrays at 2^-10 (pplacer_run.ml:331-336), then the supernodes
// pi * P(bl/2) darr * P(bl/2) parr (glv_arr.ml:46-56), written site-major
// for stage 1
template <int R, int S>
__global__ void finish_kernel(TreeArgs a, const double *statd, int E, int Epad, double *sn64,
                              int *sn_e) {
  constexpr int RS = R * S;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)a.n_nodes * a.N * R) return;
  const int r = (int)(t % R);
  const long ts = t / R;
  const int site = (int)(ts % a.N);
  const int node = (int)(ts / a.N);
  const size_t row = (size_t)node * a.N + site;
  double d[S], p[S];
#pragma unroll
  for (int k = 0; k < S; k++) { d[k] = a.darr[row * RS + r * S + k]; p[k] = a.parr[row * RS + r * S + k]; }
  int ed = a.darr_e[row], ep = a.parr_e[row];
  pull<R, S>(d, ed, -10);
  pull<R, S>(p, ep, -10);
#pragma unroll
  for (int k = 0; k < S; k++) {
    a.darr[row * RS + r * S + k] = d[k];
    a.parr[row * RS + r * S + k] = p[k];
  }
  if (r == 0) {
    a.darr_e[row] = ed;
    a.parr_e[row] = ep;
  }
  if (node < E) {
    double ud[S], up[S];
    const int sl = a.site_cat ? a.site_cat[site] : r;
    const double *Ph = a.Phalf + ((size_t)node * a.n_sl + sl) * S * S;
    evolve_rate<S>(Ph, d, ud);
    evolve_rate<S>(Ph, p, up);
    const size_t srow = (size_t)site * Epad + node;
#pragma unroll
    for (int k = 0; k < S; k++) sn64[srow * RS + r * S + k] = (statd[k] * ud[k]) * up[k];
    if (r == 0) sn_e[srow] = ed + ep;
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
