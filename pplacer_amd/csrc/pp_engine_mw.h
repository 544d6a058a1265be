// Master/worker three-taxon engine (nucleotide models, S = 4).
//
// A workgroup is one master wave (wave 0) plus W worker waves.  The master
// runs pplacer's sequential control -- baseball (core.ml:283-320),
// Three_tax.optimize (three_tax.ml:70-83), Minimization.brent with the GSL
// Brent minimiser (minimization.ml:11-100) -- on wave-uniform registers, and
// builds each probe's transition matrices and evolved pendant rows with its 64
// lanes (Diagd.multi_exp, diagd.ml:79-100; gemmish_c, linear_c.c:29-90).  Lane
// j of worker wave w holds, in registers, the distal and proximal GLV rows of
// the current edge (Glv.mask_into, gmix_model.ml:121-141) at its masked
// sites; the cached (pi*prox_ev)*dist_ev products of ten_log_like3_c
// (linear_c.c:367-424) are kept in LDS.  Each command is one LDS word and two
// barriers:
//
//   master: [tables, command] -> barrier A -> .............................. -> barrier B -> [sum]
//   worker: ..................   barrier A -> [site work, wave partial sum] -> barrier B
//
// With SPL sites per lane, worker lane j of wave w holds masked sites
// 64*(w-1)+j + k*64*W for k < SPL.  Queries longer than the registers and LDS
// of a workgroup hold take the streaming form (SPL = 0): a worker lane walks
// its sites 64*(w-1)+j + k*64*W for every k, re-reads the rows from the
// resident GLV arrays at each distal probe and keeps the cached products in a
// per-workgroup global scratch, so any masked length fits.
//
// Control state and site data live in different waves, so the register
// allocator sizes the kernel for max(control, data) rather than their sum;
// that sets how many queries a CU holds at once.
//
// Arithmetic order is the C stubs': each evolved entry is a sequential sum
// over j of P[i][j]*src[j]; each site's util the sequential (rate, state) sum
// of ((pi*prox)*dist)*pend; P entries the sequential sum over k of
// (exp(rate*bl*l_k)*x[i][k])*xit[j][k].  The FP64 form rounds every product
// and sum separately (-ffp-contract=off) and sums the per-site logs as the
// stub does, sequentially in masked-site order (linear_c.c:416-419): workers
// leave each site's log in LDS and, behind one more barrier, the first worker
// wave reproduces that sum bit for bit (wave_sequential_sum).  The FP32 form
// uses FMAs and sums the logs (in FP64) per wave, then in wave order.
#pragma once

#include <climits>
#include <type_traits>

#include "pp_engine.h"
#include "pp_engine_ctl.h"
#include "pp_exact_sum.h"

namespace pp {

// CMD_PROBE_DIST: a probe whose proximal/distal lengths changed (re-evolve)
enum { CMD_EXIT = 0, CMD_QUERY = 1, CMD_GATHER = 2, CMD_PROBE = 3, CMD_PROBE_DIST = 4 };

struct MWShared {
  double x[16], l[4], xit[16], statd[4], rates[kMaxRates];
  double lv[kMaxCodes * 4];
  int op, n_slots, M, q, loc;
  // character code -> pendant row of this query; the one-hot characters
  // (codes 0-3) always use rows 0-3
  int code_slot[kMaxCodes];
  int slot_code[kMaxCodes];
  double part[16];           // per-worker-wave sums of site logs
  long long ipart[16];       // per-worker-wave sums of site exponents
};

// LDS of one workgroup: shared header, pendant rows [n_tsl][kMaxCodes][R*4],
// P(prox) and P(dist) [n_sl][4][4], exponential staging [3][n_sl][4], then
// per masked site (cap = the longest masked length the launch takes): the
// site's log of a probe and the sequential sum's scratch (FP64 form), its
// column (the masked-site list) and in the streaming form its pendant row;
// the FP64 form also holds the sequential sum's event list.  Gamma model:
// n_tsl = 1, n_sl = R (rate r uses slice r); Price-CAT: R = 1 and
// n_tsl = n_sl = n_cats (a site uses its category's).
template <typename Real, int R, bool STRICT>
struct MWLayout {
  static constexpr int RS = R * 4;
  static constexpr size_t head = (sizeof(MWShared) + 15) & ~size_t(15);
  static constexpr bool strict = STRICT;  // site logs summed in site order
  int n_tsl, n_sl, cap;
  bool stream;
  __host__ __device__ MWLayout(int n_tsl_, int n_sl_, int cap_, bool stream_)
      : n_tsl(n_tsl_), n_sl(n_sl_), cap(cap_), stream(stream_) {}
  __host__ __device__ size_t tbytes() const { return sizeof(Real) * (size_t)n_tsl * kMaxCodes * RS; }
  __host__ __device__ size_t pbytes() const { return sizeof(Real) * (size_t)n_sl * 16; }
  __host__ __device__ size_t ebytes() const { return sizeof(double) * (size_t)3 * n_sl * 4; }
  __host__ __device__ size_t off_slog() const { return head + tbytes() + 2 * pbytes() + ebytes(); }
  __host__ __device__ size_t off_sacc() const { return off_slog() + (strict ? sizeof(double) * (size_t)cap : 0); }
  __host__ __device__ size_t off_sev() const { return off_sacc() + (strict ? sizeof(double) * (size_t)cap : 0); }
  __host__ __device__ size_t off_sidx() const {
    return off_sev() + (strict ? sizeof(double) * 2 * (size_t)kSeqSumEvents : 0);
  }
  __host__ __device__ size_t bytes() const { return off_sidx() + sizeof(int) * (size_t)cap * (stream ? 2 : 1); }
  __device__ Real *T(char *s) const { return reinterpret_cast<Real *>(s + head); }
  __device__ Real *Pp(char *s) const { return reinterpret_cast<Real *>(s + head + tbytes()); }
  __device__ Real *Pd(char *s) const { return reinterpret_cast<Real *>(s + head + tbytes() + pbytes()); }
  __device__ double *E(char *s) const {
    return reinterpret_cast<double *>(s + head + tbytes() + 2 * pbytes());
  }
  __device__ double *slog(char *s) const { return reinterpret_cast<double *>(s + off_slog()); }
  __device__ double *sacc(char *s) const { return reinterpret_cast<double *>(s + off_sacc()); }
  __device__ double *sev(char *s) const { return reinterpret_cast<double *>(s + off_sev()); }
  __device__ int *sidx(char *s) const { return reinterpret_cast<int *>(s + off_sidx()); }
  __device__ int *sslot(char *s) const { return reinterpret_cast<int *>(s + off_sidx()) + cap; }
  static __host__ __device__ size_t align(size_t off) { return (off + 15) & ~size_t(15); }
};

template <typename Real, int R, bool STRICT>
__device__ __forceinline__ MWLayout<Real, R, STRICT> mw_layout(const DevModel *dm, int cap, bool stream) {
  return MWLayout<Real, R, STRICT>(dm->cat ? dm->n_sl : 1, dm->n_sl, cap, stream);
}

// Wave-uniform branch on the wave index (wave 0 = master).
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

__device__ __forceinline__ void mw_copy_model(MWShared *sh, const DevModel *dm) {
  for (int t = threadIdx.x; t < 16; t += blockDim.x) { sh->x[t] = dm->x[t]; sh->xit[t] = dm->xit[t]; }
  for (int t = threadIdx.x; t < 4; t += blockDim.x) { sh->l[t] = dm->l[t]; sh->statd[t] = dm->statd[t]; }
  for (int t = threadIdx.x; t < kMaxRates; t += blockDim.x) sh->rates[t] = dm->rates[t];
  for (int t = threadIdx.x; t < dm->n_codes * 4; t += blockDim.x) sh->lv[t] = dm->lv[t];
}

// ---------------------------------------------------------------------------
// worker waves
// ---------------------------------------------------------------------------

// The cached products of masked site i live in LDS in groups of four sites:
// byte (i/4)*kGroupBytes + t*4*sizeof(Real) + (i%4)*sizeof(Real).  Each lane's
// sixteen products are immediate offsets from one address, and the 32-byte pad
// per group spreads a wave's accesses over all banks.
template <typename Real, int RS>
struct MWProducts {
  static constexpr int kGroupBytes = RS * 4 * (int)sizeof(Real) + 32;
  static __host__ __device__ size_t bytes(int sites) { return (size_t)((sites + 3) / 4) * kGroupBytes; }
  static __device__ __forceinline__ int base(int i) { return (i >> 2) * kGroupBytes + (i & 3) * (int)sizeof(Real); }
  static __device__ __forceinline__ Real &at(char *wl, int base, int t) {
    return *reinterpret_cast<Real *>(wl + base + t * 4 * (int)sizeof(Real));
  }
};

// One masked site of a CMD_PROBE_DIST: evolve the proximal and distal rows
// by P(prox) and P(dist) (sequential sums over j), form each (rate, state)
// product ((pi*prox_ev)*dist_ev), hand it to keep(t, value) and accumulate
// util = sum over (rate, state) of product*pend in that order.  pp / pd are
// the site's P slices (for Price-CAT its category's; else rate r's slice is
// pp + 16 r), z its evolved pendant row.
template <typename Real, int R, bool CAT, class Keep>
__device__ __forceinline__ Real site_dist_util(const Real *pp, const Real *pd, const Real *z,
                                               const double *statd, const Real (&pO)[R * 4],
                                               const Real (&dO)[R * 4], Keep &&keep) {
  using Ops = RealOps<Real>;
  Real util = 0;
#pragma unroll
  for (int r = 0; r < R; r++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const Real *ppr = pp + (CAT ? 0 : r * 16) + i * 4, *pdr = pd + (CAT ? 0 : r * 16) + i * 4;
      const Real pi = (Real)statd[i];
      const int t = r * 4 + i;
      Real xv = ppr[0] * pO[r * 4], yv = pdr[0] * dO[r * 4];
#pragma unroll
      for (int j = 1; j < 4; j++) {
        xv = Ops::madd(xv, ppr[j], pO[r * 4 + j]);
        yv = Ops::madd(yv, pdr[j], dO[r * 4 + j]);
      }
      const Real w = (pi * xv) * yv;
      keep(t, w);
      // sums start from their first product: 0 + p == p up to the sign of a
      // zero, which no later operation can observe
      util = t == 0 ? w * z[t] : Ops::madd(util, w, z[t]);
    }
  return util;
}

// CAT: Price-CAT, R == 1, site k uses P slice and pendant table sl[k] = its
// category (site_cat); otherwise slice r for rate r and pendant table 0.
// FP64 (strict): each active site's log goes to slog[site], and after a
// barrier the first worker wave adds them as the stub does, in site order
// with one rounding per term (wave_sequential_sum); FP32: the lane's logs
// are summed and reduced per wave.
template <typename Real, int R, int SPL, bool CAT, bool STRICT>
__device__ __forceinline__ void mw_worker(char *smem, const MWLayout<Real, R, STRICT> &lay, char *wl, int N,
                                          const uint8_t *codes, const Real *darr, const Real *parr,
                                          const int *darr_e, const int *parr_e, const int *site_cat) {
  constexpr int RS = R * 4;
  constexpr bool strict = STRICT;
  static_assert(!CAT || R == 1, "Price-CAT GLVs have no rate axis");
  static_assert(SPL >= 1, "register-resident form");
  using W = MWProducts<Real, RS>;
  using Ops = RealOps<Real>;
  MWShared *sh = reinterpret_cast<MWShared *>(smem);
  const Real *sT = lay.T(smem), *sPp = lay.Pp(smem), *sPd = lay.Pd(smem);
  const int *sidx = lay.sidx(smem);
  double *slog = lay.slog(smem);
  const int lane = threadIdx.x & 63;
  const int wv = wave_id() - 1;
  const int stride = (int)blockDim.x - 64;  // worker lanes
  const int my = wv * 64 + lane;
  int site[SPL], slot[SPL], wb[SPL], sl[SPL];
  bool active[SPL];
#pragma unroll
  for (int k = 0; k < SPL; k++) wb[k] = W::base(my + k * stride);
  // The command stream is QUERY (GATHER PROBE*)* ... EXIT; the nested loops
  // follow it, so each edge's rows are defined once per GATHER and stay in
  // place through that edge's probes.
  __syncthreads();  // A
  int op = sh->op;
  while (op == CMD_QUERY) {
#pragma unroll
    for (int k = 0; k < SPL; k++) {
      const int i = my + k * stride;
      active[k] = i < sh->M;
      site[k] = active[k] ? sidx[i] : 0;
      slot[k] = active[k] ? sh->code_slot[codes[(size_t)sh->q * N + site[k]]] : 0;
      sl[k] = CAT ? site_cat[site[k]] : 0;
      // pendant row of the site: table sl, character slot
      slot[k] += sl[k] * kMaxCodes;
    }
    __syncthreads();  // B
    __syncthreads();  // A
    op = sh->op;
    while (op == CMD_GATHER) {
      // Inactive slots load row 0 (any valid row).
      Real dO[SPL][RS], pO[SPL][RS];
      long long es = 0;
#pragma unroll
      for (int k = 0; k < SPL; k++) {
        const size_t row = active[k] ? (size_t)sh->loc * N + site[k] : 0;
        const Real *d = darr + row * RS;
        const Real *p = parr + row * RS;
#pragma unroll
        for (int t = 0; t < RS; t++) { dO[k][t] = d[t]; pO[k][t] = p[t]; }
        if (active[k]) es += (long long)darr_e[row] + (long long)parr_e[row];
      }
      es = wave_sum_ll(es);
      if (lane == 0) sh->ipart[wv] = es;
      __syncthreads();  // B
      __syncthreads();  // A
      op = sh->op;
      while (op == CMD_PROBE || op == CMD_PROBE_DIST) {
        Real util[SPL];
        if (op == CMD_PROBE_DIST) {
#pragma unroll
          for (int k = 0; k < SPL; k++) {
            const int o = CAT ? sl[k] * 16 : 0;
            util[k] = site_dist_util<Real, R, CAT>(sPp + o, sPd + o, sT + slot[k] * RS, sh->statd, pO[k], dO[k],
                                                   [&](int t, Real w) { W::at(wl, wb[k], t) = w; });
          }
        } else {
#pragma unroll
          for (int t = 0; t < RS; t++)
#pragma unroll
            for (int k = 0; k < SPL; k++) {
              const Real w = W::at(wl, wb[k], t), z = sT[slot[k] * RS + t];
              util[k] = t == 0 ? w * z : Ops::madd(util[k], w, z);
            }
        }
        if constexpr (strict) {
#pragma unroll
          for (int k = 0; k < SPL; k++)
            if (active[k]) slog[my + k * stride] = Ops::logs(util[k]);
          __syncthreads();  // every site's log is in slog
          if (wv == 0) {
            const double t = wave_sequential_sum(slog, sh->M, lay.sacc(smem), lay.sev(smem), nullptr, kSeqSumWaveMin);
            if (lane == 0) sh->part[0] = t;
          }
        } else {
          double lsum = 0.0;
#pragma unroll
          for (int k = 0; k < SPL; k++)
            if (active[k]) lsum += Ops::logd(util[k]);
          lsum = wave_sum(lsum);
          if (lane == 0) sh->part[wv] = lsum;
        }
        __syncthreads();  // B
        __syncthreads();  // A
        op = sh->op;
      }
    }
  }
  // op == CMD_EXIT
}

// Streaming form of the worker for masked lengths beyond the register-resident
// form: lane j of worker wave w walks sites i = 64 (w-1) + j + k W64 for every
// k while i < M.  The distal and proximal rows are read from darr / parr at
// each distal probe (one 16-value row each), and the products cached for the
// pendant-only probes live in prod[t * cap + i] (per-workgroup global scratch,
// consecutive sites in consecutive addresses).  Same arithmetic as mw_worker.
template <typename Real, int R, bool CAT, bool STRICT>
__device__ __forceinline__ void mw_worker_stream(char *smem, const MWLayout<Real, R, STRICT> &lay, Real *prod, int cap,
                                                 int N, const uint8_t *codes, const Real *darr,
                                                 const Real *parr, const int *darr_e, const int *parr_e,
                                                 const int *site_cat) {
  constexpr int RS = R * 4;
  constexpr bool strict = STRICT;
  static_assert(!CAT || R == 1, "Price-CAT GLVs have no rate axis");
  using Ops = RealOps<Real>;
  MWShared *sh = reinterpret_cast<MWShared *>(smem);
  const Real *sT = lay.T(smem), *sPp = lay.Pp(smem), *sPd = lay.Pd(smem);
  const int *sidx = lay.sidx(smem);
  int *sslot = lay.sslot(smem);
  double *slog = lay.slog(smem);
  const int lane = threadIdx.x & 63;
  const int wv = wave_id() - 1;
  const int stride = (int)blockDim.x - 64;
  const int my = wv * 64 + lane;
  __syncthreads();  // A
  int op = sh->op;
  while (op == CMD_QUERY) {
    const int M = sh->M, q = sh->q;
    for (int i = my; i < M; i += stride) {
      const int site = sidx[i];
      const int sl = CAT ? site_cat[site] : 0;
      sslot[i] = sh->code_slot[codes[(size_t)q * N + site]] + sl * kMaxCodes;
    }
    __syncthreads();  // B
    __syncthreads();  // A
    op = sh->op;
    while (op == CMD_GATHER) {
      const size_t rowbase = (size_t)sh->loc * N;
      long long es = 0;
      for (int i = my; i < M; i += stride) {
        const size_t row = rowbase + sidx[i];
        es += (long long)darr_e[row] + (long long)parr_e[row];
      }
      es = wave_sum_ll(es);
      if (lane == 0) sh->ipart[wv] = es;
      __syncthreads();  // B
      __syncthreads();  // A
      op = sh->op;
      while (op == CMD_PROBE || op == CMD_PROBE_DIST) {
        double lsum = 0.0;
        for (int i = my; i < M; i += stride) {
          const int slot = sslot[i];
          const Real *z = sT + slot * RS;
          Real util;
          if (op == CMD_PROBE_DIST) {
            const int site = sidx[i];
            const size_t row = rowbase + site;
            Real dO[RS], pO[RS];
#pragma unroll
            for (int t = 0; t < RS; t++) { dO[t] = darr[row * RS + t]; pO[t] = parr[row * RS + t]; }
            const int o = CAT ? (slot / kMaxCodes) * 16 : 0;
            util = site_dist_util<Real, R, CAT>(sPp + o, sPd + o, z, sh->statd, pO, dO,
                                                [&](int t, Real w) { prod[(size_t)t * cap + i] = w; });
          } else {
#pragma unroll
            for (int t = 0; t < RS; t++) {
              const Real w = prod[(size_t)t * cap + i];
              util = t == 0 ? w * z[t] : Ops::madd(util, w, z[t]);
            }
          }
          if constexpr (strict) slog[i] = Ops::logs(util);
          else lsum += Ops::logd(util);
        }
        if constexpr (strict) {
          __syncthreads();  // every site's log is in slog
          if (wv == 0) {
            const double t = wave_sequential_sum(slog, M, lay.sacc(smem), lay.sev(smem), nullptr, kSeqSumWaveMin);
            if (lane == 0) sh->part[0] = t;
          }
        } else {
          lsum = wave_sum(lsum);
          if (lane == 0) sh->part[wv] = lsum;
        }
        __syncthreads();  // B
        __syncthreads();  // A
        op = sh->op;
      }
    }
  }
  // op == CMD_EXIT
}

// Orders one wave's LDS writes before its later LDS reads of other lanes'
// entries (LDS executes a wave's operations in order).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// master wave
// ---------------------------------------------------------------------------
template <typename Real, int R, int SPL, bool CAT, bool STRICT>
struct MWMaster : ThreeTaxControl<MWMaster<Real, R, SPL, CAT, STRICT>> {
  using Ctl = ThreeTaxControl<MWMaster<Real, R, SPL, CAT, STRICT>>;
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
  using L = MWLayout<Real, R, STRICT>;
  static constexpr int RS = R * 4;
  MWShared *sh;
  Real *sT, *sPp, *sPd;
  double *sE;
  int *sidx;
  int lane, n_workers, cap_sites, n_codes, n_inf, n_sl;
  double log_R;

  // cap: the longest masked length the launch takes (the layout's capacity)
  __device__ __forceinline__ void init(char *smem, const DevModel *dm, int cap) {
    const L lay = mw_layout<Real, R, STRICT>(dm, cap, SPL == 0);
    sh = reinterpret_cast<MWShared *>(smem);
    sT = lay.T(smem);
    sPp = lay.Pp(smem);
    sPd = lay.Pd(smem);
    sE = lay.E(smem);
    sidx = lay.sidx(smem);
    n_sl = dm->n_sl;
    lane = threadIdx.x & 63;
    n_workers = (blockDim.x >> 6) - 1;
    cap_sites = cap;
    n_codes = dm->n_codes;
    n_inf = dm->n_inf;
    log_R = dm->log_R;
    c_pend = c_dist = c_prox = __builtin_nan("");
  }

  // A probe of the FP64 form has one more barrier, after which worker lane 0
  // adds the site logs in order.
  __device__ __forceinline__ void command(int op) {
    if (lane == 0) sh->op = op;
    __syncthreads();  // A
    if (op == CMD_EXIT) return;
    if (L::strict && (op == CMD_PROBE || op == CMD_PROBE_DIST)) __syncthreads();  // logs written
    __syncthreads();  // B
  }

  // mask_reindex of the query's informative sites (core.ml:25-28,146-156),
  // the set of characters present (their evolved rows are all a probe
  // needs), then CMD_QUERY.  Returns M.
  __device__ __forceinline__ int load_query(const uint8_t *codes, int N, int q) {
    const int chunk = (N + 63) / 64;
    const int lo = lane * chunk, hi = min(N, lo + chunk);
    int cnt = 0;
    unsigned present = 0;
    for (int k = lo; k < hi; k++) {
      const int c = codes[k];
      if (c < n_inf) { cnt++; present |= 1u << c; }
    }
    int incl = cnt;
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
      const int o = __shfl_up(incl, m, 64);
      if (lane >= m) incl += o;
    }
    const int total = __builtin_amdgcn_readfirstlane(__shfl(incl, 63, 64));
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) present |= __shfl_xor(present, m, 64);
    int pos = incl - cnt;
    for (int k = lo; k < hi; k++)
      if (codes[k] < n_inf) {
        if (pos < cap_sites) sidx[pos] = k;  // a longer query is refused below
        pos++;
      }
    if (lane == 0) {
      int ns = 4;
      for (int c = 0; c < n_codes; c++) {
        if (c < 4) { sh->code_slot[c] = c; sh->slot_code[c] = c; continue; }
        sh->code_slot[c] = (present >> c) & 1u ? ns : 0;
        if ((present >> c) & 1u) sh->slot_code[ns++] = c;
      }
      sh->n_slots = ns;
      sh->M = total;
      sh->q = q;
    }
    M = total;
    c_pend = __builtin_nan("");  // the pendant rows are per query
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

  // Diagd.multi_exp for the changed branch lengths and the evolved pendant
  // rows of the characters present, with the master's 64 lanes.
  __device__ __forceinline__ void tables(bool new_pend, bool new_dist) {
    if constexpr (CAT) {
      cat_tables(new_pend, new_dist);
      return;
    }
    // The exponentials reach the lanes forming P entries through the LDS
    // staging area sE (one write, a wave fence, independent reads).
    if (new_dist) {
      // exp(rate_r * bl * l_k) for prox (u < 4R) and dist (4R <= u < 8R)
      if (lane < 8 * R) {
        const int which = lane / (4 * R), r = (lane / 4) % R, k = lane % 4;
        sE[lane] = texp((sh->rates[r] * (which ? dist_bl : prox_bl)) * sh->l[k]);
      }
      wave_lds_sync();
#pragma unroll
      for (int base = 0; base < 2 * R * 16; base += 64) {
        const int e = base + lane;
        const int which = e / (R * 16), rem = e % (R * 16);
        const int r = rem / 16, i = (rem / 4) % 4, j = rem % 4;
        const double *ex = sE + (which * R + r) * 4;
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++) acc = acc + (ex[k] * sh->x[i * 4 + k]) * sh->xit[j * 4 + k];
        if (e < 2 * R * 16) (which ? sPd : sPp)[rem] = (Real)acc;
      }
    }
    if (new_pend) {
      double *sEp = sE + 8 * R;
      if (lane < 4 * R) {
        const int r = lane / 4, k = lane % 4;
        sEp[lane] = texp((sh->rates[r] * pend_bl) * sh->l[k]);
      }
      wave_lds_sync();
      // P_pend[r][i][j] at lane r*16 + i*4 + j (R*16 <= 64)
      double pij = 0.0;
      {
        const int r = (lane / 16) % R, i = (lane / 4) % 4, j = lane % 4;
        const double *ex = sEp + r * 4;
#pragma unroll
        for (int k = 0; k < 4; k++) pij = pij + (ex[k] * sh->x[i * 4 + k]) * sh->xit[j * 4 + k];
      }
      // One-hot characters: their evolved rows are columns of P (every other
      // product in gemmish_c is an exact zero), row j entry (r, i) = P[r][i][j].
      if (lane < R * 16) {
        const int r = lane / 16, i = (lane / 4) % 4, j = lane % 4;
        sT[j * RS + r * 4 + i] = (Real)pij;
      }
      // Ambiguity codes present in the query: rows 4.. by the stub's sum.
      const int n_t = (sh->n_slots - 4) * RS;
      for (int base = 0; base < n_t; base += 64) {
        const int t = base + lane;
        const int slot = 4 + t / RS, r = (t / 4) % R, i = t % 4;
        const int c = sh->slot_code[t < n_t ? slot : 0];
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const double p = __shfl(pij, r * 16 + i * 4 + j, 64);
          acc = acc + p * sh->lv[c * 4 + j];
        }
        if (t < n_t) sT[slot * RS + r * 4 + i] = (Real)acc;
      }
    }
  }

  // Price-CAT tables: the same sums for every category slice, staged through
  // LDS (more entries than lanes): exponentials, then P(prox), P(dist) and
  // P(pend) entries; the one-hot pendant rows are P(pend) columns and the
  // ambiguity rows are formed from them.
  // The exponentials of P(t): correctly rounded where the site logs are
  // summed in site order (the FP64 form), as the stubs' libm returns them;
  // the library's otherwise.
  static __device__ __forceinline__ double texp(double v) {
    if constexpr (STRICT) return cr::exp_call(v);
    else return exp(v);
  }

  __device__ __forceinline__ void cat_tables(bool new_pend, bool new_dist) {
    const int nd = new_dist ? 2 * n_sl * 4 : 0, np = new_pend ? n_sl * 4 : 0;
    // sE[0 .. 2 n_sl 4): prox then dist; sE[2 n_sl 4 ..): pend
    for (int t = lane; t < nd + np; t += 64) {
      const bool pend = t >= nd;
      const int u = pend ? t - nd : t;
      const int which = pend ? 2 : u / (n_sl * 4), rem = u % (n_sl * 4), s = rem / 4, k = rem % 4;
      const double bl = which == 0 ? prox_bl : which == 1 ? dist_bl : pend_bl;
      sE[which * n_sl * 4 + rem] = texp((sh->rates[s] * bl) * sh->l[k]);
    }
    wave_lds_sync();
    const int ndp = new_dist ? 2 * n_sl * 16 : 0, npp = new_pend ? n_sl * 16 : 0;
    for (int t = lane; t < ndp + npp; t += 64) {
      const bool pend = t >= ndp;
      const int u = pend ? t - ndp : t;
      const int which = pend ? 2 : u / (n_sl * 16), rem = u % (n_sl * 16);
      const int s = rem / 16, i = (rem / 4) % 4, j = rem % 4;
      const double *ex = sE + (which * n_sl + s) * 4;
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 4; k++) acc = acc + (ex[k] * sh->x[i * 4 + k]) * sh->xit[j * 4 + k];
      if (which == 2) sT[(s * kMaxCodes + j) * 4 + i] = (Real)acc;  // one-hot row j, entry i
      else (which ? sPd : sPp)[rem] = (Real)acc;
    }
    if (new_pend && sh->n_slots > 4) {
      wave_lds_sync();
      const int n_amb = sh->n_slots - 4;
      for (int t = lane; t < n_sl * n_amb * 4; t += 64) {
        const int s = t / (n_amb * 4), slot = 4 + (t / 4) % n_amb, i = t % 4;
        const int c = sh->slot_code[slot];
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 4; j++) acc = acc + (double)sT[(s * kMaxCodes + j) * 4 + i] * sh->lv[c * 4 + j];
        sT[(s * kMaxCodes + slot) * 4 + i] = (Real)acc;
      }
    }
  }

  // One likelihood probe (Model.log_like3 at the current three-taxon state).
  __device__ __forceinline__ double log_like() {
    const bool new_pend = !(pend_bl == c_pend);
    const bool new_dist = !(dist_bl == c_dist && prox_bl == c_prox);
    tables(new_pend, new_dist);
    command(new_dist ? CMD_PROBE_DIST : CMD_PROBE);
    double tot;
    if constexpr (L::strict) {
      tot = sh->part[0];
    } else {
      // partials of waves >= n_workers read as 0.0; the sum is in wave order
      double pv[16];
#pragma unroll
      for (int w = 0; w < 16; w++) pv[w] = w < n_workers ? sh->part[w] : 0.0;
      tot = pv[0];
#pragma unroll
      for (int w = 1; w < 16; w++) tot += pv[w];
    }
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
