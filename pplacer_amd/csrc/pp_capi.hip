// C ABI of libpplacer_hip.so (include/pplacer_hip.h): device context, data
// layout in HBM, and the launch sequence of the batched placement path.
//
// HBM layout of a resident tree (G nodes, N sites, E = G-1 edges, R rates):
//   darr, parr : [G][N][R*4] FP64 and FP32 copies, exponents int32 [G][N]
//                (a masked site of one edge is one contiguous row: stage 2
//                 gathers rows of the query's informative sites)
//   snodes     : [N][Epad][R*4] FP64 and FP32 (site-major: stage 1 lanes =
//                 edges read consecutive rows), exponents int32 [N][Epad]
// Queries are stored as 1-byte character codes [Q][N], sorted by first
// informative column; results are un-permuted on download.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <set>
#include <string>
#include <vector>

#include "pp_kernels.hip"
#include "pp_tree.h"

using namespace pp;

#define HIPCHECK(expr)                                                              \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) {                                                         \
      set_error(std::string(#expr) + " failed: " + hipGetErrorString(e_));          \
      return PP_EDEVICE;                                                            \
    }                                                                               \
  } while (0)

#define CHECK(cond, code, msg) \
  do {                         \
    if (!(cond)) {             \
      set_error(msg);          \
      return code;             \
    }                          \
  } while (0)

namespace {

constexpr int kTQ = 16;           // queries per stage-1 tile
constexpr int kChunk = 65536;     // queries per device batch
constexpr int kLongBucket = 30;   // stage-2 queries beyond the register-resident classes
constexpr int kBuckets = 31;      // stage-2 masked-length classes of 64 sites, then kLongBucket
constexpr int kBucketStreams = PP_MAX_BUCKET_STREAMS;  // streams for a stage-2 phase's buckets
static_assert(kChunk <= PP_MAX_CHUNK, "default chunk out of range");
// Default streams a 4-state stage-2 phase spreads its buckets over (0: the
// main stream only), by form [FP64, FP32].  Measured at 100,000 queries,
// 2,000 taxa.
constexpr int kPitchStreams[2] = {0, 8};
constexpr int kRefineStreams[2] = {4, 4};
constexpr int kProtBlock = 512;   // lanes of a 20-state stage-2 workgroup (master + workers)

struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

template <class T>
int ensure(DBuf &b, size_t count, T **out) {
  size_t need = std::max<size_t>(count * sizeof(T), 64);
  if (b.bytes < need) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    if (hipMalloc(&b.p, need) != hipSuccess) {
      set_error("hipMalloc of " + std::to_string(need) + " bytes failed");
      return PP_ENOMEM;
    }
    b.bytes = need;
  }
  *out = reinterpret_cast<T *>(b.p);
  return PP_OK;
}

#define ENSURE(buf, n, ptr)                  \
  do {                                       \
    int st_ = ensure(buf, (size_t)(n), ptr); \
    if (st_) return st_;                     \
  } while (0)


}  // namespace

struct pp_ctx {
  int device = 0;
  int n_cu = 256;
  size_t max_lds = 65536;  // bytes of LDS one workgroup may use
  hipStream_t stream = nullptr;       // main stream: stage 2 and everything else
  hipStream_t s1 = nullptr;           // stage 1 of the next chunk
  hipStream_t bs[kBucketStreams] = {};  // concurrent stage-2 buckets
  hipStream_t cur = nullptr;          // the stream the launch helpers use
  pp_launch_config launch{{kPitchStreams[0], kPitchStreams[1]}, {kRefineStreams[0], kRefineStreams[1]}, kChunk, 0};
  // graph mode: captured chunk loops by plan key (most recent last; a pass
  // and the FP64 re-placement of its FP32 failures keep one each)
  struct CapturedPass {
    uint64_t key;
    hipGraphExec_t exec;
    long long launches, pitch_launches;
  };
  std::vector<CapturedPass> graphs;
  HostModel model;
  DBuf b_model;
  DevModel *d_model = nullptr;
  // tree
  bool has_tree = false;
  int G = 0, N = 0, E = 0, Epad = 0;
  std::vector<double> bl;
  DBuf b_bl, b_d64, b_p64, b_d32, b_p32, b_de, b_pe, b_s64, b_se;
  DBuf b_de32, b_pe32;  // exponents of the FP32 copies
  DBuf b_lt;            // stage-1 log-util table [N][n_inf+1][Epad]
  DBuf b_scratch;       // 20-state stage-2 pair scratch
  DBuf b_sscr;          // 4-state streaming-form product scratch [blocks][R*4][long_cap]
  int sscr_blocks = 0;  // workgroups b_sscr holds
  DBuf b_cat;           // Price-CAT: category of each pre-masked site [N]
  // -p prior (Core.prior): kind, flat mean or per-edge means; device [G] means
  int prior_kind = PP_PRIOR_FLAT_EXP;
  bool prior_set = false;
  std::vector<double> prior_means;
  DBuf b_prior;
  std::vector<double> h_ttab;
  // resident queries (sorted by first informative column)
  int Q = 0;
  std::vector<int> order;     // sorted position -> caller's index
  std::vector<int> qM, qlo, qhi, qstat;
  DBuf b_codes, b_qM;
  // stage 2: longest masked length of the register-resident classes, and the
  // longest resident query beyond it (0: none; those take the streaming form)
  int reg_limit = 0, long_cap = 0;
  // work buffers
  DBuf b_tcodes, b_tile_lo, b_tile_hi, b_tile_q, b_ttab, b_scores, b_res, b_nres, b_status,
      b_kept, b_nkeep, b_out, b_nout, b_qlist, b_counter, b_stats;
  // the FP64 re-placement of an FP32 pass's assertion failures: its query
  // set and outputs while the pass's own are set aside
  DBuf b_rcodes, b_rqM, b_rout, b_rnout, b_rstatus, b_ridx;
  // results staging (sorted order)
  std::vector<pp_placement> h_out;
  std::vector<int> h_nout, h_status;
  // device GLVs made on this context (pp_glv_make); freed with it
  std::set<pp_glv *> glvs;
  int K_last = 0;
};

namespace {

int set_device(pp_ctx *c) {
  HIPCHECK(hipSetDevice(c->device));
  return PP_OK;
}

// stage-1 table: every character's LV evolved by start_pend (the rows of
// full_query_evolv, core.ml:137-141), [n_codes][R*S], FP64 as the stubs do
// [n_tsl][n_codes][R*S]: n_tsl = 1 for the gamma model (rate r uses matrix
// r); with Price-CAT one table per category (R == 1).
std::vector<double> query_table(const HostModel &m, double bl) {
  const int R = m.R, S = m.S, n_sl = m.n_sl(), n_tsl = m.n_cats > 0 ? m.n_cats : 1;
  std::vector<double> P((size_t)n_sl * S * S), T((size_t)n_tsl * m.n_codes * R * S);
  multi_exp(m, bl, P.data());
  for (int tsl = 0; tsl < n_tsl; tsl++)
    for (int c = 0; c < m.n_codes; c++)
      for (int r = 0; r < R; r++) {
        const double *Pr = P.data() + (size_t)(m.n_cats > 0 ? tsl : r) * S * S;
        for (int i = 0; i < S; i++) {
          double acc = 0;
          for (int j = 0; j < S; j++) acc += Pr[i * S + j] * m.lv[c * S + j];
          T[(((size_t)tsl * m.n_codes + c) * R + r) * S + i] = acc;
        }
      }
  return T;
}

template <typename Real>
void to_device_layout(const TreeGlvs &g, int G, int N, int R, int S, int Epad,
                      std::vector<Real> &darr, std::vector<Real> &parr, std::vector<Real> &sn) {
  const int RS = R * S;
  darr.assign((size_t)G * N * RS, 0);
  parr.assign((size_t)G * N * RS, 0);
  sn.assign((size_t)N * Epad * RS, 0);
  for (int i = 0; i < G; i++)
    for (int r = 0; r < R; r++)
      for (int n = 0; n < N; n++)
        for (int s = 0; s < S; s++) {
          const size_t src = (((size_t)i * R + r) * N + n) * S + s;
          const size_t dst = ((size_t)i * N + n) * RS + r * S + s;
          darr[dst] = (Real)g.darr[src];
          parr[dst] = (Real)g.parr[src];
          if (i < G - 1) sn[((size_t)n * Epad + i) * RS + r * S + s] = (Real)g.snodes[src];
        }
}

int make_fp32_copies(pp_ctx *c);

// Price-CAT: the site categories on the device (nullptr for the gamma model);
// upload_site_cat checks them against the tree's site count.
const int *site_cat_ptr(const pp_ctx *c) {
  return c->model.n_cats > 0 ? reinterpret_cast<const int *>(c->b_cat.p) : nullptr;
}

int upload_site_cat(pp_ctx *c, int N) {
  if (c->model.n_cats == 0) return PP_OK;
  CHECK((int)c->model.site_cat.size() == N, PP_ESTATE,
        "Price-CAT model: pp_site_categories must give one category per site");
  int *d;
  ENSURE(c->b_cat, std::max(N, 1), &d);
  HIPCHECK(hipMemcpy(d, c->model.site_cat.data(), (size_t)N * 4, hipMemcpyHostToDevice));
  return PP_OK;
}

int upload_glvs(pp_ctx *c, const TreeGlvs &g) {
  const int G = c->G, N = c->N, R = c->model.R, S = c->model.S, Epad = c->Epad;
  if (S != 4 && S != 20) { set_error("the GPU placement path supports n_states 4 and 20"); return PP_EINVAL; }
  const int st_cat = upload_site_cat(c, N);
  if (st_cat) return st_cat;
  const size_t glv = (size_t)G * N * R * S, snn = (size_t)N * Epad * R * S;
  double *d64, *p64, *s64;
  int *de, *pe, *se;
  ENSURE(c->b_d64, glv, &d64);
  ENSURE(c->b_p64, glv, &p64);
  ENSURE(c->b_s64, snn, &s64);
  ENSURE(c->b_de, (size_t)G * N, &de);
  ENSURE(c->b_pe, (size_t)G * N, &pe);
  ENSURE(c->b_se, (size_t)N * Epad, &se);
  {
    std::vector<double> a, b, s;
    to_device_layout<double>(g, G, N, R, S, Epad, a, b, s);
    HIPCHECK(hipMemcpy(d64, a.data(), a.size() * 8, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(p64, b.data(), b.size() * 8, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(s64, s.data(), s.size() * 8, hipMemcpyHostToDevice));
  }
  std::vector<int> ed((size_t)G * N), ep((size_t)G * N), es((size_t)N * Epad, 0);
  for (size_t i = 0; i < ed.size(); i++) { ed[i] = (int)g.darr_e[i]; ep[i] = (int)g.parr_e[i]; }
  for (int i = 0; i < G - 1; i++)
    for (int n = 0; n < N; n++) es[(size_t)n * Epad + i] = (int)g.snodes_e[(size_t)i * N + n];
  HIPCHECK(hipMemcpy(de, ed.data(), ed.size() * 4, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(pe, ep.data(), ep.size() * 4, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(se, es.data(), es.size() * 4, hipMemcpyHostToDevice));
  int st = make_fp32_copies(c);
  if (st) return st;
  double *dbl;
  ENSURE(c->b_bl, G, &dbl);
  HIPCHECK(hipMemcpy(dbl, c->bl.data(), G * 8, hipMemcpyHostToDevice));
  c->has_tree = true;
  return PP_OK;
}

// FP32 copies of the resident FP64 GLVs with their own exponents
int make_fp32_copies(pp_ctx *c) {
  const int G = c->G, N = c->N, RS = c->model.R * c->model.S;
  float *d32, *p32;
  int *de32, *pe32;
  ENSURE(c->b_d32, (size_t)G * N * RS, &d32);
  ENSURE(c->b_p32, (size_t)G * N * RS, &p32);
  ENSURE(c->b_de32, (size_t)G * N, &de32);
  ENSURE(c->b_pe32, (size_t)G * N, &pe32);
  const dim3 grid(4096), block(256);
  hipLaunchKernelGGL(fp32_rescale_kernel, grid, block, 0, c->stream, (const double *)c->b_d64.p,
                     (const int *)c->b_de.p, d32, de32, (long)G * N, RS);
  hipLaunchKernelGGL(fp32_rescale_kernel, grid, block, 0, c->stream, (const double *)c->b_p64.p,
                     (const int *)c->b_pe.p, p32, pe32, (long)G * N, RS);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(c->stream));
  return PP_OK;
}

// GPU reference-tree precompute (see pre:: kernels).  Node levels come from
// the post-order ids: heights for the distal pass, depths for the proximal.
template <int R, int S>
int gpu_precompute_t(pp_ctx *c, const TreeDesc &t, const std::vector<uint8_t> &codes,
                     const std::vector<int> &parent, const std::vector<int> &height,
                     const std::vector<int> &depth) {
  constexpr int RS = R * S;
  const int G = c->G, N = c->N, E = c->E, Epad = c->Epad;
  const size_t glv = (size_t)G * N * RS, snn = (size_t)N * Epad * RS;
  double *d64, *p64, *s64, *P, *Ph, *lv, *statd;
  int *de, *pe, *se, *cs, *ch, *par, *lvl;
  uint8_t *lc;
  ENSURE(c->b_d64, glv, &d64);
  ENSURE(c->b_p64, glv, &p64);
  ENSURE(c->b_s64, snn, &s64);
  ENSURE(c->b_de, (size_t)G * N, &de);
  ENSURE(c->b_pe, (size_t)G * N, &pe);
  ENSURE(c->b_se, (size_t)N * Epad, &se);
  DBuf b_evd, b_evde, b_P, b_Ph, b_cs, b_ch, b_par, b_lvl, b_lc, b_lv, b_sd;
  struct Freer {
    std::vector<DBuf *> v;
    ~Freer() { for (DBuf *b : v) if (b->p) (void)hipFree(b->p); }
  } freer{{&b_evd, &b_evde, &b_P, &b_Ph, &b_cs, &b_ch, &b_par, &b_lvl, &b_lc, &b_lv, &b_sd}};
  double *evd;
  int *evde;
  ENSURE(b_evd, glv, &evd);
  ENSURE(b_evde, (size_t)G * N, &evde);
  const int n_sl = c->model.n_sl();
  ENSURE(b_P, (size_t)G * n_sl * S * S, &P);
  ENSURE(b_Ph, (size_t)G * n_sl * S * S, &Ph);
  ENSURE(b_cs, G + 1, &cs);
  ENSURE(b_ch, std::max(1, t.child_start[G]), &ch);
  ENSURE(b_par, G, &par);
  ENSURE(b_lvl, G, &lvl);
  ENSURE(b_lc, (size_t)G * N, &lc);
  ENSURE(b_lv, c->model.lv.size(), &lv);
  ENSURE(b_sd, S, &statd);
  // transition matrices P(bl) and P(bl/2) for every node, FP64 on the host
  std::vector<double> hP((size_t)G * n_sl * S * S), hPh((size_t)G * n_sl * S * S);
  for (int i = 0; i < G; i++) {
    multi_exp(c->model, t.bl[i], hP.data() + (size_t)i * n_sl * S * S);
    multi_exp(c->model, t.bl[i] / 2., hPh.data() + (size_t)i * n_sl * S * S);
  }
  HIPCHECK(hipMemcpy(P, hP.data(), hP.size() * 8, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(Ph, hPh.data(), hPh.size() * 8, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(cs, t.child_start, (G + 1) * 4, hipMemcpyHostToDevice));
  if (t.child_start[G] > 0) HIPCHECK(hipMemcpy(ch, t.children, t.child_start[G] * 4, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(par, parent.data(), G * 4, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(lc, codes.data(), codes.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(lv, c->model.lv.data(), c->model.lv.size() * 8, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(statd, c->model.statd.data(), S * 8, hipMemcpyHostToDevice));
  pre::TreeArgs a;
  a.N = N;
  a.n_nodes = G;
  a.child_start = cs;
  a.children = ch;
  a.parent = par;
  a.P = P;
  a.Phalf = Ph;
  a.site_cat = site_cat_ptr(c);
  a.n_sl = n_sl;
  a.leaf_codes = lc;
  a.lv = lv;
  a.darr = d64;
  a.parr = p64;
  a.evd = evd;
  a.darr_e = de;
  a.parr_e = pe;
  a.evd_e = evde;
  // post-order by height
  const int max_h = *std::max_element(height.begin(), height.end());
  std::vector<int> order(G);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return height[x] < height[y]; });
  HIPCHECK(hipMemcpy(lvl, order.data(), G * 4, hipMemcpyHostToDevice));
  for (int h = 0, pos = 0; h <= max_h; h++) {
    int n = 0;
    while (pos + n < G && height[order[pos + n]] == h) n++;
    if (!n) continue;
    a.level_nodes = lvl + pos;
    a.n_level = n;
    const long work = (long)n * N * R;
    hipLaunchKernelGGL((pre::distal_level_kernel<R, S>), dim3((work + 255) / 256), dim3(256), 0, c->stream, a);
    HIPCHECK(hipGetLastError());
    pos += n;
  }
  hipLaunchKernelGGL(pre::fill_unit_kernel, dim3(256), dim3(256), 0, c->stream, p64 + (size_t)(G - 1) * N * RS,
                     pe + (size_t)(G - 1) * N, N, RS);
  HIPCHECK(hipGetLastError());
  // pre-order by depth (children of depth-d nodes are at depth d+1)
  const int max_d = *std::max_element(depth.begin(), depth.end());
  std::vector<int> dorder;
  for (int i = 0; i < G - 1; i++) dorder.push_back(i);
  std::stable_sort(dorder.begin(), dorder.end(), [&](int x, int y) { return depth[x] < depth[y]; });
  DBuf b_dl;
  int *dl;
  ENSURE(b_dl, std::max<size_t>(dorder.size(), 1), &dl);
  if (!dorder.empty()) HIPCHECK(hipMemcpy(dl, dorder.data(), dorder.size() * 4, hipMemcpyHostToDevice));
  for (int d = 1, pos = 0; d <= max_d; d++) {
    int n = 0;
    while (pos + n < (int)dorder.size() && depth[dorder[pos + n]] == d) n++;
    if (!n) continue;
    a.level_nodes = dl + pos;
    a.n_level = n;
    const long work = (long)n * N * R;
    hipLaunchKernelGGL((pre::proximal_level_kernel<R, S>), dim3((work + 255) / 256), dim3(256), 0, c->stream, a);
    HIPCHECK(hipGetLastError());
    pos += n;
  }
  HIPCHECK(hipMemsetAsync(s64, 0, snn * 8, c->stream));
  HIPCHECK(hipMemsetAsync(se, 0, (size_t)N * Epad * 4, c->stream));
  const long work = (long)G * N * R;
  hipLaunchKernelGGL((pre::finish_kernel<R, S>), dim3((work + 255) / 256), dim3(256), 0, c->stream, a, statd, E,
                     Epad, s64, se);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(c->stream));
  (void)hipFree(b_dl.p);
  int st = make_fp32_copies(c);
  if (st) return st;
  double *dbl;
  ENSURE(c->b_bl, G, &dbl);
  HIPCHECK(hipMemcpy(dbl, c->bl.data(), G * 8, hipMemcpyHostToDevice));
  c->has_tree = true;
  return PP_OK;
}

int gpu_precompute(pp_ctx *c, const TreeDesc &t, const char *leaf_seqs) {
  const int G = t.n_nodes, N = t.n_sites, S = c->model.S, R = c->model.R;
  std::vector<int> parent(G, -1), height(G, 0), depth(G, 0);
  for (int id = 0; id < G; id++)
    for (int k = t.child_start[id]; k < t.child_start[id + 1]; k++) {
      const int ch = t.children[k];
      CHECK(ch >= 0 && ch < id, PP_EINVAL, "node ids must be post-order (children before parents)");
      CHECK(parent[ch] < 0, PP_EINVAL, "a node has two parents");
      parent[ch] = id;
      height[id] = std::max(height[id], height[ch] + 1);
    }
  for (int id = 0; id < G - 1; id++) CHECK(parent[id] >= 0, PP_EINVAL, "the root must be the last node id");
  for (int id = G - 2; id >= 0; id--) depth[id] = depth[parent[id]] + 1;
  std::vector<uint8_t> codes((size_t)G * N, 0);
  for (int id = 0; id < G; id++) {
    if (t.child_start[id + 1] != t.child_start[id]) continue;
    for (int k = 0; k < N; k++) {
      const char ch = leaf_seqs[(size_t)id * N + k];
      const int code = c->model.code_of[(unsigned char)ch];
      CHECK(code >= 0, PP_EINVAL, std::string(1, ch) + " not a known base in leaf " + std::to_string(id));
      codes[(size_t)id * N + k] = (uint8_t)code;
    }
  }
  const int st_cat = upload_site_cat(c, N);
  if (st_cat) return st_cat;
  if (S == 4 && R == 4) return gpu_precompute_t<4, 4>(c, t, codes, parent, height, depth);
  if (S == 4 && R == 1) return gpu_precompute_t<1, 4>(c, t, codes, parent, height, depth);
  if (S == 20 && R == 4) return gpu_precompute_t<4, 20>(c, t, codes, parent, height, depth);
  if (S == 20 && R == 1) return gpu_precompute_t<1, 20>(c, t, codes, parent, height, depth);
  if (S != 4 && S != 20) { set_error("the GPU placement path supports n_states 4 and 20"); return PP_EINVAL; }
  set_error("n_rates must be 1 or 4 on the GPU path");
  return PP_EINVAL;
}

int bucket_of(int M) { return M <= 0 ? 0 : (M - 1) / 64; }

// Masked sites per worker lane in stage 2 for bucket b.  The FP64 kernels
// hold one site's GLV rows per lane (two sites' rows spill at the 4 waves per
// SIMD the kernels are built for); queries beyond 15 worker waves of one site
// take two.  The FP32 kernels always hold two.
int spl_for(bool fp64, int b) { return fp64 && b < 15 ? 1 : 2; }
// one master wave + ceil(64 (b + 1) / (64 spl)) worker waves
int block_for(int b, int spl) { return 64 * (1 + (b + spl) / spl); }

// The stage-1 log-util table for one start_pend (see code_log_kernel).
// Built from the FP64 supernodes for both forms; the FP32 form stores it in
// FP32.
template <typename Real>
int build_code_logs(pp_ctx *c, double start_pend) {
  const int N = c->N, Epad = c->Epad, n_inf = c->model.n_inf, R = c->model.R;
  c->h_ttab = query_table(c->model, start_pend);
  double *d_ttab;
  Real *lt;
  ENSURE(c->b_ttab, c->h_ttab.size(), &d_ttab);
  ENSURE(c->b_lt, (size_t)N * (n_inf + 1) * Epad, &lt);
  HIPCHECK(hipMemcpyAsync(d_ttab, c->h_ttab.data(), c->h_ttab.size() * 8, hipMemcpyHostToDevice,
                          c->stream));
  CodeLogArgs a;
  a.sn = reinterpret_cast<const double *>(c->b_s64.p);
  a.N = N;
  a.Epad = Epad;
  a.n_inf = n_inf;
  a.ttab = d_ttab;
  a.n_codes = c->model.n_codes;
  a.site_cat = site_cat_ptr(c);
  a.lt64 = std::is_same<Real, double>::value ? reinterpret_cast<double *>(lt) : nullptr;
  a.lt32 = std::is_same<Real, float>::value ? reinterpret_cast<float *>(lt) : nullptr;
  const dim3 grid(Epad / 256, N);
  const int RS = R * c->model.S;
  if (RS == 16) hipLaunchKernelGGL(code_log_kernel<16>, grid, dim3(256), 0, c->stream, a);
  else if (RS == 4) hipLaunchKernelGGL(code_log_kernel<4>, grid, dim3(256), 0, c->stream, a);
  else if (RS == 80) hipLaunchKernelGGL(code_log_kernel<80>, grid, dim3(256), 0, c->stream, a);
  else if (RS == 20) hipLaunchKernelGGL(code_log_kernel<20>, grid, dim3(256), 0, c->stream, a);
  else { set_error("n_rates must be 1 or 4 on the GPU path"); return PP_EINVAL; }
  HIPCHECK(hipGetLastError());
  return PP_OK;
}

// Stage-1 tiles of queries [q0, q0 + nq): kTQ queries per tile
// (chunk-relative indices, -1 for empty slots and failed queries) and the
// span of columns their informative characters cover, appended to lo/hi/q.
// A tile's kernel walks every column of its span.  The resident queries are
// stored by start column, so 16 consecutive ones start together but their
// tile spans the longest of them; here they are ordered by start column in
// windows of W columns (W sized so that a window holds about 128 queries),
// then by end column, so that a tile's queries also end together and its
// span is about their common length plus W.  Each (query, edge) sum is the
// same whichever tile holds the query.
void stage1_tiles(const pp_ctx *c, int q0, int nq, std::vector<int> &lo_v, std::vector<int> &hi_v,
                  std::vector<int> &q_v) {
  const int N = c->N;
  const int n_tiles = (nq + kTQ - 1) / kTQ;
  std::vector<int> ok;
  ok.reserve(nq);
  for (int l = 0; l < nq; l++)
    if (c->qstat[q0 + l] == PP_Q_OK) ok.push_back(l);
  const long long n_ok = std::max<long long>(1, (long long)ok.size());
  const int W = (int)std::max(1LL, ((long long)N * 128 + n_ok - 1) / n_ok);
  {
    // two stable counting sorts (end column, then window): the order of a
    // stable sort by (window, end), in O(queries + N)
    const int n_win = (N - 1) / W + 1;
    std::vector<int> tmp(ok.size()), cnt((size_t)std::max(N + 1, n_win) + 1, 0);
    for (int l : ok) cnt[c->qhi[q0 + l] + 1]++;
    for (int i = 1; i <= N + 1; i++) cnt[i] += cnt[i - 1];
    for (int l : ok) tmp[cnt[c->qhi[q0 + l]]++] = l;
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int l : tmp) cnt[c->qlo[q0 + l] / W + 1]++;
    for (int i = 1; i <= n_win; i++) cnt[i] += cnt[i - 1];
    for (int l : tmp) ok[cnt[c->qlo[q0 + l] / W]++] = l;
  }
  for (int t = 0; t < n_tiles; t++) {
    int lo = N, hi = 0;
    for (int j = 0; j < kTQ; j++) {
      const size_t i = (size_t)t * kTQ + j;
      if (i >= ok.size()) {
        q_v.push_back(-1);
        continue;
      }
      const int l = ok[i], q = q0 + l;
      q_v.push_back(l);
      lo = std::min(lo, c->qlo[q]);
      hi = std::max(hi, c->qhi[q]);
    }
    if (hi < lo) { lo = 0; hi = 0; }
    lo_v.push_back(lo);
    hi_v.push_back(hi);
  }
}

// Stage-1 scores of queries [q0, q0 + nq) into scores[nq][E] from tiles
// already on the device; build_code_logs<Real> must have run for the same
// start_pend.  Launches on c->cur.
template <typename Real>
int launch_stage1_tiles(pp_ctx *c, int q0, int nq, double *scores, const int *d_tlo,
                        const int *d_thi, const int *d_tq) {
  const int N = c->N, R = c->model.R;
  const int n_tiles = (nq + kTQ - 1) / kTQ;
  uint8_t *d_tc;
  ENSURE(c->b_tcodes, (size_t)n_tiles * N * kTQ, &d_tc);
  const uint8_t *codes = reinterpret_cast<const uint8_t *>(c->b_codes.p) + (size_t)q0 * N;
  hipLaunchKernelGGL(tile_codes_kernel, dim3(2048), dim3(256), 0, c->cur, codes, d_tq, n_tiles,
                     N, kTQ, (uint8_t)c->model.n_inf, d_tc);
  Stage1Args a;
  a.lt = c->b_lt.p;
  a.N = N;
  a.E = c->E;
  a.Epad = c->Epad;
  a.C1 = c->model.n_inf + 1;
  a.sn_e = reinterpret_cast<const int *>(c->b_se.p);
  a.tcodes = d_tc;
  a.tile_lo = d_tlo;
  a.tile_hi = d_thi;
  a.tile_q = d_tq;
  a.qM = reinterpret_cast<const int *>(c->b_qM.p) + q0;
  a.log_R = std::log((double)(float)R);
  a.scores = scores;
  a.n_tiles = n_tiles;
  const long long items = (long long)n_tiles * (c->Epad / 256);
  CHECK(items < (1LL << 30), PP_EINVAL, "stage-1 grid too large");
  a.per_xcd = (int)((items + 7) / 8);
  if (items == 0) return PP_OK;
  hipLaunchKernelGGL((stage1_kernel<Real, kTQ>), dim3(8 * a.per_xcd), dim3(256), 0, c->cur, a);
  HIPCHECK(hipGetLastError());
  return PP_OK;
}

// Stage-1 scores of queries [q0, q0 + nq) into scores[nq][E], tiles uploaded
// here (synchronous copies).
template <typename Real>
int launch_stage1(pp_ctx *c, int q0, int nq, double *scores) {
  std::vector<int> lo, hi, tq;
  stage1_tiles(c, q0, nq, lo, hi, tq);
  const int n_tiles = (int)lo.size();
  int *d_tlo, *d_thi, *d_tq;
  ENSURE(c->b_tile_lo, std::max(n_tiles, 1), &d_tlo);
  ENSURE(c->b_tile_hi, std::max(n_tiles, 1), &d_thi);
  ENSURE(c->b_tile_q, std::max(n_tiles, 1) * kTQ, &d_tq);
  HIPCHECK(hipStreamSynchronize(c->cur));
  HIPCHECK(hipMemcpy(d_tlo, lo.data(), n_tiles * 4, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d_thi, hi.data(), n_tiles * 4, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d_tq, tq.data(), (size_t)n_tiles * kTQ * 4, hipMemcpyHostToDevice));
  return launch_stage1_tiles<Real>(c, q0, nq, scores, d_tlo, d_thi, d_tq);
}

template <typename Real, int R>
int occupancy_grid(const void *kern, int bd, size_t lds, int n_cu, int n_items) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, bd, lds) != hipSuccess || nb < 1) nb = 1;
  return std::max(1, std::min(n_items, nb * n_cu));
}

// Stage-2 kernels: one master wave + ceil(M / (64 spl)) worker waves.  Three
// launch-bound classes (512, 768, 1024 lanes) bound the block size; the
// kernels are built for 4 waves per SIMD (MWOcc).  spl == 0 is the streaming
// form: 1024 lanes, any masked length up to stream_cap.
// LDS: header and tables, per-site logs / list / pendant rows (up to the
// launch's capacity), the pitch kernel's per-edge flags, and in the
// register-resident forms the cached products of the worker lanes' sites.
template <typename Real, int R, bool STRICT>
size_t stage2_lds(const pp_ctx *c, int bd, size_t taken, int spl, int stream_cap) {
  const int cap = spl == 0 ? stream_cap : (bd - 64) * spl;
  const MWLayout<Real, R, STRICT> lay(c->model.n_cats > 0 ? c->model.n_cats : 1, c->model.n_sl(), cap, spl == 0);
  return lay.align(lay.align(lay.bytes()) + taken) + (spl ? MWProducts<Real, R * 4>::bytes(cap) : 0);
}
size_t stage2_lds_fp64(const pp_ctx *c, int bd, size_t taken, int spl, int stream_cap) {
  return c->model.R == 4 ? stage2_lds<double, 4, true>(c, bd, taken, spl, stream_cap)
                         : stage2_lds<double, 1, true>(c, bd, taken, spl, stream_cap);
}

// Whether the FP32 form runs stage 2 in FP32.  The 20-state engine evaluates
// evolves in the eigenbasis of Q, which rebuilds small GLV entries from
// eigen coordinates of larger magnitude; in FP32 that cancellation loses
// them, so 20-state stage 2 always runs in FP64 (the FP32 form keeps its
// FP32 stage-1 table).
// The FP32 stage-2 engines: 4 states (all models), 20 states with gamma rates
// (the state-basis MFMA form needs one rate per block of four pairs).
bool stage2_fp32_ok(const pp_ctx *c) {
  return c->model.S == 4 || (c->model.S == 20 && c->model.n_cats == 0 && (c->model.R == 1 || c->model.R == 4));
}

// Longest masked length of the register-resident 4-state classes whose
// workgroup fits in LDS (FP64, the larger form, with the pitch kernel's flags).
int register_limit(const pp_ctx *c) {
  int m = 0;
  for (int b = 0; b < kLongBucket; b++) {
    const int spl = spl_for(true, b), bd = block_for(b, spl);
    if (bd > 1024) break;
    if (stage2_lds_fp64(c, bd, taken_bytes(c->E), spl, 0) > c->max_lds) break;
    m = 64 * (b + 1);
  }
  return m;
}

// Longest masked length the GPU path places: N for 20 states (pair state in
// global scratch); for 4 states the streaming form's, whose LDS holds 16
// bytes per masked site (log, column, pendant row) besides the header and
// flags, so about 9,000 sites in the 160 KB of a gfx950 workgroup.  Longer
// queries get PP_Q_RANK_EXHAUSTED ("query too long for the GPU path").
int max_masked_sites(const pp_ctx *c) {
  if (c->model.S == 20) return c->N;
  int lo = 0, hi = c->N;
  while (lo < hi) {  // largest cap whose streaming workgroup fits
    const int mid = lo + (hi - lo + 1) / 2;
    if (stage2_lds_fp64(c, 1024, taken_bytes(c->E), 0, mid) <= c->max_lds) lo = mid;
    else hi = mid - 1;
  }
  return std::max(lo, std::min(register_limit(c), c->N));
}

// Stage-2 class of a query of masked length M (4 states): its 64-site class
// when the register-resident form holds it, else the streaming class.
// 20 states: one class (the kernel takes any masked length up to the
// launch's capacity), so a chunk's stage 2 is one launch per phase.
int stage2_bucket(const pp_ctx *c, int M) {
  if (c->model.S == 20) return 0;
  return M <= c->reg_limit ? bucket_of(M) : kLongBucket;
}

int ensure_prot_scratch(pp_ctx *c);

// Product scratch of the streaming form: one slice of long_cap sites per
// workgroup the device can hold at once (1024-lane blocks: at most two per
// CU); sized once before a pass, so that no launch finds it reallocated.
// The 20-state kernels' scratch is reserved here too.
int ensure_stream_scratch(pp_ctx *c) {
  if (c->model.S == 20) return ensure_prot_scratch(c);
  if (c->model.S != 4 || c->long_cap == 0) return PP_OK;
  const int blocks = 2 * c->n_cu;
  double *p;
  ENSURE(c->b_sscr, (size_t)blocks * c->long_cap * c->model.R * 4, &p);
  c->sscr_blocks = blocks;
  return PP_OK;
}

// Grid of a 4-state stage-2 launch: the workgroups the device holds at once
// (the work queue spreads the items over them); the streaming form also
// hands each workgroup its product scratch slice.
template <typename Args, int SPL>
int stage2_grid(pp_ctx *c, const void *kern, int bd, size_t lds, int n_items, Args &a) {
  int grid = occupancy_grid<double, 4>(kern, bd, lds, c->n_cu, n_items);
  if constexpr (SPL == 0) {
    grid = std::min(grid, c->sscr_blocks);
    a.scratch = c->b_sscr.p;
    a.cap_sites = c->long_cap;
  } else {
    a.cap_sites = 0;
  }
  return grid;
}

template <typename Real, int R, int MAXBD, int SPL, bool CAT, bool STRICT>
int launch_pitch_b(pp_ctx *c, PitchArgs a, int bd) {
  const size_t lds = stage2_lds<Real, R, STRICT>(c, bd, taken_bytes(c->E), SPL, c->long_cap);
  const void *k = reinterpret_cast<const void *>(&pitch_mw_kernel<Real, R, MAXBD, SPL, CAT, STRICT>);
  const int grid = stage2_grid<PitchArgs, SPL>(c, k, bd, lds, a.nq, a);
  CHECK(grid > 0, PP_ESTATE, "streaming-form scratch not allocated");
  hipLaunchKernelGGL((pitch_mw_kernel<Real, R, MAXBD, SPL, CAT, STRICT>), dim3(grid), dim3(bd), lds, c->cur, a);
  HIPCHECK(hipGetLastError());
  return PP_OK;
}

template <typename Real, int R, int MAXBD, int SPL, bool CAT, bool STRICT>
int launch_opt_b(pp_ctx *c, OptArgs a, int bd) {
  const size_t lds = stage2_lds<Real, R, STRICT>(c, bd, 0, SPL, c->long_cap);
  const void *k = reinterpret_cast<const void *>(&optimize_mw_kernel<Real, R, MAXBD, SPL, CAT, STRICT>);
  const int grid = stage2_grid<OptArgs, SPL>(c, k, bd, lds, a.n_jobs, a);
  CHECK(grid > 0, PP_ESTATE, "streaming-form scratch not allocated");
  hipLaunchKernelGGL((optimize_mw_kernel<Real, R, MAXBD, SPL, CAT, STRICT>), dim3(grid), dim3(bd), lds, c->cur, a);
  HIPCHECK(hipGetLastError());
  return PP_OK;
}

// Bucket b's launch: FP64 with one site per lane up to 15 worker waves, two
// beyond (blocks of more than 512 lanes); FP32 with two; the long bucket in
// the streaming form.
template <typename Real, int R, bool STRICT, bool CAT = false>
int launch_pitch(pp_ctx *c, PitchArgs a, int b) {
  if (b == kLongBucket) return launch_pitch_b<Real, R, 1024, 0, CAT, STRICT>(c, a, 1024);
  const int spl = spl_for(sizeof(Real) == 8, b), bd = block_for(b, spl);
  if (spl == 1) {
    if constexpr (sizeof(Real) == 8) {
      if (bd <= 512) return launch_pitch_b<Real, R, 512, 1, CAT, STRICT>(c, a, bd);
      if (bd <= 768) return launch_pitch_b<Real, R, 768, 1, CAT, STRICT>(c, a, bd);
      return launch_pitch_b<Real, R, 1024, 1, CAT, STRICT>(c, a, bd);
    }
  }
  if constexpr (sizeof(Real) == 4) {
    if (bd <= 512) return launch_pitch_b<Real, R, 512, 2, CAT, STRICT>(c, a, bd);
  }
  if (bd <= 768) return launch_pitch_b<Real, R, 768, 2, CAT, STRICT>(c, a, bd);
  return launch_pitch_b<Real, R, 1024, 2, CAT, STRICT>(c, a, bd);
}

template <typename Real, int R, bool STRICT, bool CAT = false>
int launch_opt(pp_ctx *c, OptArgs a, int b) {
  if (b == kLongBucket) return launch_opt_b<Real, R, 1024, 0, CAT, STRICT>(c, a, 1024);
  const int spl = spl_for(sizeof(Real) == 8, b), bd = block_for(b, spl);
  if (spl == 1) {
    if constexpr (sizeof(Real) == 8) {
      if (bd <= 512) return launch_opt_b<Real, R, 512, 1, CAT, STRICT>(c, a, bd);
      if (bd <= 768) return launch_opt_b<Real, R, 768, 1, CAT, STRICT>(c, a, bd);
      return launch_opt_b<Real, R, 1024, 1, CAT, STRICT>(c, a, bd);
    }
  }
  if constexpr (sizeof(Real) == 4) {
    if (bd <= 512) return launch_opt_b<Real, R, 512, 2, CAT, STRICT>(c, a, bd);
  }
  if (bd <= 768) return launch_opt_b<Real, R, 768, 2, CAT, STRICT>(c, a, bd);
  return launch_opt_b<Real, R, 1024, 2, CAT, STRICT>(c, a, bd);
}

template <int R, int MAXBD, int SPL, bool CAT, bool STRICT>
int launch_marg_b(pp_ctx *c, OptArgs a, int bd) {
  const size_t lds = stage2_lds<double, R, STRICT>(c, bd, 0, SPL, c->long_cap);
  const void *k = reinterpret_cast<const void *>(&marginal_mw_kernel<R, MAXBD, SPL, CAT, STRICT>);
  const int grid = stage2_grid<OptArgs, SPL>(c, k, bd, lds, a.n_jobs, a);
  CHECK(grid > 0, PP_ESTATE, "streaming-form scratch not allocated");
  hipLaunchKernelGGL((marginal_mw_kernel<R, MAXBD, SPL, CAT, STRICT>), dim3(grid), dim3(bd), lds, c->cur, a);
  HIPCHECK(hipGetLastError());
  return PP_OK;
}

template <int R, bool STRICT, bool CAT = false>
int launch_marg(pp_ctx *c, OptArgs a, int b) {
  if (b == kLongBucket) return launch_marg_b<R, 1024, 0, CAT, STRICT>(c, a, 1024);
  const int spl = spl_for(true, b), bd = block_for(b, spl);
  if (spl == 1) {
    if (bd <= 512) return launch_marg_b<R, 512, 1, CAT, STRICT>(c, a, bd);
    if (bd <= 768) return launch_marg_b<R, 768, 1, CAT, STRICT>(c, a, bd);
    return launch_marg_b<R, 1024, 1, CAT, STRICT>(c, a, bd);
  }
  if (bd <= 768) return launch_marg_b<R, 768, 2, CAT, STRICT>(c, a, bd);
  return launch_marg_b<R, 1024, 2, CAT, STRICT>(c, a, bd);
}

// 20-state stage 2: one master + 7 worker waves per query.  FP64 form: pairs
// per lane as the longest masked length needs, per-workgroup pair state (a, a,
// G); FP32 form: 64-pair tiles, per-workgroup w of every pair.
template <typename Real, int R>
size_t prot_lds(const pp_ctx *c) {
  if constexpr (std::is_same<Real, float>::value) return ProtLayout::align(Prot32Layout<R>::bytes(c->N));
  return ProtLayout::align(ProtLayout::bytes(c->N));
}

template <typename Real, int R>
int prot_cap_pairs(const pp_ctx *c) {
  int max_m = 1;
  for (int q = 0; q < c->Q; q++)
    if (c->qstat[q] == PP_Q_OK) max_m = std::max(max_m, c->qM[q]);
  if constexpr (std::is_same<Real, float>::value) {
    constexpr int st = Prot32Layout<R>::kTileSites;
    return (max_m + st - 1) / st * 64;
  }
  constexpr int worker_lanes = kProtBlock - 64;
  const int pairs = max_m * R;
  return (pairs + worker_lanes - 1) / worker_lanes * worker_lanes;
}

template <typename Real, int R>
size_t prot_scratch_bytes(int cap_pairs) {
  if constexpr (std::is_same<Real, float>::value)
    return sizeof(float) * Prot32Layout<R>::scratch_floats(cap_pairs / 64);
  return sizeof(double) * ProtLayout::scratch_reals(cap_pairs);
}

// Scratch of a 20-state kernel's full-occupancy grid.
template <typename Real, int R>
size_t prot_scratch_need(const pp_ctx *c, const void *kern, size_t lds) {
  const int grid_full = occupancy_grid<Real, R>(kern, kProtBlock, lds, c->n_cu, INT_MAX);
  return prot_scratch_bytes<Real, R>(prot_cap_pairs<Real, R>(c)) * (size_t)grid_full;
}

template <int R>
size_t prot_scratch_max(const pp_ctx *c) {
  const size_t lp32 = prot_lds<float, R>(c) + taken_bytes(c->E), lp64 = prot_lds<double, R>(c) + taken_bytes(c->E);
  size_t n = 0;
  n = std::max(n, prot_scratch_need<float, R>(c, reinterpret_cast<const void *>(&pitch_prot_kernel<float, R>), lp32));
  n = std::max(n, prot_scratch_need<double, R>(c, reinterpret_cast<const void *>(&pitch_prot_kernel<double, R>), lp64));
  n = std::max(n, prot_scratch_need<float, R>(c, reinterpret_cast<const void *>(&optimize_prot_kernel<float, R>),
                                              prot_lds<float, R>(c)));
  n = std::max(n, prot_scratch_need<double, R>(c, reinterpret_cast<const void *>(&optimize_prot_kernel<double, R>),
                                               prot_lds<double, R>(c)));
  n = std::max(n, prot_scratch_need<double, R>(c, reinterpret_cast<const void *>(&marginal_prot_kernel<R>),
                                               prot_lds<double, R>(c)));
  return n;
}

// The scratch every 20-state kernel of the current queries may use, allocated
// before a pass, so that no launch (nor a captured graph) sees it move: the
// kernels of one pass share the buffer and run on one stream.
int ensure_prot_scratch(pp_ctx *c) {
  if (c->model.S != 20) return PP_OK;
  const size_t n = c->model.R == 4 ? prot_scratch_max<4>(c) : prot_scratch_max<1>(c);
  char *p;
  ENSURE(c->b_scratch, n, &p);
  return PP_OK;
}

template <typename Real, int R>
int prot_scratch(pp_ctx *c, const void *kern, size_t lds, int cap_pairs, void **out) {
  const int grid_full = occupancy_grid<Real, R>(kern, kProtBlock, lds, c->n_cu, INT_MAX);
  char *p;
  const size_t bytes = prot_scratch_bytes<Real, R>(cap_pairs) * (size_t)grid_full;
  ENSURE(c->b_scratch, bytes, &p);
  *out = p;
  return PP_OK;
}

template <typename Real, int R>
int launch_pitch_prot(pp_ctx *c, PitchArgs a) {
  const size_t lds = prot_lds<Real, R>(c) + taken_bytes(c->E);
  const void *k = reinterpret_cast<const void *>(&pitch_prot_kernel<Real, R>);
  const int grid = occupancy_grid<Real, R>(k, kProtBlock, lds, c->n_cu, a.nq);
  a.cap_pairs = prot_cap_pairs<Real, R>(c);
  int st = prot_scratch<Real, R>(c, k, lds, a.cap_pairs, &a.scratch);
  if (st) return st;
  hipLaunchKernelGGL((pitch_prot_kernel<Real, R>), dim3(grid), dim3(kProtBlock), lds, c->cur, a);
  HIPCHECK(hipGetLastError());
  return PP_OK;
}

template <typename Real, int R>
int launch_opt_prot(pp_ctx *c, OptArgs a) {
  const size_t lds = prot_lds<Real, R>(c);
  const void *k = reinterpret_cast<const void *>(&optimize_prot_kernel<Real, R>);
  const int grid = occupancy_grid<Real, R>(k, kProtBlock, lds, c->n_cu, a.n_jobs);
  a.cap_pairs = prot_cap_pairs<Real, R>(c);
  int st = prot_scratch<Real, R>(c, k, lds, a.cap_pairs, &a.scratch);
  if (st) return st;
  hipLaunchKernelGGL((optimize_prot_kernel<Real, R>), dim3(grid), dim3(kProtBlock), lds, c->cur, a);
  HIPCHECK(hipGetLastError());
  return PP_OK;
}

template <int R>
int launch_marg_prot(pp_ctx *c, OptArgs a) {
  const size_t lds = prot_lds<double, R>(c);
  const void *k = reinterpret_cast<const void *>(&marginal_prot_kernel<R>);
  const int grid = occupancy_grid<double, R>(k, kProtBlock, lds, c->n_cu, a.n_jobs);
  a.cap_pairs = prot_cap_pairs<double, R>(c);
  int st = prot_scratch<double, R>(c, k, lds, a.cap_pairs, &a.scratch);
  if (st) return st;
  hipLaunchKernelGGL((marginal_prot_kernel<R>), dim3(grid), dim3(kProtBlock), lds, c->cur, a);
  HIPCHECK(hipGetLastError());
  return PP_OK;
}

// The -p prior on the device: the exponential mean of every edge (the flat
// prior repeats its mean); checks that one is known for this tree.
int upload_prior(pp_ctx *c, OptArgs *o) {
  CHECK(c->prior_set, PP_ESTATE, "no prior for -p: call pp_set_prior (pp_tree_upload installs none)");
  std::vector<double> m((size_t)c->G, 0.);
  if (c->prior_kind == PP_PRIOR_INFORMATIVE) {
    CHECK((int)c->prior_means.size() == c->G, PP_ESTATE, "informative prior: one mean per node is needed");
    m = c->prior_means;
  } else if (c->prior_kind == PP_PRIOR_FLAT_EXP) {
    std::fill(m.begin(), m.end(), c->prior_means[0]);
  }
  double *d;
  ENSURE(c->b_prior, (size_t)c->G, &d);
  HIPCHECK(hipMemcpyAsync(d, m.data(), m.size() * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  o->prior_kind = c->prior_kind == PP_PRIOR_UNIFORM ? PRIOR_UNIFORM : PRIOR_EXP;
  o->prior_mean = d;
  return PP_OK;
}

template <bool STRICT>
int dispatch_marg(pp_ctx *c, OptArgs a, int b) {
  a.site_cat = site_cat_ptr(c);
  if (c->model.S == 20) {
    if (c->model.R == 4) return launch_marg_prot<4>(c, a);
    if (c->model.R == 1) return launch_marg_prot<1>(c, a);
    set_error("n_rates must be 1 or 4 on the GPU path");
    return PP_EINVAL;
  }
  if (c->model.n_cats > 0) return launch_marg<1, STRICT, true>(c, a, b);
  if (c->model.R == 4) return launch_marg<4, STRICT>(c, a, b);
  if (c->model.R == 1) return launch_marg<1, STRICT>(c, a, b);
  set_error("n_rates must be 1 or 4 on the GPU path");
  return PP_EINVAL;
}

template <typename Real, bool STRICT>
int dispatch_pitch(pp_ctx *c, PitchArgs a, int b) {
  a.site_cat = site_cat_ptr(c);
  if (c->model.S == 20) {
    if (c->model.R == 4) return launch_pitch_prot<Real, 4>(c, a);
    if (c->model.R == 1) return launch_pitch_prot<Real, 1>(c, a);
    set_error("n_rates must be 1 or 4 on the GPU path");
    return PP_EINVAL;
  }
  if (c->model.n_cats > 0) return launch_pitch<Real, 1, STRICT, true>(c, a, b);
  if (c->model.R == 4) return launch_pitch<Real, 4, STRICT>(c, a, b);
  if (c->model.R == 1) return launch_pitch<Real, 1, STRICT>(c, a, b);
  set_error("n_rates must be 1 or 4 on the GPU path");
  return PP_EINVAL;
}
template <typename Real, bool STRICT>
int dispatch_opt(pp_ctx *c, OptArgs a, int b) {
  a.site_cat = site_cat_ptr(c);
  if (c->model.S == 20) {
    if (c->model.R == 4) return launch_opt_prot<Real, 4>(c, a);
    if (c->model.R == 1) return launch_opt_prot<Real, 1>(c, a);
    set_error("n_rates must be 1 or 4 on the GPU path");
    return PP_EINVAL;
  }
  if (c->model.n_cats > 0) return launch_opt<Real, 1, STRICT, true>(c, a, b);
  if (c->model.R == 4) return launch_opt<Real, 4, STRICT>(c, a, b);
  if (c->model.R == 1) return launch_opt<Real, 1, STRICT>(c, a, b);
  set_error("n_rates must be 1 or 4 on the GPU path");
  return PP_EINVAL;
}

// Encode and upload queries (sorted by first informative column).
int queries_upload(pp_ctx *c, int n_queries, const char *queries) {
  CHECK(c->model.built, PP_ESTATE, "no model: call pp_model_build first");
  CHECK(c->has_tree, PP_ESTATE, "no tree: call pp_tree_set or pp_tree_upload first");
  CHECK(n_queries >= 0, PP_EINVAL, "n_queries < 0");
  const int N = c->N;
  c->Q = n_queries;
  std::vector<int> lo(n_queries), hi(n_queries), M(n_queries), st(n_queries, PP_Q_OK);
  std::vector<uint8_t> codes((size_t)n_queries * N);
  const int n_inf = c->model.n_inf;
  const int max_m = max_masked_sites(c);
  c->reg_limit = c->model.S == 4 ? register_limit(c) : 0;
  c->long_cap = 0;
  for (int q = 0; q < n_queries; q++) {
    const char *s = queries + (size_t)q * N;
    uint8_t *o = codes.data() + (size_t)q * N;
    int l = N, h = 0, m = 0;
    for (int k = 0; k < N; k++) {
      int code = c->model.code_of[(unsigned char)s[k]];
      if (code < 0) { st[q] = PP_Q_BAD_CHAR; code = n_inf; }
      o[k] = (uint8_t)code;
      if (code < n_inf) { m++; l = std::min(l, k); h = k + 1; }
    }
    if (st[q] == PP_Q_OK && m == 0) st[q] = PP_Q_NO_INFORMATIVE;
    if (st[q] == PP_Q_OK && m > max_m) st[q] = PP_Q_RANK_EXHAUSTED;
    if (st[q] == PP_Q_OK && c->model.S == 4 && m > c->reg_limit) c->long_cap = std::max(c->long_cap, m);
    lo[q] = m ? l : 0;
    hi[q] = m ? h : 0;
    M[q] = m;
  }
  c->order.resize(n_queries);
  std::iota(c->order.begin(), c->order.end(), 0);
  std::stable_sort(c->order.begin(), c->order.end(), [&](int a, int b) { return lo[a] < lo[b]; });
  c->qM.resize(n_queries);
  c->qlo.resize(n_queries);
  c->qhi.resize(n_queries);
  c->qstat.resize(n_queries);
  std::vector<uint8_t> sorted((size_t)n_queries * N);
  for (int i = 0; i < n_queries; i++) {
    int q = c->order[i];
    c->qM[i] = M[q];
    c->qlo[i] = lo[q];
    c->qhi[i] = hi[q];
    c->qstat[i] = st[q];
    std::memcpy(sorted.data() + (size_t)i * N, codes.data() + (size_t)q * N, N);
  }
  uint8_t *d_codes;
  int *d_qM;
  ENSURE(c->b_codes, (size_t)n_queries * N, &d_codes);
  ENSURE(c->b_qM, n_queries, &d_qM);
  if (n_queries > 0) {
    HIPCHECK(hipMemcpy(d_codes, sorted.data(), sorted.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_qM, c->qM.data(), (size_t)n_queries * 4, hipMemcpyHostToDevice));
  }
  return PP_OK;
}

// Events of one placement pass, destroyed with it.
struct EventPool {
  std::vector<hipEvent_t> ev;
  hipEvent_t get() {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ev.push_back(e);
    return e;
  }
  ~EventPool() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
};

// FNV-1a over the bytes of the values a pass's launches derive from (the key
// of its captured graph).
struct PlanKey {
  uint64_t v = 1469598103934665603ull;
  void bytes(const void *p, size_t n) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; i++) {
      v ^= b[i];
      v *= 1099511628211ull;
    }
  }
  template <class T>
  void add(const T &x) { bytes(&x, sizeof(T)); }
  template <class T>
  void add_vec(const std::vector<T> &x) {
    add(x.size());
    if (!x.empty()) bytes(x.data(), x.size() * sizeof(T));
  }
};

// Restores the context's launch stream on every exit path.
struct LaunchStream {
  pp_ctx *c;
  LaunchStream(pp_ctx *c_, hipStream_t s) : c(c_) { c->cur = s; }
  void set(hipStream_t s) { c->cur = s; }
  ~LaunchStream() { c->cur = c->stream; }
};

// One placement pass over the resident queries, in chunks.  Stage 1 of every
// chunk runs on its own stream into one of two score buffers, so it overlaps
// stage 2 of the previous chunk; stage 2 (baseball, keep, refinement, -p,
// finalize) runs chunk after chunk on the main stream; in the 4-state form a
// phase's buckets can run concurrently on up to kBucketStreams streams (the
// tails of their work queues overlap), joined before the next phase.
int place_pass(pp_ctx *c, const pp_prefs *prefs, int precision, pp_stats *stats) {
  CHECK(prefs, PP_EINVAL, "prefs is null");
  CHECK(precision == PP_FP64 || precision == PP_FP32, PP_EINVAL, "precision must be PP_FP64 or PP_FP32");
  CHECK(prefs->keep_at_most >= 0 && prefs->keep_at_most <= 64, PP_EINVAL, "keep_at_most must be in 0..64");
  int st = set_device(c);
  if (st) return st;
  const int Q = c->Q, E = c->E, N = c->N, K = std::max(1, prefs->keep_at_most);
  c->K_last = K;
  long long t_max_pitches, t_max_strikes;
  if (prefs->max_strikes == 0) { t_max_pitches = 1LL << 62; t_max_strikes = 1LL << 62; }
  else { t_max_pitches = prefs->max_pitches; t_max_strikes = prefs->max_strikes; }
  const int cap = (int)std::min<long long>(E, std::max<long long>(1, t_max_pitches + 1));
  const int chunk = std::max(1, std::min(Q, (int)std::min<long long>(c->launch.chunk, (1LL << 31) / ((long long)E * 8 + 1))));
  // Chunk boundaries cq0[0 .. n_chunks]: the fewest chunks of at most `chunk`
  // queries, of even size.  Against full chunks and a short last one, at
  // 100,000 queries: FP32 250 -> 245 ms per pass, FP64 356 -> 359 ms; a short
  // first chunk (its stage 1 overlaps nothing) measured no faster.
  const int n_chunks = Q > 0 ? (Q + chunk - 1) / chunk : 0;
  std::vector<int> cq0(n_chunks + 1, 0);
  for (int k = 1; k <= n_chunks; k++) cq0[k] = (int)((long long)Q * k / n_chunks);
  pp_placement *d_out;
  int *d_nout, *d_status;
  ENSURE(c->b_out, (size_t)std::max(Q, 1) * K, &d_out);
  ENSURE(c->b_nout, std::max(Q, 1), &d_nout);
  ENSURE(c->b_status, std::max(Q, 1), &d_status);
  double *d_scores;
  Res *d_res, *d_kept;
  int *d_nres, *d_nkeep, *d_qlist, *d_counter;
  unsigned long long *d_stats;
  const int n_sbuf = n_chunks > 1 ? 2 : 1;
  ENSURE(c->b_scores, (size_t)n_sbuf * chunk * E, &d_scores);
  ENSURE(c->b_res, (size_t)chunk * cap, &d_res);
  ENSURE(c->b_nres, chunk, &d_nres);
  ENSURE(c->b_kept, (size_t)chunk * K, &d_kept);
  ENSURE(c->b_nkeep, chunk, &d_nkeep);
  ENSURE(c->b_counter, 128, &d_counter);
  ENSURE(c->b_stats, 6, &d_stats);
  st = ensure_stream_scratch(c);
  if (st) return st;
  HIPCHECK(hipMemsetAsync(d_stats, 0, 6 * sizeof(unsigned long long), c->stream));
  if (Q > 0) HIPCHECK(hipMemcpyAsync(d_status, c->qstat.data(), (size_t)Q * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemsetAsync(d_nout, 0, (size_t)std::max(Q, 1) * 4, c->stream));
  long long launches = 0, pitch_launches = 0;
  OptArgs marg_proto{};
  std::memset(&marg_proto, 0, sizeof(marg_proto));
  if (prefs->calc_pp) {
    st = upload_prior(c, &marg_proto);
    if (st) return st;
  }
  EventPool pool;
  LaunchStream ls(c, c->stream);
  const bool fp32 = precision == PP_FP32;
  const bool s2_fp32 = fp32 && stage2_fp32_ok(c);
  // Concurrent buckets, measured at 100,000 queries: the FP32 form 323 -> 274
  // ms per pass with 4 streams, 258 -> 251 ms with 8 for its baseball; the
  // refinement with 4 streams (FP64 364 -> 356 ms; with 50,000-query chunks
  // the FP32 form's 273-279 ms with 8 -> 267-270 ms with 4).  The FP64
  // baseball kernels stay on the main stream by default, so that each
  // launch's duration is the kernel's own (4 streams measured within 1% at
  // 50,000-query chunks), and a tuner may set them.  The 20-state kernels
  // share one scratch buffer and do not fork.
  const int pitch_streams = c->model.S == 4 ? c->launch.pitch_streams[s2_fp32 ? 1 : 0] : 0;
  const int refine_streams = c->model.S == 4 ? c->launch.refine_streams[s2_fp32 ? 1 : 0] : 0;
  hipEvent_t e_t0 = pool.get(), e_t1 = pool.get();
  CHECK(e_t0 && e_t1, PP_EDEVICE, "hipEventCreate failed");
  HIPCHECK(hipEventRecord(e_t0, c->stream));
  st = fp32 ? build_code_logs<float>(c, prefs->start_pend) : build_code_logs<double>(c, prefs->start_pend);
  if (st) return st;
  launches++;
  HIPCHECK(hipEventRecord(e_t1, c->stream));
  HIPCHECK(hipStreamWaitEvent(c->s1, e_t1, 0));  // stage 1 reads the table
  // Host-side plans of every chunk, uploaded once with synchronous copies so
  // that no asynchronous copy reads host memory: stage-1 tiles, and the
  // chunk's queries bucketed by masked length (qlist + per-bucket offsets).
  // qlist[q0 .. q0 + nq) holds chunk q0's queries by bucket, boffs[ci] the
  // bucket offsets within it
  std::vector<int> t_lo, t_hi, t_q, t_off(n_chunks + 1, 0), qlist_all(std::max(Q, 1), -1);
  std::vector<std::vector<int>> boffs(n_chunks, std::vector<int>(kBuckets + 1, 0));
  for (int ci = 0; ci < n_chunks; ci++) {
    const int q0 = cq0[ci], nq = cq0[ci + 1] - q0;
    t_off[ci] = (int)t_lo.size();
    stage1_tiles(c, q0, nq, t_lo, t_hi, t_q);
    std::vector<std::vector<int>> buckets(kBuckets);
    for (int l = 0; l < nq; l++)
      if (c->qstat[q0 + l] == PP_Q_OK) buckets[stage2_bucket(c, c->qM[q0 + l])].push_back(l);
    // the single 20-state class: longest queries first, so that the launch's
    // tail holds the short ones
    if (c->model.S == 20)
      std::stable_sort(buckets[0].begin(), buckets[0].end(),
                       [&](int x, int y) { return c->qM[q0 + x] > c->qM[q0 + y]; });
    std::vector<int> &boff = boffs[ci];
    int n = 0;
    for (int b = 0; b < kBuckets; b++) {
      boff[b] = n;
      for (int l : buckets[b]) qlist_all[(size_t)q0 + n++] = l;
    }
    boff[kBuckets] = n;
  }
  t_off[n_chunks] = (int)t_lo.size();
  int *d_tlo, *d_thi, *d_tq;
  const int n_tiles_all = std::max(1, (int)t_lo.size());
  ENSURE(c->b_tile_lo, n_tiles_all, &d_tlo);
  ENSURE(c->b_tile_hi, n_tiles_all, &d_thi);
  ENSURE(c->b_tile_q, (size_t)n_tiles_all * kTQ, &d_tq);
  uint8_t *d_tc_prealloc;
  ENSURE(c->b_tcodes, (size_t)((std::min(chunk, std::max(Q, 1)) + kTQ - 1) / kTQ) * N * kTQ, &d_tc_prealloc);
  (void)d_tc_prealloc;
  ENSURE(c->b_qlist, std::max(Q, 1), &d_qlist);
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (!t_lo.empty()) {
    HIPCHECK(hipMemcpy(d_tlo, t_lo.data(), t_lo.size() * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_thi, t_hi.data(), t_hi.size() * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_tq, t_q.data(), t_q.size() * 4, hipMemcpyHostToDevice));
  }
  if (Q > 0) HIPCHECK(hipMemcpy(d_qlist, qlist_all.data(), (size_t)Q * 4, hipMemcpyHostToDevice));
  // per-chunk intervals, read after the pass: stage 1 (s1 stream), pitch,
  // keep + refine, -p, finalize (main stream)
  struct Marks { hipEvent_t s1a, s1b, p0, p1, r1, m1, f1; };
  std::vector<Marks> marks(n_chunks);
  std::vector<hipEvent_t> scores_free(n_sbuf, nullptr);  // last pitch phase reading each buffer
  // each baseball launch's own interval, on the stream it was launched on
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pitch_ev;
  // Runs launch(b) for every non-empty bucket, round-robin over the first
  // n_streams bucket streams (none: the main stream), and joins them back
  // into the main stream.  Buckets go longest queries first: streams share
  // the device's few hardware queues, so a launch made late may wait behind
  // another on its queue, and the phase then ends on a short bucket's tail
  // rather than a long bucket's whole run.
  auto phase = [&](const std::vector<int> &boff, int n_streams, auto &&launch) -> int {
    if (n_streams <= 0) {
      ls.set(c->stream);
      for (int b = kBuckets - 1; b >= 0; b--)
        if (boff[b + 1] > boff[b]) {
          const int r = launch(b);
          if (r) return r;
        }
      return PP_OK;
    }
    hipEvent_t ef = pool.get();
    CHECK(ef, PP_EDEVICE, "hipEventCreate failed");
    HIPCHECK(hipEventRecord(ef, c->stream));
    bool used[kBucketStreams] = {};
    int k = 0;
    for (int b = kBuckets - 1; b >= 0; b--) {
      if (boff[b + 1] == boff[b]) continue;
      const int si = k++ % std::min(n_streams, kBucketStreams);
      if (!used[si]) {
        HIPCHECK(hipStreamWaitEvent(c->bs[si], ef, 0));
        used[si] = true;
      }
      ls.set(c->bs[si]);
      const int r = launch(b);
      if (r) return r;
    }
    ls.set(c->stream);
    for (int si = 0; si < kBucketStreams; si++) {
      if (!used[si]) continue;
      hipEvent_t ej = pool.get();
      CHECK(ej, PP_EDEVICE, "hipEventCreate failed");
      HIPCHECK(hipEventRecord(ej, c->bs[si]));
      HIPCHECK(hipStreamWaitEvent(c->stream, ej, 0));
    }
    return PP_OK;
  };
  // The chunk loop: launched eagerly, or (graph mode) captured once into a
  // graph that later passes with the same plan replay.  The key covers
  // everything the launches' arguments and shapes derive from.
  const bool graph = c->launch.graph != 0;
  uint64_t key = 0;
  if (graph) {
    PlanKey h;
    h.add(Q); h.add(K); h.add(precision); h.add(*prefs); h.add(c->launch); h.add(n_chunks);
    h.add_vec(cq0); h.add_vec(t_off);
    for (const auto &bo : boffs) h.add_vec(bo);
    const void *ptrs[] = {d_out, d_nout, d_status, d_scores, d_res, d_nres, d_kept, d_nkeep, d_qlist,
                          d_counter, d_stats, d_tlo, d_thi, d_tq, c->b_tcodes.p, c->b_codes.p, c->b_qM.p,
                          c->b_lt.p, c->b_se.p, c->b_d64.p, c->b_p64.p, c->b_d32.p, c->b_p32.p, c->b_de.p,
                          c->b_pe.p, c->b_de32.p, c->b_pe32.p, c->b_bl.p, c->b_sscr.p, c->d_model, c->b_cat.p,
                          c->b_prior.p, c->b_scratch.p};
    for (const void *ptr : ptrs) h.add(ptr);
    h.add(c->reg_limit); h.add(c->long_cap); h.add(c->sscr_blocks);
    key = h.v;
  }
  hipEvent_t g0 = pool.get(), g1 = pool.get();
  CHECK(g0 && g1, PP_EDEVICE, "hipEventCreate failed");
  if (graph) HIPCHECK(hipEventRecord(g0, c->stream));
  constexpr size_t kMaxGraphs = 4;
  auto cached = std::find_if(c->graphs.begin(), c->graphs.end(), [&](const auto &g) { return g.key == key; });
  if (graph && cached != c->graphs.end()) {
    HIPCHECK(hipGraphLaunch(cached->exec, c->stream));
    launches += cached->launches;
    pitch_launches += cached->pitch_launches;
  } else {
    const long long l0 = launches, pl0 = pitch_launches;
    if (graph) {
      if (c->graphs.size() >= kMaxGraphs) {
        (void)hipGraphExecDestroy(c->graphs.front().exec);
        c->graphs.erase(c->graphs.begin());
      }
      HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
      // the stage-1 stream joins the capture
      hipEvent_t ecap = pool.get();
      int cst = ecap ? PP_OK : PP_EDEVICE;
      if (!cst && hipEventRecord(ecap, c->stream) != hipSuccess) cst = PP_EDEVICE;
      if (!cst && hipStreamWaitEvent(c->s1, ecap, 0) != hipSuccess) cst = PP_EDEVICE;
      if (cst) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        set_error("graph capture could not fork the stage-1 stream");
        return cst;
      }
    }
    auto run_chunks = [&]() -> int {
    for (int ci = 0; ci < n_chunks; ci++) {
      const int q0 = cq0[ci], nq = cq0[ci + 1] - q0;
      double *scores = d_scores + (size_t)(ci % n_sbuf) * chunk * E;
      Marks &mk = marks[ci];
      for (hipEvent_t *e : {&mk.s1a, &mk.s1b, &mk.p0, &mk.p1, &mk.r1, &mk.m1, &mk.f1}) {
        *e = pool.get();
        CHECK(*e, PP_EDEVICE, "hipEventCreate failed");
      }
      // stage 1 on its stream, once the baseball phase two chunks back has
      // finished with this score buffer
      if (scores_free[ci % n_sbuf]) HIPCHECK(hipStreamWaitEvent(c->s1, scores_free[ci % n_sbuf], 0));
      ls.set(c->s1);
      if (!graph) HIPCHECK(hipEventRecord(mk.s1a, c->s1));
      const int to = t_off[ci];
      st = fp32 ? launch_stage1_tiles<float>(c, q0, nq, scores, d_tlo + to, d_thi + to, d_tq + (size_t)to * kTQ)
                : launch_stage1_tiles<double>(c, q0, nq, scores, d_tlo + to, d_thi + to, d_tq + (size_t)to * kTQ);
      if (st) return st;
      launches += 2;
      HIPCHECK(hipEventRecord(mk.s1b, c->s1));
      ls.set(c->stream);
      const std::vector<int> &boff = boffs[ci];
      int *qlist = d_qlist + q0;
      hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(128), 0, c->stream, d_counter, 128);
      HIPCHECK(hipStreamWaitEvent(c->stream, mk.s1b, 0));
      if (!graph) HIPCHECK(hipEventRecord(mk.p0, c->stream));
      const uint8_t *codes = reinterpret_cast<const uint8_t *>(c->b_codes.p) + (size_t)q0 * N;
      int *status = d_status + q0;
      st = phase(boff, pitch_streams, [&](int b) -> int {
        PitchArgs p{};
        p.model = c->d_model;
        p.codes = codes;
        p.scores = scores;
        p.qlist = qlist + boff[b];
        p.nq = boff[b + 1] - boff[b];
        p.counter = d_counter + b;
        p.darr = s2_fp32 ? c->b_d32.p : c->b_d64.p;
        p.parr = s2_fp32 ? c->b_p32.p : c->b_p64.p;
        p.darr_e = reinterpret_cast<const int *>(s2_fp32 ? c->b_de32.p : c->b_de.p);
        p.parr_e = reinterpret_cast<const int *>(s2_fp32 ? c->b_pe32.p : c->b_pe.p);
        p.bl = reinterpret_cast<const double *>(c->b_bl.p);
        p.N = N;
        p.E = E;
        p.start_pend = prefs->start_pend;
        p.max_pend = prefs->max_pend;
        p.initial_tol = prefs->initial_tolerance;
        p.strike_box = prefs->strike_box;
        p.t_max_pitches = t_max_pitches;
        p.t_max_strikes = t_max_strikes;
        p.cap = cap;
        p.res = d_res;
        p.n_res = d_nres;
        p.status = status;
        p.n_calls = d_stats;
        p.n_pitched = d_stats + 1;
        launches++;
        pitch_launches++;
        if (graph) return s2_fp32 ? dispatch_pitch<float, false>(c, p, b) : dispatch_pitch<double, true>(c, p, b);
        hipEvent_t k0 = pool.get(), k1 = pool.get();
        CHECK(k0 && k1, PP_EDEVICE, "hipEventCreate failed");
        HIPCHECK(hipEventRecord(k0, c->cur));
        const int r = s2_fp32 ? dispatch_pitch<float, false>(c, p, b) : dispatch_pitch<double, true>(c, p, b);
        HIPCHECK(hipEventRecord(k1, c->cur));
        pitch_ev.emplace_back(k0, k1);
        return r;
      });
      if (st) return st;
      HIPCHECK(hipEventRecord(mk.p1, c->stream));
      scores_free[ci % n_sbuf] = mk.p1;
      KeepArgs ka;
      ka.Q = nq;
      ka.cap = cap;
      ka.K = K;
      ka.keep = prefs->keep_at_most;
      ka.res = d_res;
      ka.n_res = d_nres;
      ka.status = status;
      ka.log_keep_factor = std::log(prefs->keep_factor);
      ka.kept = d_kept;
      ka.n_keep = d_nkeep;
      hipLaunchKernelGGL(keep_kernel, dim3((nq + 255) / 256), dim3(256), 0, c->stream, ka);
      HIPCHECK(hipGetLastError());
      st = phase(boff, refine_streams, [&](int b) -> int {
        OptArgs o{};
        std::memset(&o, 0, sizeof(o));
        o.model = c->d_model;
        o.codes = codes;
        o.N = N;
        o.mode = 1;
        o.n_jobs = (boff[b + 1] - boff[b]) * K;
        o.K = K;
        o.qlist = qlist + boff[b];
        o.counter = d_counter + 32 + b;
        o.darr = c->b_d64.p;
        o.parr = c->b_p64.p;
        o.darr_e = reinterpret_cast<const int *>(c->b_de.p);
        o.parr_e = reinterpret_cast<const int *>(c->b_pe.p);
        o.bl = reinterpret_cast<const double *>(c->b_bl.p);
        o.tol = 1e-5;  // core.ml:14 final_tolerance
        o.max_pend = prefs->max_pend;
        o.start_pend = prefs->start_pend;
        o.kept = d_kept;
        o.n_keep = d_nkeep;
        o.status = status;
        o.n_calls = d_stats;
        launches++;
        // the FP64 form sums site logs in site order; the FP32 form's FP64
        // refinement sums them per wave
        return fp32 ? dispatch_opt<double, false>(c, o, b) : dispatch_opt<double, true>(c, o, b);
      });
      if (st) return st;
      if (!graph) HIPCHECK(hipEventRecord(mk.r1, c->stream));
      if (prefs->calc_pp) {  // core.ml:378-392, one job per kept placement
        // on the main stream: forked over the refinement's streams, this phase
        // measured 196 -> 215-221 ms (FP64) at 20,000 queries
        st = phase(boff, 0, [&](int b) -> int {
          OptArgs o = marg_proto;
          o.model = c->d_model;
          o.codes = codes;
          o.N = N;
          o.mode = 1;
          o.n_jobs = (boff[b + 1] - boff[b]) * K;
          o.K = K;
          o.qlist = qlist + boff[b];
          o.counter = d_counter + 64 + b;
          o.darr = c->b_d64.p;
          o.parr = c->b_p64.p;
          o.darr_e = reinterpret_cast<const int *>(c->b_de.p);
          o.parr_e = reinterpret_cast<const int *>(c->b_pe.p);
          o.bl = reinterpret_cast<const double *>(c->b_bl.p);
          o.max_pend = prefs->max_pend;
          o.pp_rel_err = prefs->pp_rel_err;
          o.kept = d_kept;
          o.n_keep = d_nkeep;
          o.status = status;
          o.n_calls = d_stats;
          launches++;
          return fp32 ? dispatch_marg<false>(c, o, b) : dispatch_marg<true>(c, o, b);
        });
        if (st) return st;
      }
      if (!graph) HIPCHECK(hipEventRecord(mk.m1, c->stream));
      FinalArgs fa;
      fa.Q = nq;
      fa.K = K;
      fa.calc_pp = prefs->calc_pp;
      fa.kept = d_kept;
      fa.n_keep = d_nkeep;
      fa.status = status;
      fa.out = d_out + (size_t)q0 * K;
      fa.n_out = d_nout + q0;
      hipLaunchKernelGGL(finalize_kernel, dim3((nq + 255) / 256), dim3(256), 0, c->stream, fa);
      HIPCHECK(hipGetLastError());
      launches += 2;
      if (!graph) HIPCHECK(hipEventRecord(mk.f1, c->stream));
    }
      return PP_OK;
    };
    st = run_chunks();
    if (graph) {
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(c->stream, &g);
      if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
      }
      if (ec != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        set_error(std::string("hipStreamEndCapture failed: ") + hipGetErrorString(ec));
        return PP_EDEVICE;
      }
      hipGraphExec_t exec = nullptr;
      const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess || !exec) {
        set_error(std::string("hipGraphInstantiate failed: ") + hipGetErrorString(ei));
        return PP_EDEVICE;
      }
      c->graphs.push_back({key, exec, launches - l0, pitch_launches - pl0});
      HIPCHECK(hipGraphLaunch(exec, c->stream));
    } else if (st) {
      return st;
    }
  }
  if (graph) HIPCHECK(hipEventRecord(g1, c->stream));
  HIPCHECK(hipStreamSynchronize(c->s1));
  HIPCHECK(hipStreamSynchronize(c->stream));
  float t_table = 0, t_graph = 0;
  double t_s1 = 0, t_pitch = 0, t_ref = 0, t_pp = 0;
  HIPCHECK(hipEventElapsedTime(&t_table, e_t0, e_t1));
  if (graph) HIPCHECK(hipEventElapsedTime(&t_graph, g0, g1));
  for (const Marks &mk : marks) {
    if (graph) break;  // no per-phase events in graph mode
    float a1 = 0, a2 = 0, a3 = 0, a4 = 0, a5 = 0;
    HIPCHECK(hipEventElapsedTime(&a1, mk.s1a, mk.s1b));
    HIPCHECK(hipEventElapsedTime(&a2, mk.p0, mk.p1));
    HIPCHECK(hipEventElapsedTime(&a3, mk.p1, mk.r1));
    HIPCHECK(hipEventElapsedTime(&a4, mk.r1, mk.m1));
    HIPCHECK(hipEventElapsedTime(&a5, mk.m1, mk.f1));
    t_s1 += a1;
    t_pitch += a2;
    t_ref += a3 + a5;
    t_pp += a4;
  }
  double t_pitch_k = 0;
  for (const auto &kv : pitch_ev) {
    float a = 0;
    HIPCHECK(hipEventElapsedTime(&a, kv.first, kv.second));
    t_pitch_k += a;
  }
  if (stats) {
    unsigned long long hs[6];
    HIPCHECK(hipMemcpy(hs, d_stats, sizeof(hs), hipMemcpyDeviceToHost));
    stats->t_stage1_ms = t_s1;
    stats->t_table_ms = t_table;
    stats->t_pitch_ms = t_pitch;
    stats->t_refine_ms = t_ref;
    stats->t_ranking_ms = t_table + t_s1;
    stats->t_ml_ms = graph ? t_graph : t_pitch + t_ref;
    stats->n_logdots = 0;
    stats->n_site_pairs = 0;
    for (int q = 0; q < Q; q++)
      if (c->qstat[q] == PP_Q_OK) {
        stats->n_logdots += E;
        stats->n_site_pairs += (int64_t)c->qM[q] * E;
      }
    stats->n_like_calls = (int64_t)hs[0];
    stats->n_pitched = (int64_t)hs[1];
    stats->n_site_probes_pend = (int64_t)hs[2];
    stats->n_site_probes_dist = (int64_t)hs[3];
    stats->n_site_probes_pend_pitch = (int64_t)hs[4];
    stats->n_site_probes_dist_pitch = (int64_t)hs[5];
    stats->n_pitch_launches = pitch_launches;
    stats->t_pp_ms = t_pp;
    stats->t_pitch_kernel_ms = t_pitch_k;
    stats->n_launches = launches;
    stats->n_fp64_replaced = 0;
    stats->t_fp64_replace_ms = 0;
  }
  return PP_OK;
}

// Queries of an FP32 pass that failed on the prox_bl >= 0 assertion
// (three_tax.ml:41-45) are placed again in the FP64 form: FP32 likelihoods
// steer Brent along other paths, where the one-ulp drift of dist + prox can
// trip the assertion on queries the reference places; the FP64 form fails
// exactly the reference's.  idx lists them (sorted positions, ascending).
// The resident set and the pass's outputs are set aside while a pass over
// those queries alone runs, whose results are then scattered back.
int replace_in_fp64(pp_ctx *c, const pp_prefs *prefs, const std::vector<int> &idx, pp_stats *stats) {
  const int n = (int)idx.size(), N = c->N, K = c->K_last;
  int *d_idx;
  uint8_t *d_rc;
  int *d_rqM;
  ENSURE(c->b_ridx, n, &d_idx);
  ENSURE(c->b_rcodes, (size_t)n * N, &d_rc);
  ENSURE(c->b_rqM, n, &d_rqM);
  HIPCHECK(hipMemcpy(d_idx, idx.data(), (size_t)n * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(gather_codes_kernel, dim3(std::min(2048, (int)(((size_t)n * N + 255) / 256))), dim3(256), 0,
                     c->stream, reinterpret_cast<const uint8_t *>(c->b_codes.p), d_idx, n, N, d_rc);
  HIPCHECK(hipGetLastError());
  std::vector<int> sub_M(n), sub_lo(n), sub_hi(n);
  for (int j = 0; j < n; j++) {
    sub_M[j] = c->qM[idx[j]];
    sub_lo[j] = c->qlo[idx[j]];
    sub_hi[j] = c->qhi[idx[j]];
  }
  HIPCHECK(hipMemcpy(d_rqM, sub_M.data(), (size_t)n * 4, hipMemcpyHostToDevice));
  // set the resident set and outputs aside (the positions stay sorted by
  // first informative column, as stage-1 tiling expects)
  const int Q = c->Q;
  std::vector<int> order_s = std::move(c->order), qM_s = std::move(c->qM), qlo_s = std::move(c->qlo),
                   qhi_s = std::move(c->qhi), qstat_s = std::move(c->qstat);
  auto swap_bufs = [c]() {
    std::swap(c->b_codes, c->b_rcodes);
    std::swap(c->b_qM, c->b_rqM);
    std::swap(c->b_out, c->b_rout);
    std::swap(c->b_nout, c->b_rnout);
    std::swap(c->b_status, c->b_rstatus);
  };
  swap_bufs();
  c->Q = n;
  c->order.resize(n);
  std::iota(c->order.begin(), c->order.end(), 0);
  c->qM = sub_M;
  c->qlo = sub_lo;
  c->qhi = sub_hi;
  c->qstat.assign(n, PP_Q_OK);
  // A handful of queries spread over the stage-2 classes: their launches run
  // concurrently on the FP32 form's streams rather than one after another.
  const pp_launch_config launch = c->launch;
  c->launch.pitch_streams[0] = launch.pitch_streams[1];
  c->launch.refine_streams[0] = launch.refine_streams[1];
  pp_stats sub{};
  int st = place_pass(c, prefs, PP_FP64, &sub);
  c->launch = launch;
  swap_bufs();
  c->Q = Q;
  c->order = std::move(order_s);
  c->qM = std::move(qM_s);
  c->qlo = std::move(qlo_s);
  c->qhi = std::move(qhi_s);
  c->qstat = std::move(qstat_s);
  c->K_last = K;
  if (st) return st;
  hipLaunchKernelGGL(scatter_results_kernel, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                     reinterpret_cast<const pp_placement *>(c->b_rout.p), reinterpret_cast<const int *>(c->b_rnout.p),
                     reinterpret_cast<const int *>(c->b_rstatus.p), d_idx, n, K,
                     reinterpret_cast<pp_placement *>(c->b_out.p), reinterpret_cast<int *>(c->b_nout.p),
                     reinterpret_cast<int *>(c->b_status.p));
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (stats) {
    stats->n_fp64_replaced += n;
    stats->t_fp64_replace_ms += sub.t_ranking_ms + sub.t_ml_ms;
    stats->t_ml_ms += sub.t_ranking_ms + sub.t_ml_ms;
  }
  return PP_OK;
}

// A placement pass over the resident queries; in the 4-state FP32 form the
// queries failing the prox_bl assertion are then placed again in FP64.
int place_resident(pp_ctx *c, const pp_prefs *prefs, int precision, pp_stats *stats) {
  int st = place_pass(c, prefs, precision, stats);
  if (st || precision != PP_FP32 || !stage2_fp32_ok(c) || c->Q == 0) return st;
  std::vector<int> hs(c->Q), idx;
  HIPCHECK(hipMemcpy(hs.data(), c->b_status.p, (size_t)c->Q * 4, hipMemcpyDeviceToHost));
  for (int i = 0; i < c->Q; i++)
    if (hs[i] == PP_Q_ASSERT) idx.push_back(i);
  if (idx.empty()) return PP_OK;
  return replace_in_fp64(c, prefs, idx, stats);
}

int results_download(pp_ctx *c, pp_placement *out, int *n_out, int *status) {
  const int Q = c->Q, K = c->K_last;
  std::vector<pp_placement> h((size_t)Q * K);
  std::vector<int> hn(Q), hs(Q);
  if (Q) {
    HIPCHECK(hipMemcpy(h.data(), c->b_out.p, h.size() * sizeof(pp_placement), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(hn.data(), c->b_nout.p, Q * 4, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(hs.data(), c->b_status.p, Q * 4, hipMemcpyDeviceToHost));
  }
  for (int i = 0; i < Q; i++) {
    const int q = c->order[i];
    n_out[q] = hn[i];
    status[q] = hs[i];
    std::memcpy(out + (size_t)q * K, h.data() + (size_t)i * K, sizeof(pp_placement) * K);
  }
  return PP_OK;
}

// host <-> device helper for the per-call twins
struct Scratch {
  std::vector<void *> ptrs;
  ~Scratch() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <class T>
  T *up(const T *h, size_t n) {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n * sizeof(T), 8)) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    if (h && n) (void)hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice);
    return reinterpret_cast<T *>(p);
  }
};

}  // namespace

extern "C" {

int pp_ctx_create(int device, pp_ctx **out) {
  CHECK(out, PP_EINVAL, "out is null");
  int n = 0;
  HIPCHECK(hipGetDeviceCount(&n));
  CHECK(device >= 0 && device < n, PP_EINVAL, "no such GPU device");
  pp_ctx *c = new pp_ctx();
  c->device = device;
  HIPCHECK(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIPCHECK(hipGetDeviceProperties(&prop, device));
  c->n_cu = prop.multiProcessorCount;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
    c->max_lds = (size_t)lds;
  HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHECK(hipStreamCreateWithFlags(&c->s1, hipStreamNonBlocking));
  for (hipStream_t &b : c->bs) HIPCHECK(hipStreamCreateWithFlags(&b, hipStreamNonBlocking));
  c->cur = c->stream;
  *out = c;
  return PP_OK;
}

int pp_ctx_destroy(pp_ctx *c) {
  if (!c) return PP_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->s1) (void)hipStreamSynchronize(c->s1);
  for (hipStream_t b : c->bs)
    if (b) (void)hipStreamSynchronize(b);
  while (!c->glvs.empty()) pp_glv_free(*c->glvs.begin());
  for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.exec);
  DBuf *bufs[] = {&c->b_model, &c->b_bl, &c->b_d64, &c->b_p64, &c->b_d32, &c->b_p32, &c->b_de,
                  &c->b_pe, &c->b_s64, &c->b_se, &c->b_de32, &c->b_pe32, &c->b_lt, &c->b_scratch, &c->b_cat, &c->b_prior, &c->b_codes, &c->b_qM, &c->b_tcodes,
                  &c->b_tile_lo, &c->b_tile_hi, &c->b_tile_q, &c->b_ttab, &c->b_scores, &c->b_res,
                  &c->b_nres, &c->b_status, &c->b_kept, &c->b_nkeep, &c->b_out, &c->b_nout,
                  &c->b_qlist, &c->b_counter, &c->b_stats, &c->b_sscr, &c->b_rcodes, &c->b_rqM,
                  &c->b_rout, &c->b_rnout, &c->b_rstatus, &c->b_ridx};
  for (DBuf *b : bufs)
    if (b->p) (void)hipFree(b->p);
  (void)hipStreamDestroy(c->stream);
  if (c->s1) (void)hipStreamDestroy(c->s1);
  for (hipStream_t b : c->bs)
    if (b) (void)hipStreamDestroy(b);
  delete c;
  return PP_OK;
}

static int upload_model(pp_ctx *c) {
  DevModel dm;
  c->model.fill_dev(&dm);
  DevModel *d;
  ENSURE(c->b_model, 1, &d);
  HIPCHECK(hipMemcpy(d, &dm, sizeof(dm), hipMemcpyHostToDevice));
  c->d_model = d;
  c->has_tree = false;
  return PP_OK;
}

int pp_model_build(pp_ctx *c, int n_states, int n_rates, const double *exch, const double *statd,
                   const double *rates) {
  CHECK(c && exch && statd && rates, PP_EINVAL, "null argument");
  int st = set_device(c);
  if (st) return st;
  std::string err;
  st = build_model(c->model, n_states, n_rates, exch, statd, rates, err);
  if (st) { set_error(err); return st; }
  return upload_model(c);
}

int pp_model_build_cat(pp_ctx *c, int n_states, int n_cats, const double *exch, const double *statd,
                       const double *cat_rates) {
  CHECK(c && exch && statd && cat_rates, PP_EINVAL, "null argument");
  CHECK(n_cats >= 1 && n_cats <= kMaxRates, PP_EINVAL, "n_cats must be in 1..32");
  for (int k = 0; k < n_cats; k++) CHECK(cat_rates[k] >= 0., PP_EINVAL, "negative category rate");
  int st = set_device(c);
  if (st) return st;
  std::string err;
  const double one = 1.0;
  st = build_model(c->model, n_states, 1, exch, statd, &one, err);
  if (st) { set_error(err); return st; }
  c->model.n_cats = n_cats;
  c->model.cat_rates.assign(cat_rates, cat_rates + n_cats);
  return upload_model(c);
}

int pp_site_categories(pp_ctx *c, int n_sites, const int *site_cat) {
  CHECK(c && site_cat && n_sites >= 0, PP_EINVAL, "bad arguments");
  CHECK(c->model.built && c->model.n_cats > 0, PP_ESTATE, "no CAT model: call pp_model_build_cat first");
  for (int i = 0; i < n_sites; i++)
    CHECK(site_cat[i] >= 0 && site_cat[i] < c->model.n_cats, PP_EINVAL, "site category out of range");
  c->model.site_cat.assign(site_cat, site_cat + n_sites);
  c->has_tree = false;
  return PP_OK;
}

int pp_model_get(pp_ctx *c, double *x, double *l, double *xit) {
  CHECK(c && c->model.built, PP_ESTATE, "no model");
  const int S = c->model.S;
  std::memcpy(x, c->model.x.data(), S * S * 8);
  std::memcpy(l, c->model.l.data(), S * 8);
  std::memcpy(xit, c->model.xit.data(), S * S * 8);
  return PP_OK;
}

int pp_tree_set(pp_ctx *c, int n_nodes, const int *child_start, const int *children,
                const double *bl, int n_sites, const char *leaf_seqs) {
  CHECK(c && child_start && children && bl && leaf_seqs, PP_EINVAL, "null argument");
  CHECK(c->model.built, PP_ESTATE, "no model: call pp_model_build first");
  CHECK(n_nodes >= 2 && n_sites >= 1, PP_EINVAL, "need n_nodes >= 2 and n_sites >= 1");
  int st = set_device(c);
  if (st) return st;
  TreeDesc t{n_nodes, n_sites, child_start, children, bl};
  c->has_tree = false;
  c->G = n_nodes;
  c->N = n_sites;
  c->E = n_nodes - 1;
  c->Epad = (c->E + 255) / 256 * 256;
  c->bl.assign(bl, bl + n_nodes);
  // the default prior: exponential with mean = tree length / edges
  // (pplacer_run.ml:386-389)
  c->prior_kind = PP_PRIOR_FLAT_EXP;
  c->prior_means.assign(1, tree_length(t) / (double)(n_nodes - 1));
  c->prior_set = true;
  return gpu_precompute(c, t, leaf_seqs);
}

int pp_set_prior(pp_ctx *c, int kind, int n, const double *means) {
  CHECK(c, PP_EINVAL, "null context");
  CHECK(kind == PP_PRIOR_FLAT_EXP || kind == PP_PRIOR_UNIFORM || kind == PP_PRIOR_INFORMATIVE,
        PP_EINVAL, "unknown prior kind");
  if (kind != PP_PRIOR_UNIFORM) {
    CHECK(means && n >= 1, PP_EINVAL, "the exponential priors need their means");
    if (kind == PP_PRIOR_FLAT_EXP) n = 1;
    c->prior_means.assign(means, means + n);
  } else {
    c->prior_means.clear();
  }
  c->prior_kind = kind;
  c->prior_set = true;
  return PP_OK;
}

int pp_tree_upload(pp_ctx *c, int n_nodes, const double *bl, int n_sites, const double *darr,
                   const int64_t *darr_e, const double *parr, const int64_t *parr_e,
                   const double *snodes, const int64_t *snodes_e) {
  CHECK(c && bl && darr && darr_e && parr && parr_e && snodes && snodes_e, PP_EINVAL, "null argument");
  CHECK(c->model.built, PP_ESTATE, "no model: call pp_model_build first");
  CHECK(n_nodes >= 2 && n_sites >= 1, PP_EINVAL, "need n_nodes >= 2 and n_sites >= 1");
  int st = set_device(c);
  if (st) return st;
  const size_t L = (size_t)n_nodes * c->model.R * n_sites * c->model.S;
  TreeGlvs g;
  g.darr.assign(darr, darr + L);
  g.parr.assign(parr, parr + L);
  g.snodes.assign(snodes, snodes + L);
  g.darr_e.assign(darr_e, darr_e + (size_t)n_nodes * n_sites);
  g.parr_e.assign(parr_e, parr_e + (size_t)n_nodes * n_sites);
  g.snodes_e.assign(snodes_e, snodes_e + (size_t)n_nodes * n_sites);
  c->G = n_nodes;
  c->N = n_sites;
  c->E = n_nodes - 1;
  c->Epad = (c->E + 255) / 256 * 256;
  c->bl.assign(bl, bl + n_nodes);
  c->prior_set = false;  // no tree shape: the host sets the prior for -p
  return upload_glvs(c, g);
}

int pp_tree_download(pp_ctx *c, double *darr, int64_t *darr_e, double *parr, int64_t *parr_e,
                     double *snodes, int64_t *snodes_e) {
  CHECK(c && c->has_tree, PP_ESTATE, "no tree");
  int st = set_device(c);
  if (st) return st;
  const int G = c->G, N = c->N, R = c->model.R, S = c->model.S, RS = R * S, Epad = c->Epad;
  std::vector<double> a((size_t)G * N * RS), b((size_t)G * N * RS), s((size_t)N * Epad * RS);
  std::vector<int> ea((size_t)G * N), eb((size_t)G * N), es((size_t)N * Epad);
  HIPCHECK(hipMemcpy(a.data(), c->b_d64.p, a.size() * 8, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(b.data(), c->b_p64.p, b.size() * 8, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(s.data(), c->b_s64.p, s.size() * 8, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(ea.data(), c->b_de.p, ea.size() * 4, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(eb.data(), c->b_pe.p, eb.size() * 4, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(es.data(), c->b_se.p, es.size() * 4, hipMemcpyDeviceToHost));
  for (int i = 0; i < G; i++)
    for (int r = 0; r < R; r++)
      for (int n = 0; n < N; n++)
        for (int k = 0; k < S; k++) {
          const size_t dst = (((size_t)i * R + r) * N + n) * S + k;
          const size_t src = ((size_t)i * N + n) * RS + r * S + k;
          if (darr) darr[dst] = a[src];
          if (parr) parr[dst] = b[src];
          if (snodes) snodes[dst] = i < G - 1 ? s[((size_t)n * Epad + i) * RS + r * S + k] : 0.0;
        }
  for (int i = 0; i < G; i++)
    for (int n = 0; n < N; n++) {
      if (darr_e) darr_e[(size_t)i * N + n] = ea[(size_t)i * N + n];
      if (parr_e) parr_e[(size_t)i * N + n] = eb[(size_t)i * N + n];
      if (snodes_e) snodes_e[(size_t)i * N + n] = i < G - 1 ? es[(size_t)n * Epad + i] : 0;
    }
  return PP_OK;
}

int pp_queries_upload(pp_ctx *c, int n_queries, const char *queries) {
  CHECK(c && (queries || n_queries == 0), PP_EINVAL, "null argument");
  int st = set_device(c);
  if (st) return st;
  return queries_upload(c, n_queries, queries);
}

int pp_place_resident(pp_ctx *c, const pp_prefs *prefs, int precision, pp_stats *stats) {
  CHECK(c, PP_EINVAL, "null context");
  CHECK(c->has_tree, PP_ESTATE, "no tree");
  return place_resident(c, prefs, precision, stats);
}

int pp_results_download(pp_ctx *c, pp_placement *out, int *n_out, int *status) {
  CHECK(c && out && n_out && status, PP_EINVAL, "null argument");
  int st = set_device(c);
  if (st) return st;
  return results_download(c, out, n_out, status);
}

int pp_synchronize(pp_ctx *c) {
  CHECK(c, PP_EINVAL, "null context");
  HIPCHECK(hipStreamSynchronize(c->stream));
  return PP_OK;
}

int pp_launch_config_get(pp_ctx *c, pp_launch_config *out) {
  CHECK(c && out, PP_EINVAL, "null argument");
  *out = c->launch;
  return PP_OK;
}

int pp_launch_config_set(pp_ctx *c, const pp_launch_config *cfg) {
  CHECK(c && cfg, PP_EINVAL, "null argument");
  for (int f = 0; f < 2; f++) {
    CHECK(cfg->pitch_streams[f] >= 0 && cfg->pitch_streams[f] <= kBucketStreams, PP_EINVAL,
          "pitch_streams must be in 0..PP_MAX_BUCKET_STREAMS");
    CHECK(cfg->refine_streams[f] >= 0 && cfg->refine_streams[f] <= kBucketStreams, PP_EINVAL,
          "refine_streams must be in 0..PP_MAX_BUCKET_STREAMS");
  }
  CHECK(cfg->chunk >= 1 && cfg->chunk <= PP_MAX_CHUNK, PP_EINVAL, "chunk must be in 1..PP_MAX_CHUNK");
  CHECK(cfg->graph == 0 || cfg->graph == 1, PP_EINVAL, "graph must be 0 or 1");
  c->launch = *cfg;
  return PP_OK;
}

int pp_place_batch(pp_ctx *c, int n_queries, const char *queries, const pp_prefs *prefs,
                   int precision, pp_placement *out, int *n_out, int *status, pp_stats *stats) {
  CHECK(c && prefs && (queries || n_queries == 0) && out && n_out && status, PP_EINVAL, "null argument");
  int st = pp_queries_upload(c, n_queries, queries);
  if (st) return st;
  st = pp_place_resident(c, prefs, precision, stats);
  if (st) return st;
  st = pp_results_download(c, out, n_out, status);
  if (st) return st;
  for (int q = 0; q < n_queries; q++)
    if (status[q] != PP_Q_OK) {
      set_error("query " + std::to_string(q) + " failed with status " + std::to_string(status[q]));
      return PP_EQUERY;
    }
  return PP_OK;
}

int pp_stage1_scores(pp_ctx *c, int n_queries, const char *queries, double start_pend,
                     int precision, double *scores) {
  CHECK(c && scores, PP_EINVAL, "null argument");
  CHECK(precision == PP_FP64 || precision == PP_FP32, PP_EINVAL, "bad precision");
  int st = pp_queries_upload(c, n_queries, queries);
  if (st) return st;
  const int E = c->E;
  double *d;
  ENSURE(c->b_scores, (size_t)std::max(n_queries, 1) * E, &d);
  st = precision == PP_FP32 ? build_code_logs<float>(c, start_pend) : build_code_logs<double>(c, start_pend);
  if (st) return st;
  st = precision == PP_FP32 ? launch_stage1<float>(c, 0, n_queries, d) : launch_stage1<double>(c, 0, n_queries, d);
  if (st) return st;
  std::vector<double> h((size_t)n_queries * E);
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipMemcpy(h.data(), d, h.size() * 8, hipMemcpyDeviceToHost));
  for (int i = 0; i < n_queries; i++) {
    const int q = c->order[i];
    if (c->qstat[i] != PP_Q_OK) {
      for (int e = 0; e < E; e++) scores[(size_t)q * E + e] = NAN;
    } else {
      std::memcpy(scores + (size_t)q * E, h.data() + (size_t)i * E, (size_t)E * 8);
    }
  }
  return PP_OK;
}

int pp_rank_topk(pp_ctx *c, int n_queries, int n_edges, const double *scores, int k, int *ranked) {
  CHECK(c && scores && ranked, PP_EINVAL, "null argument");
  CHECK(n_edges >= 1 && k >= 1 && k <= n_edges, PP_EINVAL, "need 1 <= k <= n_edges");
  int st = set_device(c);
  if (st) return st;
  Scratch s;
  double *d_s = s.up(scores, (size_t)n_queries * n_edges);
  int *d_r = s.up<int>(nullptr, (size_t)n_queries * k);
  CHECK(d_s && d_r, PP_ENOMEM, "device allocation failed");
  const size_t lds = 4 * ((n_edges + 31) / 32);
  hipLaunchKernelGGL(topk_kernel, dim3(std::max(n_queries, 1)), dim3(256), lds, c->stream, d_s, n_queries,
                     n_edges, k, d_r);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipMemcpy(ranked, d_r, (size_t)n_queries * k * 4, hipMemcpyDeviceToHost));
  return PP_OK;
}

int pp_optimize_batch(pp_ctx *c, int n_jobs, const char *queries, const int *query_of_job,
                      const int *locs, const double *pend0, const double *dist0, double tolerance,
                      double max_pend, int precision, double *ll, double *pend, double *dist,
                      int *n_calls, int *status) {
  CHECK(c && queries && query_of_job && locs && pend0 && dist0 && ll && pend && dist && n_calls && status,
        PP_EINVAL, "null argument");
  CHECK(precision == PP_FP64 || precision == PP_FP32, PP_EINVAL, "bad precision");
  int n_queries = 0;
  for (int j = 0; j < n_jobs; j++) n_queries = std::max(n_queries, query_of_job[j] + 1);
  int st = pp_queries_upload(c, n_queries, queries);
  if (st) return st;
  for (int j = 0; j < n_jobs; j++) CHECK(locs[j] >= 0 && locs[j] < c->E, PP_EINVAL, "loc out of range");
  std::vector<int> pos(n_queries);
  for (int i = 0; i < n_queries; i++) pos[c->order[i]] = i;
  std::vector<int> jq(n_jobs);
  for (int j = 0; j < n_jobs; j++) jq[j] = pos[query_of_job[j]];
  Scratch s;
  int *d_jq = s.up(jq.data(), n_jobs), *d_loc = s.up(locs, n_jobs);
  double *d_p0 = s.up(pend0, n_jobs), *d_d0 = s.up(dist0, n_jobs);
  double *d_ll = s.up<double>(nullptr, n_jobs), *d_pe = s.up<double>(nullptr, n_jobs),
         *d_di = s.up<double>(nullptr, n_jobs);
  int *d_calls = s.up<int>(nullptr, n_jobs), *d_st = s.up<int>(nullptr, n_jobs);
  unsigned long long *d_stats = s.up<unsigned long long>(nullptr, 4);
  int *d_counter = s.up<int>(nullptr, 32);
  CHECK(d_stats && d_counter, PP_ENOMEM, "device allocation failed");
  std::vector<int> init_st(n_jobs, PP_Q_OK);
  for (int j = 0; j < n_jobs; j++)
    if (c->qstat[jq[j]] != PP_Q_OK) init_st[j] = 100 + c->qstat[jq[j]];
  HIPCHECK(hipMemcpy(d_st, init_st.data(), n_jobs * 4, hipMemcpyHostToDevice));
  HIPCHECK(hipMemset(d_counter, 0, 128));
  HIPCHECK(hipMemset(d_stats, 0, 32));
  std::vector<std::vector<int>> buckets(kBuckets);
  for (int j = 0; j < n_jobs; j++)
    if (init_st[j] == PP_Q_OK) buckets[stage2_bucket(c, c->qM[jq[j]])].push_back(j);
  std::vector<int> flat;
  std::vector<int> boff(kBuckets + 1, 0);
  for (int b = 0; b < kBuckets; b++) {
    boff[b] = (int)flat.size();
    flat.insert(flat.end(), buckets[b].begin(), buckets[b].end());
  }
  boff[kBuckets] = (int)flat.size();
  int *d_list = s.up(flat.data(), flat.size());
  st = ensure_stream_scratch(c);
  if (st) return st;
  for (int b = 0; b < kBuckets; b++) {
    const int n = boff[b + 1] - boff[b];
    if (!n) continue;
    OptArgs o{};
    std::memset(&o, 0, sizeof(o));
    o.model = c->d_model;
    o.codes = reinterpret_cast<const uint8_t *>(c->b_codes.p);
    o.N = c->N;
    o.mode = 0;
    o.n_jobs = n;
    o.K = 1;
    o.qlist = d_list + boff[b];
    o.counter = d_counter + b;
    const bool fp32 = precision == PP_FP32 && stage2_fp32_ok(c);
    o.darr = fp32 ? c->b_d32.p : c->b_d64.p;
    o.parr = fp32 ? c->b_p32.p : c->b_p64.p;
    o.darr_e = reinterpret_cast<const int *>(fp32 ? c->b_de32.p : c->b_de.p);
    o.parr_e = reinterpret_cast<const int *>(fp32 ? c->b_pe32.p : c->b_pe.p);
    o.bl = reinterpret_cast<const double *>(c->b_bl.p);
    o.tol = tolerance;
    o.max_pend = max_pend;
    o.start_pend = 0.1;
    o.job_q = d_jq;
    o.job_loc = d_loc;
    o.job_pend = d_p0;
    o.job_dist = d_d0;
    o.out_ll = d_ll;
    o.out_pend = d_pe;
    o.out_dist = d_di;
    o.out_calls = d_calls;
    o.out_status = d_st;
    o.n_calls = d_stats;
    st = fp32 ? dispatch_opt<float, false>(c, o, b) : dispatch_opt<double, true>(c, o, b);
    if (st) return st;
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipMemcpy(ll, d_ll, n_jobs * 8, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(pend, d_pe, n_jobs * 8, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(dist, d_di, n_jobs * 8, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(n_calls, d_calls, n_jobs * 4, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(status, d_st, n_jobs * 4, hipMemcpyDeviceToHost));
  return PP_OK;
}

int pp_marginal_batch(pp_ctx *c, int n_jobs, const char *queries, const int *query_of_job,
                      const int *locs, const double *pend, const double *dist, double rel_err,
                      double max_pend, double *marg, int *n_calls, int *status) {
  CHECK(c && queries && query_of_job && locs && pend && dist && marg && n_calls && status,
        PP_EINVAL, "null argument");
  int n_queries = 0;
  for (int j = 0; j < n_jobs; j++) n_queries = std::max(n_queries, query_of_job[j] + 1);
  int st = pp_queries_upload(c, n_queries, queries);
  if (st) return st;
  for (int j = 0; j < n_jobs; j++) CHECK(locs[j] >= 0 && locs[j] < c->E, PP_EINVAL, "loc out of range");
  OptArgs proto{};
  std::memset(&proto, 0, sizeof(proto));
  st = upload_prior(c, &proto);
  if (st) return st;
  std::vector<int> pos(n_queries);
  for (int i = 0; i < n_queries; i++) pos[c->order[i]] = i;
  std::vector<int> jq(n_jobs);
  for (int j = 0; j < n_jobs; j++) jq[j] = pos[query_of_job[j]];
  Scratch s;
  int *d_jq = s.up(jq.data(), n_jobs), *d_loc = s.up(locs, n_jobs);
  double *d_p0 = s.up(pend, n_jobs), *d_d0 = s.up(dist, n_jobs);
  double *d_m = s.up<double>(nullptr, n_jobs);
  int *d_calls = s.up<int>(nullptr, n_jobs), *d_st = s.up<int>(nullptr, n_jobs);
  unsigned long long *d_stats = s.up<unsigned long long>(nullptr, 4);
  int *d_counter = s.up<int>(nullptr, 32);
  CHECK(d_jq && d_loc && d_p0 && d_d0 && d_m && d_calls && d_st && d_stats && d_counter, PP_ENOMEM,
        "device allocation failed");
  std::vector<int> init_st(n_jobs, PP_Q_OK);
  for (int j = 0; j < n_jobs; j++)
    if (c->qstat[jq[j]] != PP_Q_OK) init_st[j] = 100 + c->qstat[jq[j]];
  HIPCHECK(hipMemcpy(d_st, init_st.data(), n_jobs * 4, hipMemcpyHostToDevice));
  HIPCHECK(hipMemset(d_counter, 0, 128));
  HIPCHECK(hipMemset(d_stats, 0, 32));
  std::vector<std::vector<int>> buckets(kBuckets);
  for (int j = 0; j < n_jobs; j++)
    if (init_st[j] == PP_Q_OK) buckets[stage2_bucket(c, c->qM[jq[j]])].push_back(j);
  std::vector<int> flat;
  std::vector<int> boff(kBuckets + 1, 0);
  for (int b = 0; b < kBuckets; b++) {
    boff[b] = (int)flat.size();
    flat.insert(flat.end(), buckets[b].begin(), buckets[b].end());
  }
  boff[kBuckets] = (int)flat.size();
  int *d_list = s.up(flat.data(), flat.size());
  st = ensure_stream_scratch(c);
  if (st) return st;
  for (int b = 0; b < kBuckets; b++) {
    const int n = boff[b + 1] - boff[b];
    if (!n) continue;
    OptArgs o = proto;
    o.model = c->d_model;
    o.codes = reinterpret_cast<const uint8_t *>(c->b_codes.p);
    o.N = c->N;
    o.mode = 0;
    o.n_jobs = n;
    o.K = 1;
    o.qlist = d_list + boff[b];
    o.counter = d_counter + b;
    o.darr = c->b_d64.p;
    o.parr = c->b_p64.p;
    o.darr_e = reinterpret_cast<const int *>(c->b_de.p);
    o.parr_e = reinterpret_cast<const int *>(c->b_pe.p);
    o.bl = reinterpret_cast<const double *>(c->b_bl.p);
    o.max_pend = max_pend;
    o.pp_rel_err = rel_err;
    o.job_q = d_jq;
    o.job_loc = d_loc;
    o.job_pend = d_p0;
    o.job_dist = d_d0;
    o.out_ll = d_m;
    o.out_calls = d_calls;
    o.out_status = d_st;
    o.n_calls = d_stats;
    st = dispatch_marg<true>(c, o, b);
    if (st) return st;
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipMemcpy(marg, d_m, n_jobs * 8, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(n_calls, d_calls, n_jobs * 4, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(status, d_st, n_jobs * 4, hipMemcpyDeviceToHost));
  return PP_OK;
}

// ---- linear_c.c twins ------------------------------------------------------

int pp_gemmish(double *dst, const double *a, const double *b, int n_sites, int n_states) {
  CHECK(dst && a && b && n_sites >= 0 && n_states > 0, PP_EINVAL, "bad arguments");
  Scratch s;
  const size_t n = (size_t)n_sites * n_states;
  double *da = s.up(a, (size_t)n_states * n_states), *db = s.up(b, n), *dd = s.up<double>(nullptr, n);
  CHECK(da && db && dd, PP_ENOMEM, "device allocation failed");
  if (n) hipLaunchKernelGGL(gemmish_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, dd, da, db, n_sites, n_states);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpy(dst, dd, n * 8, hipMemcpyDeviceToHost));
  return PP_OK;
}

int pp_dediagonalize(double *dst, const double *u, const double *lambda, const double *uit, int n) {
  CHECK(dst && u && lambda && uit && n > 0, PP_EINVAL, "bad arguments");
  Scratch s;
  double *du = s.up(u, (size_t)n * n), *dl = s.up(lambda, n), *dv = s.up(uit, (size_t)n * n),
         *dd = s.up<double>(nullptr, (size_t)n * n);
  CHECK(du && dl && dv && dd, PP_ENOMEM, "device allocation failed");
  hipLaunchKernelGGL(dediag_kernel, dim3((n * n + 255) / 256), dim3(256), 0, 0, dd, du, dl, dv, n);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpy(dst, dd, (size_t)n * n * 8, hipMemcpyDeviceToHost));
  return PP_OK;
}

static int twin_prod(double *dst, const double *statd, const double *x, const double *y, int R, int N,
                     int S) {
  CHECK(dst && x && y && R > 0 && N >= 0 && S > 0, PP_EINVAL, "bad arguments");
  Scratch s;
  const size_t n = (size_t)R * N * S;
  double *dx = s.up(x, n), *dy = s.up(y, n), *dd = s.up<double>(nullptr, n);
  double *dst_d = statd ? s.up(statd, S) : nullptr;
  CHECK(dx && dy && dd, PP_ENOMEM, "device allocation failed");
  if (n) hipLaunchKernelGGL(pairwise_kernel, dim3(std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0, 0, dd,
                            dx, dy, (long)n, dst_d, S);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpy(dst, dd, n * 8, hipMemcpyDeviceToHost));
  return PP_OK;
}

int pp_ten_pairwise_prod(double *dst, const double *x, const double *y, int n_rates, int n_sites,
                         int n_states) {
  return twin_prod(dst, nullptr, x, y, n_rates, n_sites, n_states);
}

int pp_ten_statd_pairwise_prod(const double *statd, double *dst, const double *a, const double *b,
                               int n_rates, int n_sites, int n_states) {
  CHECK(statd, PP_EINVAL, "statd is null");
  return twin_prod(dst, statd, a, b, n_rates, n_sites, n_states);
}

static int twin_logdot(const double *statd, const double *x, const double *y, const double *z,
                       const uint16_t *mask, int first, int last, double *util, int R, int N, int S,
                       double *out, int rate_term = 1) {
  CHECK(x && y && out && R > 0 && N >= 0 && S > 0, PP_EINVAL, "bad arguments");
  CHECK(first >= 0 && last < N && first <= last + 1, PP_EINVAL, "bad site interval");
  const int n_used = last - first + 1;
  Scratch s;
  const size_t n = (size_t)R * N * S;
  double *dx = s.up(x, n), *dy = s.up(y, n), *dz = z ? s.up(z, n) : nullptr;
  double *ds = statd ? s.up(statd, S) : nullptr;
  uint16_t *dm = mask ? s.up(mask, N) : nullptr;
  double *du = s.up<double>(nullptr, std::max(n_used, 1)), *dout = s.up<double>(nullptr, 1);
  CHECK(dx && dy && du && dout, PP_ENOMEM, "device allocation failed");
  if (n_used > 0)
    hipLaunchKernelGGL(util_kernel, dim3((n_used + 255) / 256), dim3(256), 0, 0, du, ds, dx, dy, dz, dm, first,
                       n_used, R, N, S);
  hipLaunchKernelGGL(logsum_kernel, dim3(1), dim3(64), 0, 0, du, dm, first, n_used, R, rate_term, dout);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpy(out, dout, 8, hipMemcpyDeviceToHost));
  if (util && n_used > 0) HIPCHECK(hipMemcpy(util, du, (size_t)n_used * 8, hipMemcpyDeviceToHost));
  return PP_OK;
}

int pp_ten_log_like3(const double *statd, const double *x, const double *y, const double *z,
                     double *util, int n_rates, int n_sites, int n_states, double *out) {
  CHECK(statd && z, PP_EINVAL, "null argument");
  return twin_logdot(statd, x, y, z, nullptr, 0, n_sites - 1, util, n_rates, n_sites, n_states, out);
}

int pp_ten_masked_logdot(const double *x, const double *y, const uint16_t *mask, double *util,
                         int n_rates, int n_sites, int n_states, double *out) {
  CHECK(mask, PP_EINVAL, "mask is null");
  return twin_logdot(nullptr, x, y, nullptr, mask, 0, n_sites - 1, util, n_rates, n_sites, n_states, out);
}

int pp_ten_bounded_logdot(const double *x, const double *y, int first, int last, double *util,
                          int n_rates, int n_sites, int n_states, double *out) {
  return twin_logdot(nullptr, x, y, nullptr, nullptr, first, last, util, n_rates, n_sites, n_states, out);
}

// Test hook, not part of the drop-in ABI: the FP64 engine's wave-parallel
// sequential sum (wave_sequential_sum) over x[0..n), copied into LDS, beside
// the one-lane chain.  out[0] = the wave's sum, out[1] = the chain's, out[2] =
// the number of events (-1 where the wave fell back to the chain).
static __global__ void seqsum_test_kernel(const double *x, int n, double *out) {
  extern __shared__ __align__(16) double sm[];
  double *v = sm, *acc = sm + n, *ev = sm + 2 * (size_t)n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v[i] = x[i];
  __syncthreads();
  int n_events = 0;
  const double s = wave_sequential_sum(v, n, acc, ev, &n_events);
  const double c = threadIdx.x == 0 ? ordered_sum(v, n) : 0.0;
  if (threadIdx.x == 0) {
    out[0] = s;
    out[1] = c;
    out[2] = n_events;
  }
}

int pp_test_sequential_sum(const double *x, int n, double *out) {
  CHECK(out && n >= 0 && (x || n == 0), PP_EINVAL, "bad arguments");
  int dev = 0, max_lds = 65536;
  HIPCHECK(hipGetDevice(&dev));
  if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) max_lds = 65536;
  const size_t lds = sizeof(double) * (2 * (size_t)n + 2 * kSeqSumEvents);
  CHECK(lds <= (size_t)max_lds, PP_EINVAL, "n exceeds one workgroup's LDS");
  Scratch s;
  double *dx = s.up(x, std::max(n, 1)), *dout = s.up<double>(nullptr, 3);
  CHECK(dx && dout, PP_ENOMEM, "device allocation failed");
  hipLaunchKernelGGL(seqsum_test_kernel, dim3(1), dim3(64), lds, 0, dx, n, dout);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpy(out, dout, 3 * sizeof(double), hipMemcpyDeviceToHost));
  return PP_OK;
}

// Test hook, not part of the drop-in ABI: the correctly rounded exp and log
// (pp_crmath.h) on the device, out_exp[i] = exp(x[i]), out_log[i] = log(x[i]).
static __global__ void cr_math_test_kernel(const double *x, long long n, double *oe, double *ol) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    oe[i] = cr::exp(x[i]);
    ol[i] = cr::log(x[i]);
  }
}

int pp_test_cr_math(long long n, const double *x, double *out_exp, double *out_log) {
  CHECK(n >= 0 && (n == 0 || (x && out_exp && out_log)), PP_EINVAL, "bad arguments");
  if (n == 0) return PP_OK;
  Scratch s;
  double *dx = s.up(x, (size_t)n), *de = s.up<double>(nullptr, (size_t)n), *dl = s.up<double>(nullptr, (size_t)n);
  CHECK(dx && de && dl, PP_ENOMEM, "device allocation failed");
  const int grid = (int)std::min<long long>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(cr_math_test_kernel, dim3(grid), dim3(256), 0, 0, dx, n, de, dl);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpy(out_exp, de, (size_t)n * 8, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(out_log, dl, (size_t)n * 8, hipMemcpyDeviceToHost));
  return PP_OK;
}

// The mat_* stubs: the rate-free [site][state] GLVs of Gcat_model; the same
// kernels with one rate and no n log R term.
int pp_mat_log_like3(const double *statd, const double *x, const double *y, const double *z, int n_sites,
                     int n_states, double *out) {
  CHECK(statd && z, PP_EINVAL, "null argument");
  return twin_logdot(statd, x, y, z, nullptr, 0, n_sites - 1, nullptr, 1, n_sites, n_states, out, 0);
}

int pp_mat_pairwise_prod(double *dst, const double *x, const double *y, int n_sites, int n_states) {
  return twin_prod(dst, nullptr, x, y, 1, n_sites, n_states);
}

int pp_mat_statd_pairwise_prod(const double *statd, double *dst, const double *a, const double *b,
                               int n_sites, int n_states) {
  CHECK(statd, PP_EINVAL, "statd is null");
  return twin_prod(dst, statd, a, b, 1, n_sites, n_states);
}

int pp_mat_masked_logdot(const double *x, const double *y, const uint16_t *mask, int n_sites, int n_states,
                         double *out) {
  CHECK(mask, PP_EINVAL, "mask is null");
  return twin_logdot(nullptr, x, y, nullptr, mask, 0, n_sites - 1, nullptr, 1, n_sites, n_states, out, 0);
}

int pp_mat_bounded_logdot(const double *x, const double *y, int first, int last, int n_sites, int n_states,
                          double *out) {
  return twin_logdot(nullptr, x, y, nullptr, nullptr, first, last, nullptr, 1, n_sites, n_states, out, 0);
}

}  // extern "C"

#include "pp_glvm.hip"
