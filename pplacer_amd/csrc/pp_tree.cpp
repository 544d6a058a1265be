// Reference-tree precompute on the host, FP64, in pplacer's GLV layout
// [G][R][N][S] + int64 exponents [G][N]:
//   Like_stree.calc_distal_and_proximal   like_stree.ml:43-110
//   Glv.perhaps_pull_exponent (-50, -10)  gmix_model.ml:163-187, pplacer_run.ml:331-336
//   Glv_arr.prep_supernodes (bl/2)        glv_arr.ml:46-56, pplacer_run.ml:337-341
// Nodes of one post-order level are independent, so each level is split
// over host threads.
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "pp_tree.h"

namespace pp {

namespace {

struct Glv {
  double *a;
  int64_t *e;
};

struct Ctx {
  const HostModel &m;
  int G, N;
  long L;
};

void evolve(const HostModel &m, double *tensor, Glv dst, Glv src, int N, double bl) {
  const int R = m.R, S = m.S;
  std::memcpy(dst.e, src.e, sizeof(int64_t) * N);
  multi_exp(m, bl, tensor);
  for (int r = 0; r < R; r++) {
    const double *a = tensor + r * S * S;
    const double *b = src.a + (long)r * N * S;
    double *d = dst.a + (long)r * N * S;
    for (int site = 0; site < N; site++) {  // linear_c.c:29-90
      for (int i = 0; i < S; i++) {
        double acc = 0;
        for (int j = 0; j < S; j++) acc += a[i * S + j] * b[j];
        d[i] = acc;
      }
      b += S;
      d += S;
    }
  }
}

void pairwise(Glv dst, Glv x, Glv y, long L, int N) {  // gmix_model.ml:215-218
  for (int i = 0; i < N; i++) dst.e[i] = x.e[i] + y.e[i];
  for (long i = 0; i < L; i++) dst.a[i] = x.a[i] * y.a[i];
}

void listwise(Glv dst, const std::vector<Glv> &xs, long L, int N) {  // gmix_model.ml:224-233
  if (xs.size() >= 2) {
    pairwise(dst, xs[0], xs[1], L, N);
    for (size_t k = 2; k < xs.size(); k++) pairwise(dst, dst, xs[k], L, N);
  } else if (xs.size() == 1) {
    std::memcpy(dst.e, xs[0].e, sizeof(int64_t) * N);
    std::memcpy(dst.a, xs[0].a, sizeof(double) * L);
  }
}

void pull_exponent(int min_twoexp, Glv g, int R, int N, int S) {  // gmix_model.ml:163-187
  for (int site = 0; site < N; site++) {
    int max_twoexp = -2147483647;
    for (int r = 0; r < R; r++) {
      const double *s = g.a + ((long)r * N + site) * S;
      double mx = s[0];
      for (int k = 1; k < S; k++)
        if (s[k] > mx) mx = s[k];
      int te;
      (void)std::frexp(mx, &te);
      if (te > max_twoexp) max_twoexp = te;
    }
    if (max_twoexp < min_twoexp) {
      double f = std::ldexp(0.5, -max_twoexp + 1);
      for (int r = 0; r < R; r++) {
        double *s = g.a + ((long)r * N + site) * S;
        for (int k = 0; k < S; k++) s[k] *= f;
      }
      g.e[site] += max_twoexp;
    }
  }
}

template <class F>
void parallel_for(int n, F f) {
  unsigned hw = std::thread::hardware_concurrency();
  int nt = (int)std::min<unsigned>(hw ? hw : 1, 16);
  if (n < 4 * nt) nt = 1;
  if (nt <= 1) {
    for (int i = 0; i < n; i++) f(i);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      for (int i = t; i < n; i += nt) f(i);
    });
  for (auto &x : th) x.join();
}

}  // namespace

int host_precompute(const HostModel &m, const TreeDesc &t, const char *leaf_seqs,
                    TreeGlvs &out, std::string &err) {
  const int R = m.R, S = m.S, G = t.n_nodes, N = t.n_sites;
  const long L = (long)R * N * S;
  out.darr.assign((size_t)G * L, 0.);
  out.parr.assign((size_t)G * L, 0.);
  out.snodes.assign((size_t)G * L, 0.);
  out.darr_e.assign((size_t)G * N, 0);
  out.parr_e.assign((size_t)G * N, 0);
  out.snodes_e.assign((size_t)G * N, 0);
  auto at = [&](std::vector<double> &a, std::vector<int64_t> &e, int i) {
    return Glv{a.data() + (size_t)i * L, e.data() + (size_t)i * N};
  };
  // heights: leaves 0, a node 1 + max over children (children have smaller ids)
  std::vector<int> height(G, 0), depth(G, 0), parent(G, -1);
  int max_h = 0;
  for (int id = 0; id < G; id++) {
    for (int k = t.child_start[id]; k < t.child_start[id + 1]; k++) {
      int c = t.children[k];
      if (c < 0 || c >= id) { err = "node ids must be post-order (children before parents)"; return PP_EINVAL; }
      parent[c] = id;
      height[id] = std::max(height[id], height[c] + 1);
    }
    max_h = std::max(max_h, height[id]);
  }
  // leaf encoding errors
  for (int id = 0; id < G; id++) {
    if (t.child_start[id + 1] != t.child_start[id]) continue;
    const char *s = leaf_seqs + (size_t)id * N;
    for (int k = 0; k < N; k++)
      if (m.code_of[(unsigned char)s[k]] < 0) {
        err = std::string(1, s[k]) + " not a known base in leaf " + std::to_string(id);
        return PP_EINVAL;
      }
  }
  // post-order (like_stree.ml:43-69); evolv_dist goes to snodes as scratch
  std::vector<std::vector<int>> by_h(max_h + 1);
  for (int id = 0; id < G; id++) by_h[height[id]].push_back(id);
  for (int h = 0; h <= max_h; h++) {
    auto &ids = by_h[h];
    parallel_for((int)ids.size(), [&](int k) {
      int id = ids[k];
      std::vector<double> tensor(R * S * S);
      Glv distal = at(out.darr, out.darr_e, id);
      int c0 = t.child_start[id], c1 = t.child_start[id + 1];
      if (c1 > c0) {
        std::vector<Glv> xs;
        for (int j = c0; j < c1; j++) xs.push_back(at(out.snodes, out.snodes_e, t.children[j]));
        listwise(distal, xs, L, N);
        pull_exponent(-50, distal, R, N, S);  // like_stree.ml:8
      } else {
        const char *s = leaf_seqs + (size_t)id * N;  // prep_constant_rate_glv_from_lv_arr
        for (int site = 0; site < N; site++) {
          const double *lv = &m.lv[m.code_of[(unsigned char)s[site]] * S];
          for (int r = 0; r < R; r++) std::memcpy(distal.a + ((long)r * N + site) * S, lv, sizeof(double) * S);
          distal.e[site] = 0;
        }
      }
      evolve(m, tensor.data(), at(out.snodes, out.snodes_e, id), distal, N, t.bl[id]);
    });
  }
  // pre-order (like_stree.ml:74-100): proximal of the root is the unit GLV
  {
    Glv top = at(out.parr, out.parr_e, G - 1);
    for (long i = 0; i < L; i++) top.a[i] = 1.;
    for (int i = 0; i < N; i++) top.e[i] = 0;
  }
  std::vector<std::vector<int>> by_d;
  for (int id = G - 1; id >= 0; id--) {
    depth[id] = parent[id] < 0 ? 0 : depth[parent[id]] + 1;
    if ((int)by_d.size() <= depth[id]) by_d.resize(depth[id] + 1);
    if (t.child_start[id + 1] > t.child_start[id]) by_d[depth[id]].push_back(id);
  }
  for (auto &ids : by_d) {
    parallel_for((int)ids.size(), [&](int k) {
      int id = ids[k];
      std::vector<double> tensor(R * S * S);
      std::vector<double> ea(L);
      std::vector<int64_t> ee(N);
      Glv evolved{ea.data(), ee.data()};
      evolve(m, tensor.data(), evolved, at(out.parr, out.parr_e, id), N, t.bl[id]);
      int c0 = t.child_start[id], c1 = t.child_start[id + 1];
      for (int j = c0; j < c1; j++) {  // ListFuns.pull_each_out keeps the rest in order
        std::vector<Glv> xs{evolved};
        for (int o = c0; o < c1; o++)
          if (o != j) xs.push_back(at(out.snodes, out.snodes_e, t.children[o]));
        Glv below = at(out.parr, out.parr_e, t.children[j]);
        listwise(below, xs, L, N);
        pull_exponent(-50, below, R, N, S);
      }
    });
  }
  // pull exponents at -10, then supernodes (pplacer_run.ml:331-341)
  parallel_for(G, [&](int id) {
    pull_exponent(-10, at(out.darr, out.darr_e, id), R, N, S);
    pull_exponent(-10, at(out.parr, out.parr_e, id), R, N, S);
  });
  parallel_for(G, [&](int id) {
    std::vector<double> tensor(R * S * S), ud(L), up(L);
    std::vector<int64_t> ude(N), upe(N);
    double half = t.bl[id] / 2.;
    evolve(m, tensor.data(), Glv{ud.data(), ude.data()}, at(out.darr, out.darr_e, id), N, half);
    evolve(m, tensor.data(), Glv{up.data(), upe.data()}, at(out.parr, out.parr_e, id), N, half);
    Glv sn = at(out.snodes, out.snodes_e, id);
    for (int i = 0; i < N; i++) sn.e[i] = ude[i] + upe[i];
    const double *a = ud.data(), *b = up.data();
    double *d = sn.a;
    for (int r = 0; r < R; r++)  // linear_c.c:443-463
      for (int site = 0; site < N; site++)
        for (int s = 0; s < S; s++) *d++ = m.statd[s] * (*a++) * (*b++);
  });
  return PP_OK;
}

}  // namespace pp
