// Host-side parts of the placement path that run once per model / tree:
// model building (gmix_model.ml:25-37, diagd.ml:123-172, gamma.ml:52-70),
// tip encoding tables (nuc_models.ml:24-45, prot_models.ml:5-49), empirical
// frequencies (alignment.ml:186-217) and pre-masking (pplacer_run.ml:59-144).
#include <cctype>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pp_crmath.h"
#include "pp_internal.h"
#include "pp_tree.h"

namespace pp {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

// ---- incomplete gamma and its inverse (what gamma.ml takes from GSL) -----
static double gamma_inc_P(double a, double x) {
  if (x <= 0.0) return 0.0;
  double lpre = -x + a * std::log(x) - std::lgamma(a);
  if (x < a + 1.0) {
    double ap = a, sum = 1.0 / a, del = sum;
    for (int n = 0; n < 100000; n++) {
      ap += 1.0;
      del *= x / ap;
      sum += del;
      if (std::fabs(del) < std::fabs(sum) * 1e-17) break;
    }
    return sum * std::exp(lpre);
  }
  const double tiny = 1e-300;
  double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int i = 1; i < 100000; i++) {
    double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (std::fabs(d) < tiny) d = tiny;
    c = b + an / c;
    if (std::fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    double del = d * c;
    h *= del;
    if (std::fabs(del - 1.0) < 1e-17) break;
  }
  return 1.0 - std::exp(lpre) * h;
}

static double gamma_Pinv(double a, double p) {
  if (p <= 0.0) return 0.0;
  if (p >= 1.0) return INFINITY;
  double lo = 0.0, hi = a > 1.0 ? a : 1.0;
  while (gamma_inc_P(a, hi) < p) { lo = hi; hi *= 2.0; }
  for (int it = 0; it < 400; it++) {
    double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (gamma_inc_P(a, mid) < p) lo = mid; else hi = mid;
  }
  double x = 0.5 * (lo + hi);
  for (int it = 0; it < 8; it++) {
    double fx = gamma_inc_P(a, x) - p;
    double dens = std::exp(-x + (a - 1.0) * std::log(x) - std::lgamma(a));
    if (!(dens > 0)) break;
    double nx = x - fx / dens;
    if (!(nx > 0) || nx == x) break;
    x = nx;
  }
  return x;
}

// gamma.ml:52-70 (Yang 1994 mean rates, made mean one)
int discrete_gamma(int n_cat, double alpha, double *rates) {
  if (n_cat <= 0 || !(alpha > 0)) return PP_EINVAL;
  std::vector<double> freq(n_cat + 1);
  for (int i = 0; i <= n_cat; i++) {
    if (i == 0) freq[i] = 0.;
    else if (i == n_cat) freq[i] = 1.;
    else {
      double y = (double)i / (double)n_cat;
      double cut = y < gamma_inc_P(alpha, 1e-7) ? 0. : gamma_Pinv(alpha, y);
      freq[i] = gamma_inc_P(alpha + 1., cut);
    }
  }
  double tot = 0.;
  for (int i = 0; i < n_cat; i++) tot += freq[i + 1] - freq[i];
  for (int i = 0; i < n_cat; i++) rates[i] = (double)n_cat * (freq[i + 1] - freq[i]) / tot;
  return PP_OK;
}

// Cyclic Jacobi eigensolver for the symmetric D^1/2 B D^1/2 (what diagd.ml
// takes from GSL symmv): column k of u is the eigenvector of l[k].
void symm_eigs(const double *m, int n, double *l, double *u) {
  std::vector<double> a(m, m + n * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) u[i * n + j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        if (i != j) off += a[i * n + j] * a[i * n + j];
        else diag += a[i * n + j] * a[i * n + j];
      }
    if (off <= 1e-300 || off <= diag * 1e-34) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = a[p * n + q];
        if (apq == 0.0) continue;
        double app = a[p * n + p], aqq = a[q * n + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {
          double akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = c * akp - s * akq;
          a[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = c * apk - s * aqk;
          a[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double ukp = u[k * n + p], ukq = u[k * n + q];
          u[k * n + p] = c * ukp - s * ukq;
          u[k * n + q] = s * ukp + c * ukq;
        }
      }
  }
  for (int i = 0; i < n; i++) l[i] = a[i * n + i];
}

static double safe_invert(double x) { return x == 0. ? 1. : 1. / x; }  // diagd.ml:118

static void dediag(double *dst, const double *u, const double *lambda, const double *uit,
                   int n) {  // linear_c.c:92-146
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double acc = 0;
      for (int k = 0; k < n; k++) acc += lambda[k] * u[i * n + k] * uit[j * n + k];
      dst[i * n + j] = acc;
    }
}

// ---- tree summaries for the -p priors --------------------------------------

static double subtree_length(const TreeDesc &t, int i) {
  double acc = t.bl[i];  // fold_left (+.) (bl id) over the children
  for (int c = t.child_start[i]; c < t.child_start[i + 1]; c++) acc = acc + subtree_length(t, t.children[c]);
  return acc;
}

double tree_length(const TreeDesc &t) {
  const int root = t.n_nodes - 1;
  double tot = 0.;
  for (int c = t.child_start[root]; c < t.child_start[root + 1]; c++)
    tot = tot + subtree_length(t, t.children[c]);
  return tot;
}

// returns the distances from the top of edge i to the leaves below, in order
static std::vector<double> mean_map_aux(const TreeDesc &t, int i, double lower, double *means) {
  const double bl = t.bl[i];
  if (t.child_start[i] == t.child_start[i + 1]) {
    means[i] = lower + bl;
    return {bl};
  }
  std::vector<double> dist;
  for (int c = t.child_start[i]; c < t.child_start[i + 1]; c++) {
    std::vector<double> d = mean_map_aux(t, t.children[c], lower, means);
    dist.insert(dist.end(), d.begin(), d.end());
  }
  const double top = lower + bl;
  double tot = top + dist[0];  // List.reduce (+.) over the shifted list
  for (size_t k = 1; k < dist.size(); k++) tot = tot + (top + dist[k]);
  means[i] = tot / (double)dist.size();
  for (double &d : dist) d = bl + d;
  return dist;
}

void prior_mean_map(const TreeDesc &t, double lower, double *means) {
  for (int i = 0; i < t.n_nodes; i++) means[i] = 0.;
  const int root = t.n_nodes - 1;
  for (int c = t.child_start[root]; c < t.child_start[root + 1]; c++)
    (void)mean_map_aux(t, t.children[c], lower, means);
}

// diagd.ml:79-100; with Price-CAT one matrix per category (gcat_model.ml:70-82)
void multi_exp(const HostModel &m, double bl, double *dst) {
  int S = m.S;
  double util[kMaxStates];
  const double *rates = m.sl_rates();
  for (int r = 0; r < m.n_sl(); r++) {
    for (int i = 0; i < S; i++) util[i] = std::exp(rates[r] * bl * m.l[i]);
    dediag(dst + r * S * S, m.x.data(), util, m.xit.data(), S);
  }
}

// Character tables: nucleotides (nuc_models.ml:24-45, columns A C G T) and
// amino acids (prot_models.ml:5-49).  Informative codes first.
static void fill_codes(HostModel &m) {
  for (int i = 0; i < 256; i++) m.code_of[i] = -1;
  m.lv.clear();
  auto add = [&](const char *chars, const double *v) {
    for (const char *c = chars; *c; c++) {
      m.code_of[(unsigned char)*c] = m.n_codes;
      m.code_of[(unsigned char)std::tolower(*c)] = m.n_codes;
    }
    m.lv.insert(m.lv.end(), v, v + m.S);
    m.n_codes++;
  };
  m.n_codes = 0;
  if (m.S == 4) {
    static const struct { const char *c; int bits; } tab[] = {
        {"A", 1}, {"C", 2}, {"G", 4}, {"T", 8}, {"R", 5}, {"Y", 10}, {"M", 3}, {"K", 12},
        {"S", 6}, {"W", 9}, {"H", 11}, {"B", 14}, {"V", 7}, {"D", 13}, {"NX", 15}};
    for (auto &t : tab) {
      double v[4];
      for (int i = 0; i < 4; i++) v[i] = (t.bits >> i) & 1 ? 1. : 0.;
      add(t.c, v);
    }
    m.n_inf = m.n_codes;
    double ones[4] = {1, 1, 1, 1};
    add("-?", ones);
  } else {
    static const char code[] = "ARNDCQEGHILKMFPSTWYV";
    for (int i = 0; i < 20; i++) {
      double v[20] = {0};
      v[i] = 1.;
      char s[2] = {code[i], 0};
      add(s, v);
    }
    double b[20] = {0}, j[20] = {0}, z[20] = {0}, ones[20];
    b[3] = b[2] = 1.;   // B = D or N
    j[9] = j[10] = 1.;  // J = I or L
    z[6] = z[5] = 1.;   // Z = E or Q
    for (int i = 0; i < 20; i++) ones[i] = 1.;
    add("B", b);
    add("J", j);
    add("Z", z);
    add("X", ones);
    m.n_inf = m.n_codes;
    add("-?", ones);
  }
}

// gmix_model.ml:25-37 -> Diagd.normed_of_exchangeable_pair (diagd.ml:123-172)
int build_model(HostModel &m, int S, int R, const double *exch, const double *statd,
                const double *rates, std::string &err) {
  if (!(S == 4 || S == 20)) { err = "n_states must be 4 or 20"; return PP_EINVAL; }
  if (R < 1 || R > kMaxRates) { err = "n_rates must be in 1..32"; return PP_EINVAL; }
  m.S = S;
  m.R = R;
  m.n_cats = 0;
  m.cat_rates.clear();
  m.site_cat.clear();
  m.statd.assign(statd, statd + S);
  m.rates.assign(rates, rates + R);
  for (int i = 0; i < S; i++)
    if (!(statd[i] >= 0.)) { err = "negative element in the diagonal of a DB matrix!"; return PP_EINVAL; }
  std::vector<double> b(S * S), droot(S), mm(S * S), u(S * S);
  for (int i = 0; i < S; i++)
    for (int j = 0; j < S; j++) {
      if (i != j) b[i * S + j] = exch[i * S + j];
      else {
        double total = 0.;
        for (int k = 0; k < S; k++)
          if (k != i) total += exch[k * S + i] * statd[k];
        b[i * S + j] = -(total * safe_invert(statd[i]));
      }
    }
  for (int i = 0; i < S; i++) droot[i] = std::sqrt(statd[i]);
  for (int i = 0; i < S; i++)
    for (int j = 0; j < S; j++) mm[i * S + j] = droot[i] * (b[i * S + j] * droot[j]);
  m.l.assign(S, 0.);
  symm_eigs(mm.data(), S, m.l.data(), u.data());
  m.x.assign(S * S, 0.);
  m.xit.assign(S * S, 0.);
  for (int i = 0; i < S; i++)
    for (int j = 0; j < S; j++) {
      m.x[i * S + j] = safe_invert(droot[i]) * u[i * S + j];
      m.xit[i * S + j] = droot[i] * u[i * S + j];
    }
  // normalize_rate: -sum_i pi_i Q_ii = 1
  std::vector<double> q(S * S);
  dediag(q.data(), m.x.data(), m.l.data(), m.xit.data(), S);
  double rate = 0.;
  for (int i = 0; i < S; i++) rate = rate - q[i * S + i] * statd[i];
  double s = 1. / rate;
  for (int i = 0; i < S; i++) m.l[i] *= s;
  fill_codes(m);
  m.built = true;
  return PP_OK;
}

void HostModel::fill_dev(DevModel *d) const {
  std::memset(d, 0, sizeof(*d));
  d->R = R;
  d->S = S;
  d->n_codes = n_codes;
  d->n_inf = n_inf;
  d->log_R = std::log((double)(float)R);
  d->n_sl = n_sl();
  d->cat = n_cats > 0;
  for (int i = 0; i < S; i++) { d->statd[i] = statd[i]; d->l[i] = l[i]; }
  for (int i = 0; i < S * S; i++) { d->x[i] = x[i]; d->xit[i] = xit[i]; }
  for (int r = 0; r < n_sl(); r++) d->rates[r] = sl_rates()[r];
  for (int i = 0; i < n_codes * S; i++) d->lv[i] = lv[i];
}

}  // namespace pp

using namespace pp;

extern "C" {

const char *pp_last_error(void) { return g_err.c_str(); }

int pp_discrete_gamma(int n_cat, double alpha, double *rates) {
  int st = discrete_gamma(n_cat, alpha, rates);
  if (st) set_error("discrete_gamma: need n_cat > 0 and alpha > 0");
  return st;
}

int pp_emper_freq(int n_states, const char *seqs, int n_seq, int n_sites, double *freq) {
  HostModel m;
  m.S = n_states;
  if (!(n_states == 4 || n_states == 20)) { set_error("n_states must be 4 or 20"); return PP_EINVAL; }
  fill_codes(m);
  int S = n_states;
  for (int i = 0; i < S; i++) freq[i] = 0.;
  for (long s = 0; s < n_seq; s++)
    for (long k = 0; k < n_sites; k++) {
      unsigned char c = (unsigned char)seqs[s * n_sites + k];
      if (c == '-' || c == '?') continue;
      int code = m.code_of[c];
      if (code < 0) {
        set_error(std::string("'") + (char)c + "' not a known base");
        return PP_EINVAL;
      }
      const double *lv = &m.lv[code * S];
      double tot = 0.;
      for (int i = 0; i < S; i++) tot += lv[i];
      double sc = 1. / tot;
      for (int i = 0; i < S; i++) freq[i] += lv[i] * sc;
    }
  double tot = 0.;
  for (int i = 0; i < S; i++) tot += freq[i];
  double sc = 1. / tot;
  for (int i = 0; i < S; i++) freq[i] *= sc;
  return PP_OK;
}

int pp_premask(int n_states, const char *ref_seqs, int n_ref, const char *query_seqs,
               int n_query, int n_sites, int discard_nonoverlapped, uint8_t *mask,
               uint8_t *query_keep, int *masklen) {
  HostModel m;
  m.S = n_states;
  if (!(n_states == 4 || n_states == 20)) { set_error("n_states must be 4 or 20"); return PP_EINVAL; }
  fill_codes(m);
  std::vector<uint8_t> ref_mask(n_sites, 0), q_mask(n_sites, 0);
  auto inf = [](char c) { return !(c == '-' || c == '?'); };
  for (long s = 0; s < n_ref; s++)
    for (long k = 0; k < n_sites; k++) {
      char c = ref_seqs[s * n_sites + k];
      if (m.code_of[(unsigned char)c] < 0) {
        set_error(std::string(1, c) + " is not a known base");
        return PP_EINVAL;
      }
      if (inf(c)) ref_mask[k] = 1;
    }
  for (long q = 0; q < n_query; q++) {
    const char *seq = query_seqs + q * n_sites;
    bool overlaps = false;
    for (long k = 0; k < n_sites; k++)
      if (inf(seq[k]) && ref_mask[k]) { overlaps = true; break; }
    if (!overlaps && !discard_nonoverlapped) {
      set_error("Sequence " + std::to_string(q) + " doesn't overlap any reference sequences.");
      return PP_EINVAL;
    }
    query_keep[q] = overlaps;
    if (!overlaps) continue;
    for (long k = 0; k < n_sites; k++) {
      char c = seq[k];
      if (m.code_of[(unsigned char)c] < 0) {
        set_error(std::string(1, c) + " is not a known base");
        return PP_EINVAL;
      }
      if (inf(c)) q_mask[k] = 1;
    }
  }
  int n = 0;
  for (long k = 0; k < n_sites; k++) {
    mask[k] = ref_mask[k] && q_mask[k];
    n += mask[k];
  }
  *masklen = n;
  if (n == 0) {
    set_error("Sequence length cut to 0 by pre-masking; can't proceed with no information.");
    return PP_EINVAL;
  }
  return PP_OK;
}

int pp_tree_length(int n_nodes, const int *child_start, const int *children, const double *bl,
                   double *length) {
  if (!child_start || !children || !bl || !length || n_nodes < 1) {
    pp::set_error("bad arguments");
    return PP_EINVAL;
  }
  pp::TreeDesc t{n_nodes, 0, child_start, children, bl};
  *length = pp::tree_length(t);
  return PP_OK;
}

int pp_prior_mean_map(int n_nodes, const int *child_start, const int *children,
                      const double *bl, double lower, double *means) {
  if (!child_start || !children || !bl || !means || n_nodes < 1) {
    pp::set_error("bad arguments");
    return PP_EINVAL;
  }
  pp::TreeDesc t{n_nodes, 0, child_start, children, bl};
  pp::prior_mean_map(t, lower, means);
  return PP_OK;
}

// Test hook, not part of the drop-in ABI: the correctly rounded exp and log
// of pp_crmath.h evaluated on the host (the same source the kernels use).
int pp_test_cr_math_host(long long n, const double *x, double *out_exp, double *out_log) {
  if (n < 0 || (n > 0 && (!x || !out_exp || !out_log))) return PP_EINVAL;
  for (long long i = 0; i < n; i++) {
    out_exp[i] = pp::cr::exp(x[i]);
    out_log[i] = pp::cr::log(x[i]);
  }
  return PP_OK;
}

void pp_prefs_default(pp_prefs *p) {
  p->start_pend = 0.1;
  p->max_pend = 2.0;
  p->initial_tolerance = 0.01;
  p->strike_box = 3.0;
  p->keep_factor = 0.01;
  p->max_strikes = 6;
  p->max_pitches = 40;
  p->keep_at_most = 7;
  p->calc_pp = 0;
  p->pp_rel_err = 0.01;
}

}  // extern "C"
