// Sequential three-taxon control shared by the stage-2 engines: pplacer's
// Three_tax.optimize (three_tax.ml:48-85), Minimization.brent with its
// bisection start finder (minimization.ml:11-100), the GSL 2.x Brent
// minimiser it calls (restated: golden section 0.3819660, step tolerance
// sqrt(DBL_EPSILON)*|x|, parabolic step when admissible), and
// safe_ml_optimize_location (core.ml:255-281), and for -p the marginal
// likelihood of Three_tax.calc_marg_prob (three_tax.ml:109-215) with its
// nested QUADPACK QNG integrations.  Runs on wave-uniform values in the
// engine's control wave; the engine D supplies log_like(), one probe of
// Model.log_like3 at the current (pend_bl, dist_bl, prox_bl).
#pragma once

#include <cfloat>

#include "pp_device.h"

#define PP_QNG_STORAGE static __constant__ const
#include "pp_qng_tables.h"

namespace pp {

enum { QNG_OK = 0, QNG_ETOL = 1, QNG_EBADTOL = 2, QNG_ABORT = 3, QNG_ASSERT = 4 };
enum { PRIOR_EXP = 0, PRIOR_UNIFORM = 1 };

// Up to 128 wave-uniform doubles spread over the wave: value j lives in lane
// j % 64 of word j / 64.
struct LaneStore {
  double w0, w1;
  __device__ __forceinline__ void put(int j, double v) {
    if ((int)(threadIdx.x & 63) == (j & 63)) {
      if (j < 64) w0 = v;
      else w1 = v;
    }
  }
  __device__ __forceinline__ double get(int j) const { return readlane_f64(j < 64 ? w0 : w1, j & 63); }
  // the sum of the +/- pair k of the concatenated abscissa tables
  __device__ __forceinline__ double pair(int k) const { return get(1 + 2 * k) + get(2 + 2 * k); }
};

// QUADPACK's error estimate from the difference of two rules, the integral of
// |f| and of |f - mean|.
__device__ __forceinline__ double qng_rescale_error(double err, double result_abs, double result_asc) {
  err = fabs(err);
  if (result_asc != 0 && err != 0) {
    const double scale = pow(200 * err / result_asc, 1.5);
    err = scale < 1 ? result_asc * scale : result_asc;
  }
  if (result_abs > DBL_MIN / (50 * DBL_EPSILON)) {
    const double min_err = 50 * DBL_EPSILON * result_abs;
    if (min_err > err) err = min_err;
  }
  return err;
}

// QUADPACK QNG with epsabs = 0 (Integration.value_integrate, integration.ml:28):
// the 21-point Gauss-Kronrod rule checked against the 10-point Gauss rule,
// then Patterson's 43- and 87-point extensions reusing every evaluation.  f is
// evaluated at one call site, at the centre and then at the +/- pairs of the
// tables in order; f(x, err) sets err to abandon the rule.
template <class F>
__device__ __forceinline__ int qng(F &&f, double a, double b, double epsrel, LaneStore &st,
                                   double *result) {
  *result = 0;
  if (epsrel < 50 * DBL_EPSILON || epsrel < 0.5e-28) return QNG_EBADTOL;
  const double half_length = 0.5 * (b - a), abs_half_length = fabs(half_length);
  const double center = 0.5 * (b + a);
  double resabs = 0, resasc = 0, prev = 0;
  int j = 0;
#pragma unroll 1
  for (int stage = 0; stage < 3; stage++) {
    const int end = stage == 0 ? 21 : stage == 1 ? 43 : 87;
#pragma unroll 1
    for (; j < end; j++) {
      double x = center;
      if (j > 0) {
        const int p = (j - 1) >> 1;
        const double t = p < 5 ? pp_qng_x1[p] : p < 10 ? pp_qng_x2[p - 5]
                       : p < 21 ? pp_qng_x3[p - 10] : pp_qng_x4[p - 21];
        const double abscissa = half_length * t;
        x = ((j - 1) & 1) ? center - abscissa : center + abscissa;
      }
      int err = 0;
      const double v = f(x, err);
      if (err) return QNG_ABORT;
      st.put(j, v);
    }
    const double fc = st.get(0);
    double res, diff;
    if (stage == 0) {
      double res10 = 0, res21 = pp_qng_w21b[5] * fc;
      resabs = pp_qng_w21b[5] * fabs(fc);
#pragma unroll 1
      for (int k = 0; k < 5; k++) {
        const double f1 = st.get(1 + 2 * k), f2 = st.get(2 + 2 * k), fval = f1 + f2;
        res10 += pp_qng_w10[k] * fval;
        res21 += pp_qng_w21a[k] * fval;
        resabs += pp_qng_w21a[k] * (fabs(f1) + fabs(f2));
      }
#pragma unroll 1
      for (int k = 0; k < 5; k++) {
        const double f1 = st.get(11 + 2 * k), f2 = st.get(12 + 2 * k), fval = f1 + f2;
        res21 += pp_qng_w21b[k] * fval;
        resabs += pp_qng_w21b[k] * (fabs(f1) + fabs(f2));
      }
      resabs *= abs_half_length;
      const double mean = 0.5 * res21;
      resasc = pp_qng_w21b[5] * fabs(fc - mean);
#pragma unroll 1
      for (int k = 0; k < 5; k++)
        resasc += (pp_qng_w21a[k] * (fabs(st.get(1 + 2 * k) - mean) + fabs(st.get(2 + 2 * k) - mean))
                   + pp_qng_w21b[k] * (fabs(st.get(11 + 2 * k) - mean) + fabs(st.get(12 + 2 * k) - mean)));
      resasc *= abs_half_length;
      res = res21;
      diff = res21 - res10;
    } else if (stage == 1) {
      double res43 = pp_qng_w43b[11] * fc;
#pragma unroll 1
      for (int k = 0; k < 10; k++) res43 += st.pair(k) * pp_qng_w43a[k];
#pragma unroll 1
      for (int k = 0; k < 11; k++) res43 += st.pair(10 + k) * pp_qng_w43b[k];
      res = res43;
      diff = res43 - prev;
    } else {
      double res87 = pp_qng_w87b[22] * fc;
#pragma unroll 1
      for (int k = 0; k < 21; k++) res87 += st.pair(k) * pp_qng_w87a[k];
#pragma unroll 1
      for (int k = 0; k < 22; k++) res87 += pp_qng_w87b[k] * st.pair(21 + k);
      res = res87;
      diff = res87 - prev;
    }
    const double result_kronrod = res * half_length;
    const double err = qng_rescale_error(diff * half_length, resabs, resasc);
    *result = result_kronrod;
    if (err < 0. || err < epsrel * fabs(result_kronrod)) return QNG_OK;
    prev = res;
  }
  return QNG_ETOL;  // "failed to reach tolerance with highest-order rule"
}

template <class D>
struct ThreeTaxControl {
  // Three_tax state (last probe) and caches
  int M;
  double pend_bl, dist_bl, prox_bl;
  double c_pend, c_dist, c_prox;
  double esum, last_ll;
  bool have_last;
  // n_calls: Model.log_like3 calls as the reference makes them
  // (Three_tax.n_like_calls), the GSL set's re-evaluations included;
  // n_probes / n_dist_probes: probes this engine actually ran
  long long n_calls, n_probes, n_dist_probes;

  __device__ __forceinline__ D &self() { return *static_cast<D *>(this); }

  // -p: the pendant prior of the current location (core.ml:86-92) and the
  // target of find_upper_limit
  int pr_kind;
  double pr_mu, up_target;

  __device__ __forceinline__ double prior(double x) const {
    if (pr_kind == PRIOR_UNIFORM) return 1.;
    return x < 0 ? 0. : exp(-x / pr_mu) / pr_mu;  // gsl_ran_exponential_pdf
  }

  // ---- Minimization.brent with the GSL Brent minimiser ---------------------
  // kind 0: pendant (three_tax.ml:38-39), 1: distal (three_tax.ml:41-45),
  // 2: find_upper_limit's |target - (log_like + log prior)| over the pendant
  // length (three_tax.ml:142-149)
  __device__ __forceinline__ double fprobe(int kind, double v, int &err) {
    if (kind == 1) {
      const double prox = (dist_bl + prox_bl) - v;
      if (!(prox >= 0.)) err = ST_ASSERT;
      dist_bl = v;
      prox_bl = prox;
    } else {
      pend_bl = v;
    }
    const double ll = self().log_like();
    if (kind != 2) return -ll;
    const double x = fabs(up_target - (ll + log(prior(v))));
    return x == INFINITY ? 100000.0 : x;
  }

  __device__ __forceinline__ int brent(int kind, double raw_start, double left, double right, double tol,
                       int max_iters, double *result) {
    int err = 0;
    if (left >= raw_start || raw_start >= right) return ST_INVALID_START;
    const double righty = fprobe(kind, right, err);  // (f left, f right) evaluates right first
    if (err) return err;
    const double prox_r = prox_bl;
    const double lefty = fprobe(kind, left, err);
    if (err) return err;
    const double prox_l = prox_bl;
    const double smaller = lefty < righty ? left : right;
    const double miny = lefty <= righty ? lefty : righty;  // Stdlib.min
    double prev = raw_start, pv, prox_p;
    for (int iter = 1;; iter++) {
      pv = fprobe(kind, prev, err);
      if (err) return err;
      prox_p = prox_bl;
      if (pv < miny) break;
      if (fabs(prev - smaller) < tol) { *result = smaller; return ST_OK; }  // FoundMin
      if (iter > max_iters) return ST_FIND_START;
      prev = (prev + smaller) / 2.;
    }
    // gsl_min_fminimizer_set evaluates f(lower), f(upper), f(start) again.  A
    // pendant probe moves only the pendant length, so those are the values
    // just computed, and the state is already at start.  A distal probe also
    // moves the proximal length to (dist + prox) - v, whose rounding drifts
    // from probe to probe: the three moves are replayed in the set's order,
    // with their assertion, and a value is reused only where the proximal
    // length comes out bit-identical to the one it was computed at.
    double x_lower = left, x_upper = right, x_min = prev;
    double f_lower = lefty, f_upper = righty, f_min = pv;
    // (scalar selects and one probe site: no addressable locals in the loop)
#pragma nounroll
    for (int k = 0; k < 3; k++) {
      const double x = k == 0 ? left : k == 1 ? right : prev;
      double f = k == 0 ? f_lower : k == 1 ? f_upper : f_min;
      if (kind == 1) {
        const double prox = (dist_bl + prox_bl) - x;
        if (prox == (k == 0 ? prox_l : k == 1 ? prox_r : prox_p)) {
          dist_bl = x;
          prox_bl = prox;
          n_calls++;  // the reference's call, answered from the value at hand
        } else {
          f = fprobe(kind, x, err);
          if (err) return err;
        }
      } else {
        n_calls++;
      }
      if (!isfinite(f)) return ST_GSL_BADFUNC;
      if (k == 0) f_lower = f;
      else if (k == 1) f_upper = f;
      else f_min = f;
    }
    if (x_lower > x_upper || x_min >= x_upper || x_min <= x_lower) return ST_GSL;
    if (f_min >= f_lower || f_min >= f_upper) return ST_GSL;
    const double golden = 0.3819660;
    double v = x_lower + golden * (x_upper - x_lower), w_ = v;
    double sd = 0., se = 0.;
    double f_v = fprobe(kind, v, err);
    if (err) return err;
    if (!isfinite(f_v)) return ST_GSL_BADFUNC;
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
      if (!isfinite(f_u)) return ST_GSL_BADFUNC;
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
  __device__ __forceinline__ int optimize_something(int kind, double tol, double start_v, double min_v,
                                    double max_v, double *res) {
    if (min_v == max_v) { *res = start_v; return ST_OK; }
    return brent(kind, start_v, min_v, max_v, tol, 100, res);
  }

  // three_tax.ml:70-83
  __device__ __forceinline__ int optimize(double tol, double max_pend, int max_iter) {
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
  __device__ __forceinline__ void set_state(double cut_bl, double pendant, double distal) {
    pend_bl = pendant;
    dist_bl = distal;
    prox_bl = cut_bl - distal;
    have_last = false;
  }

  // core.ml:255-281: optimize; on a GSL error retry once from the default
  // branch lengths; ExceededMaxIter keeps the current state.  The result is
  // Three_tax.get_results (three_tax.ml:85) at the last probe.
  __device__ __forceinline__ int safe_optimize(double tol, double max_pend, double start_pend, double cut_bl,
                               double *ll, double *pend, double *dist) {
    for (int attempt = 0; attempt < 2; attempt++) {
      const int st = optimize(tol, max_pend, 200);
      if (st == ST_OK || st == ST_MAXITER) {
        *ll = have_last ? last_ll : self().log_like();
        *pend = pend_bl;
        *dist = dist_bl;
        return st;
      }
      if (!is_gsl(st) || attempt == 1) return st;
      set_state(cut_bl, start_pend, cut_bl / 2.);
    }
    return ST_OK;
  }

  // ---- posterior probabilities --------------------------------------------
  // three_tax.ml:121-163 (the prior is never 0 at the ML pendant length for an
  // exponential prior short of exp underflow, handled as the reference does)
  __device__ __forceinline__ int find_upper_limit(double max_pend, double orig_ll, double *limit) {
    const double orig_pend_bl = pend_bl;
    if (prior(orig_pend_bl) == 0.) {
      double curr = orig_pend_bl;
#pragma unroll 1
      for (;;) {
        const double next = curr / 2.;
        if (prior(next) > exp(-10.)) { *limit = curr; return ST_OK; }
        curr = next;
      }
    }
    up_target = -10. + orig_ll + log(prior(orig_pend_bl));
    const double start = orig_pend_bl + (max_pend - orig_pend_bl) / 2.;
    const int st = brent(2, start, orig_pend_bl, max_pend, 0.02, 10, limit);
    if (st == ST_GSL) { *limit = max_pend; return ST_OK; }  // Gsl_exn EINVAL
    return st;
  }

  // three_tax.ml:169-215 at the current state: *marg = base_ll +
  // log(cut_bl^-1 \int_0^cut \int_0^upper exp(ll - base_ll) prior(pend)), the
  // inner range halved on every QNG tolerance failure (at most 10 times).
  __device__ __forceinline__ int calc_marg_prob(double rel_err, double max_pend, double *marg) {
    const double base_ll = self().log_like();
    const double cut_bl = dist_bl + prox_bl;
    double upper;
    int st = find_upper_limit(max_pend, base_ll, &upper);
    if (st) return st;
    LaneStore s_in, s_out;
#pragma unroll 1
    for (int n_exceptions = 0;; n_exceptions++) {
      if (n_exceptions >= 10) { *marg = base_ll; return ST_OK; }
      int fail = QNG_OK;  // status of the failed rule, or QNG_ASSERT
      auto inner = [&](double pend, int &err) -> double {
        (void)err;
        pend_bl = pend;
        return exp(self().log_like() - base_ll) * prior(pend);
      };
      auto inner_integral = [&](int &err) -> double {
        double r;
        const int q = qng(inner, 0., upper, rel_err, s_in, &r);
        if (q != QNG_OK) { err = 1; fail = q; }
        return r;
      };
      double v;
      if (fabs(cut_bl - 0.) < 1e-5) {  // cut_bl =~ 0. (ppatteries.ml:96)
        const double prox = (dist_bl + prox_bl) - cut_bl / 2.;
        if (!(prox >= 0.)) return ST_ASSERT;
        dist_bl = cut_bl / 2.;
        prox_bl = prox;
        int err = 0;
        v = inner_integral(err);
      } else {
        auto outer = [&](double dist, int &err) -> double {
          const double prox = (dist_bl + prox_bl) - dist;
          if (!(prox >= 0.)) { err = 1; fail = QNG_ASSERT; return 0.; }
          dist_bl = dist;
          prox_bl = prox;
          return inner_integral(err);
        };
        double r;
        const int q = qng(outer, 0., cut_bl, rel_err, s_out, &r);
        if (q != QNG_OK && q != QNG_ABORT) fail = q;
        v = r / cut_bl;
      }
      if (fail == QNG_OK) { *marg = base_ll + log(v); return ST_OK; }
      if (fail == QNG_ASSERT) return ST_ASSERT;
      if (fail != QNG_ETOL) return ST_GSL;  // an uncaught Gsl_exn (EBADTOL)
      upper = upper / 2.;
    }
  }
};

}  // namespace pp
