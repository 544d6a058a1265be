// Sequential three-taxon control shared by the stage-2 engines: pplacer's
// Three_tax.optimize (three_tax.ml:48-85), Minimization.brent with its
// bisection start finder (minimization.ml:11-100), the GSL 2.x Brent
// minimiser it calls (restated: golden section 0.3819660, step tolerance
// sqrt(DBL_EPSILON)*|x|, parabolic step when admissible), and
// safe_ml_optimize_location (core.ml:255-281).  Runs on wave-uniform values in
// the engine's control wave; the engine D supplies log_like(), one probe of
// Model.log_like3 at the current (pend_bl, dist_bl, prox_bl).
#pragma once

#include "pp_device.h"

namespace pp {

template <class D>
struct ThreeTaxControl {
  // Three_tax state (last probe) and caches
  int M;
  double pend_bl, dist_bl, prox_bl;
  double c_pend, c_dist, c_prox;
  double esum, last_ll;
  bool have_last;
  long long n_calls, n_dist_probes;

  __device__ __forceinline__ D &self() { return *static_cast<D *>(this); }

  // ---- Minimization.brent with the GSL Brent minimiser ---------------------
  // kind 0: pendant (three_tax.ml:38-39), 1: distal (three_tax.ml:41-45)
  __device__ __forceinline__ double fprobe(int kind, double v, int &err) {
    if (kind == 0) {
      pend_bl = v;
    } else {
      const double prox = (dist_bl + prox_bl) - v;
      if (!(prox >= 0.)) err = ST_ASSERT;
      dist_bl = v;
      prox_bl = prox;
    }
    return -self().log_like();
  }

  __device__ __forceinline__ int brent(int kind, double raw_start, double left, double right, double tol,
                       int max_iters, double *result) {
    int err = 0;
    if (left >= raw_start || raw_start >= right) return ST_INVALID_START;
    const double righty = fprobe(kind, right, err);  // (f left, f right) evaluates right first
    if (err) return err;
    const double lefty = fprobe(kind, left, err);
    if (err) return err;
    const double smaller = lefty < righty ? left : right;
    const double miny = lefty <= righty ? lefty : righty;  // Stdlib.min
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
      if (st != ST_GSL || attempt == 1) return st;
      set_state(cut_bl, start_pend, cut_bl / 2.);
    }
    return ST_OK;
  }
};

}  // namespace pp
