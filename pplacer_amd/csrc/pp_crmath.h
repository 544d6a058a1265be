// Correctly rounded FP64 exp and log, for the FP64 form's transition
// matrices and site logs.
//
// The C stubs call the host libm (glibc), whose exp and log return the
// correctly rounded result in all but a tiny fraction of cases; the device's
// library routines are faithful (within one ulp) but differ from the
// correctly rounded value on a few percent of exp and a fraction of a
// percent of log arguments.  Each such ulp can flip a Brent comparison at a
// flat optimum.  These routines return the correctly rounded result, so the
// device agrees with the stubs' libm wherever that libm rounds correctly.
//
// Method (Ziv's strategy, table-driven, own constants from
// gen_crmath_tables.py):
// * exp: x = k ln2/128 + r, |r| <= ln2/256 (+ rounding), exp(x) =
//   2^(k >> 7) * 2^((k & 127)/128) * exp(r), the table entry a double-double;
//   r is reduced in double-double (ln2/128 = C1 + C2 + C3, k*C1 exact), and
//   exp(r) = 1 + r + r^2/2 + r^3 q(r) with r^2 split exactly.  Relative error
//   below 2^-77.
// * log: x = 2^e m, m in [0.75, 1.5); m's interval j of width 1/128 gives
//   u = m inv[j] - 1, |u| < 2^-6, and
//   log x = e ln2 + L[j] + log1p(u) with L[j] = -log(inv[j]) a double-double;
//   inv[j] (1/(interval centre) rounded to 8 significant bits) is formed
//   arithmetically, so that u = m inv[j] - 1 (as an exact double-double) is
//   ready while L[j] loads; log1p(u) = u - u^2/2 + u^3/3 + u^4 P(u) with u^2
//   and u^3 kept exactly.  Relative error below 2^-73.
// * The double-double result (h, l) is rounded to nearest and accepted when
//   h + l +/- eps |h| round alike (eps = 2^-74 for exp, 2^-71 for log, above
//   the error bounds); otherwise (rarely: the value lies that close to a
//   rounding boundary) the value is
//   recomputed in double-double throughout (error about 2^-100) and rounded.
// * Outside the fast range -- exp of x <= -708.3 (subnormal results) or
//   >= 709.7, log of zero, negatives, subnormals, infinities, NaN -- the
//   library routine answers.
// Every operation is a separately rounded IEEE operation or an fma: the
// library is built with -ffp-contract=off, so the error-free transformations
// below hold on the host and the device alike.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define PP_CR_FN __device__ __forceinline__
#define PP_CR_SLOW __device__ __attribute__((noinline))
#define PP_CR_TAB static __constant__ const
#else
#define PP_CR_FN static inline
#define PP_CR_SLOW static
#define PP_CR_TAB static const
#endif

#include "pp_crmath_tables.h"

namespace pp {
namespace cr {

struct DD {
  double hi, lo;
};

PP_CR_FN DD two_sum(double a, double b) {
  const double s = a + b, bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
// |a| >= |b| (or a == 0)
PP_CR_FN DD fast_two_sum(double a, double b) {
  const double s = a + b;
  return {s, b - (s - a)};
}
PP_CR_FN DD two_prod(double a, double b) {
  const double p = a * b;
  return {p, fma(a, b, -p)};
}
PP_CR_FN DD dd_add(DD a, DD b) {
  DD s = two_sum(a.hi, b.hi);
  const DD t = two_sum(a.lo, b.lo);
  s = fast_two_sum(s.hi, s.lo + t.hi);
  return fast_two_sum(s.hi, s.lo + t.lo);
}
PP_CR_FN DD dd_mul(DD a, DD b) {
  const DD p = two_prod(a.hi, b.hi);
  return fast_two_sum(p.hi, p.lo + (a.hi * b.lo + a.lo * b.hi));
}
PP_CR_FN DD dd_mul_d(DD a, double b) {
  const DD p = two_prod(a.hi, b);
  return fast_two_sum(p.hi, p.lo + a.lo * b);
}

// h + l rounded to nearest; ok when every value within eps |h| of h + l
// rounds to the same double (rounding is monotone, so the two ends decide).
PP_CR_FN double round_checked(double h, double l, double eps, bool &ok) {
  const double r = h + l, d = eps * fabs(h);
  ok = (h + (l + d)) == r && (h + (l - d)) == r;
  return r;
}

// exp(r) * T in double-double: Horner over 1/k!, k = 13..0 (a call: taken
// rarely, it stays out of the callers' registers)
PP_CR_SLOW double exp_slow(DD r, DD t) {
  DD p = {pp_cr_exp_pdd[26], pp_cr_exp_pdd[27]};
  for (int k = 12; k >= 0; k--) p = dd_add(dd_mul(p, r), DD{pp_cr_exp_pdd[2 * k], pp_cr_exp_pdd[2 * k + 1]});
  const DD y = dd_mul(p, t);
  bool ok;
  return round_checked(y.hi, y.lo, 0x1p-100, ok);
}

PP_CR_FN double exp(double x) {
  if (!(x > -708.3 && x < 709.7)) return ::exp(x);
  const double kf = rint(x * pp_cr_exp_c[0]);
  const int k = (int)kf, j = k & 127, m = (k - j) >> 7;  // k = 128 m + j
  // r = x - k ln2/128: k C1 is exact and x - k C1 exact (Sterbenz)
  const double r1 = x - kf * pp_cr_exp_c[1];
  const DD p2 = two_prod(kf, pp_cr_exp_c[2]);
  DD r = two_sum(r1, -p2.hi);
  r = fast_two_sum(r.hi, (r.lo - p2.lo) - kf * pp_cr_exp_c[3]);
  const DD t = {pp_cr_exp_t[2 * j], pp_cr_exp_t[2 * j + 1]};
  // exp(r) - 1 = rh + sq.hi/2 + tail, sq = rh^2 exactly
  const double rh = r.hi, rl = r.lo;
  const DD sq = two_prod(rh, rh);
  const double q = pp_cr_exp_p[3] +
                   rh * (pp_cr_exp_p[4] + rh * (pp_cr_exp_p[5] + rh * (pp_cr_exp_p[6] + rh * pp_cr_exp_p[7])));
  const double tail = 0.5 * sq.lo + (sq.hi * rh) * q + (rl + rl * rh);
  const double half_sq = 0.5 * sq.hi;
  // T (1 + E) = t.hi + t.hi rh + t.hi sq/2 + (t.hi tail + t.lo + t.lo E)
  const DD a = two_prod(t.hi, rh), b = two_prod(t.hi, half_sq);
  const double c = t.hi * tail + t.lo + t.lo * (rh + half_sq);
  const DD s1 = fast_two_sum(t.hi, a.hi), s2 = fast_two_sum(s1.hi, b.hi);
  const DD y = fast_two_sum(s2.hi, (s1.lo + s2.lo) + ((a.lo + b.lo) + c));
  bool ok;
  double res = round_checked(y.hi, y.lo, 0x1p-74, ok);
  if (!ok) res = exp_slow(r, t);
  return ldexp(res, m);
}

// exp as a call: for callers that inline their probe code at many sites, one
// copy of the routine instead of one per site
PP_CR_SLOW double exp_call(double x) { return exp(x); }

// log1p(u) + L + e ln2 in double-double: Horner over (-1)^(k+1)/k, k = 18..1
// (a call, as exp_slow)
PP_CR_SLOW double log_slow(DD u, DD l, double ef) {
  DD p = {pp_cr_log_pdd[34], pp_cr_log_pdd[35]};
  for (int k = 16; k >= 0; k--) p = dd_add(dd_mul(p, u), DD{pp_cr_log_pdd[2 * k], pp_cr_log_pdd[2 * k + 1]});
  p = dd_mul(p, u);
  const DD e2 = two_prod(ef, pp_cr_ln2[1]);
  DD en = fast_two_sum(ef * pp_cr_ln2[0], e2.hi);  // e hi is exact
  en = fast_two_sum(en.hi, en.lo + (e2.lo + ef * pp_cr_ln2[2]));
  const DD y = dd_add(dd_add(en, l), p);
  bool ok;
  return round_checked(y.hi, y.lo, 0x1p-100, ok);
}

PP_CR_FN double log(double x) {
  if (!(x >= 0x1p-1022 && x < INFINITY)) return ::log(x);
  int e;
  double m = 2. * frexp(x, &e);  // x = m 2^(e-1), m in [1, 2)
  e -= 1;
  if (m >= 1.5) {
    m *= 0.5;
    e += 1;
  }
  const int j = (int)((m - 0.75) * 128.);  // m - 0.75 exact
  // inv[j]: 1/(interval centre) rounded to 8 significant bits; 1 next to 1
  const double q = 1. / (0.75 + ((double)j + 0.5) * 0.0078125);
  const double inv = (j == 31 || j == 32) ? 1. : q < 1. ? rint(q * 256.) * 0x1p-8 : rint(q * 128.) * 0x1p-7;
  // u + ul = m inv - 1 exactly (p - 1 is exact: p is within 2^-6 of 1)
  const double p = m * inv;
  const DD uu = two_sum(p - 1., fma(m, inv, -p));
  const double u = uu.hi, ul = uu.lo;
  const DD l = {pp_cr_log_l[2 * j], pp_cr_log_l[2 * j + 1]};
  const double ef = (double)e;
  // u^2 and u^3 exactly as double-doubles
  const DD u2 = two_prod(u, u);
  DD u3 = two_prod(u2.hi, u);
  u3.lo += u2.lo * u;
  // u^3/3
  const DD c3 = two_prod(u3.hi, pp_cr_log_pdd[4]);
  const double c3lo = c3.lo + (u3.hi * pp_cr_log_pdd[5] + u3.lo * pp_cr_log_pdd[4]);
  // P(u) = sum_{k=3..15} c_k u^(k-3) by Estrin's scheme (four levels)
  const double *c = pp_cr_log_p;
  const double w2 = u2.hi, w4 = w2 * w2, w8 = w4 * w4;
  const double q0 = (c[3] + u * c[4]) + w2 * (c[5] + u * c[6]);
  const double q1 = (c[7] + u * c[8]) + w2 * (c[9] + u * c[10]);
  const double q2 = (c[11] + u * c[12]) + w2 * (c[13] + u * c[14]);
  const double pp = (q0 + w4 * q1) + w8 * (q2 + w4 * c[15]);
  const double r4 = w4 * pp;  // u^4 P(u)
  // e ln2 + L + u - u^2/2 + u^3/3 + u^4 P(u): |e ln2| >= ln2 > |L|, |L| >= 2^-8
  // or L = 0 (and then |u| >= 2^-53 |u^2|): only u's term needs a two_sum
  DD s = fast_two_sum(ef * pp_cr_ln2[0], l.hi);
  double lo = s.lo;
  s = two_sum(s.hi, u);
  lo += s.lo;
  s = fast_two_sum(s.hi, -0.5 * u2.hi);
  lo += s.lo;
  s = fast_two_sum(s.hi, c3.hi);
  lo += s.lo;
  // ul / (1 + u) for the part of u past one double
  lo += ((ef * pp_cr_ln2[1] + ef * pp_cr_ln2[2]) + l.lo) + ((-0.5 * u2.lo + c3lo) + r4) + (ul - ul * u);
  const DD y = fast_two_sum(s.hi, lo);
  bool ok;
  const double res = round_checked(y.hi, y.lo, 0x1p-71, ok);
  return ok ? res : log_slow(uu, l, ef);
}

}  // namespace cr
}  // namespace pp
