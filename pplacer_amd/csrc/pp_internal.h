// Internal declarations shared by the host code and the HIP kernels of
// libpplacer_hip.so.  Not part of the public C ABI (include/pplacer_hip.h).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "pplacer_hip.h"

namespace pp {

constexpr int kMaxStates = 20;
constexpr int kMaxRates = 32;  // gamma rates, or Price-CAT categories
constexpr int kMaxCodes = 32;

// Model as the kernels see it.  Character codes: index into lv[]; codes
// >= n_inf are the non-informative '-' and '?' (alignment.ml:235).
// Transition-matrix slices: one per gamma rate (R of them, all used at every
// site), or with Price-CAT one per category (R == 1, each site uses the slice
// of its category).  rates[] holds the slices' rates.
struct DevModel {
  int R, S, n_codes, n_inf;
  int n_sl, cat;
  double log_R;  // log((float)R) as the C stubs compute it (linear_c.c:421)
  double statd[kMaxStates];
  double x[kMaxStates * kMaxStates];
  double l[kMaxStates];
  double xit[kMaxStates * kMaxStates];
  double rates[kMaxRates];
  double lv[kMaxCodes * kMaxStates];  // [code][state] 0/1 likelihood vectors
};

// ---- host-side model and tree algorithms (pp_host.cpp) -------------------
struct HostModel {
  int R = 0, S = 0;
  std::vector<double> statd, rates, x, l, xit;  // diagd.ml:33-39
  // Price-CAT (gcat_model.ml): category rates and the category of each
  // pre-masked site; R == 1 and rates == {1} then
  int n_cats = 0;
  std::vector<double> cat_rates;
  std::vector<int> site_cat;
  int n_sl() const { return n_cats > 0 ? n_cats : R; }
  const double *sl_rates() const { return n_cats > 0 ? cat_rates.data() : rates.data(); }
  // character code table
  int code_of[256];
  int n_codes = 0, n_inf = 0;
  std::vector<double> lv;  // [n_codes][S]
  bool built = false;
  void fill_dev(DevModel *m) const;
};

int build_model(HostModel &m, int S, int R, const double *exch, const double *statd,
                const double *rates, std::string &err);
int discrete_gamma(int n_cat, double alpha, double *rates);
void symm_eigs(const double *m, int n, double *l, double *u);
void multi_exp(const HostModel &m, double bl, double *dst /*[n_sl][S][S]*/);

void set_error(const std::string &msg);

}  // namespace pp
