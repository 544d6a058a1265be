/*
 * pp_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64 restatement of pplacer's per-(query, edge) placement path,
 * following the reference loop and summation order.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker or the timed CPU baseline.  The product (pplacer_amd/)
 * never links, imports or calls anything here.
 *
 * GLV layout (pplacer_src/linear_c.c:5-12): a[rate][site][state] float64,
 * row-major, plus a per-site exponent e[site] (OCaml native int, int64 here,
 * pplacer_src/gmix_model.ml:58-61).
 */
#ifndef PP_ORACLE_H
#define PP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes, mirroring the OCaml exceptions on the path */
enum {
  OR_OK = 0,
  OR_GSL_EBADFUNC = 1,      /* Gsl_exn: non-finite function value (GSL SAFE_FUNC_CALL) */
  OR_GSL_EINVAL = 2,        /* Gsl_exn: bad bracket handed to the Brent minimiser */
  OR_EXCEEDED_MAX_ITER = 3, /* Minimization.ExceededMaxIter */
  OR_FIND_START_FAILURE = 4,/* Minimization.FindStartFailure */
  OR_INVALID_START = 5,     /* Minimization.InvalidStartValues */
  OR_BAD_INPUT = 6,         /* failwith / invalid_arg on inputs */
  OR_NO_RESULTS = 7         /* "empty results for <query>" */
};

/* ---- linear_c.c restatements ------------------------------------------ */
void or_gemmish(double *dst, const double *a, const double *b, int n_sites, int n_states);
void or_dediagonalize(double *dst, const double *u, const double *lambda,
                      const double *uit, int n);
double or_ten_log_like3(const double *statd, const double *x, const double *y,
                        const double *z, double *util, int n_rates, int n_sites,
                        int n_states);
void or_ten_pairwise_prod(double *dst, const double *x, const double *y, long size);
void or_ten_statd_pairwise_prod(const double *statd, double *dst, const double *a,
                                const double *b, int n_rates, int n_sites, int n_states);
double or_ten_masked_logdot(const double *x, const double *y, const uint16_t *mask,
                            double *util, int n_rates, int n_sites, int n_states);
double or_ten_bounded_logdot(const double *x, const double *y, int first, int last,
                             double *util, int n_rates, int n_sites, int n_states);
double or_mat_masked_logdot(const double *x, const double *y, const uint16_t *mask,
                            int n_sites, int n_states);
double or_mat_log_like3(const double *statd, const double *x, const double *y,
                        const double *z, int n_sites, int n_states);
double or_mat_bounded_logdot(const double *x, const double *y, int first, int last, int n_states);

/* ---- gstar_support.ml --------------------------------------------------- */
int or_get_twoexp(double x);
double or_of_twoexp(int i);
double or_total_twoexp(const int64_t *e, int n);
double or_masked_total_twoexp(const int64_t *e, const uint16_t *mask, int n);
double or_bounded_total_twoexp(const int64_t *e, int start, int last);
void or_perhaps_pull_exponent(int min_allowed_twoexp, double *a, int64_t *e,
                              int n_rates, int n_sites, int n_states);

/* ---- gamma.ml ----------------------------------------------------------- */
double or_gamma_inc_P(double a, double x);
double or_gamma_Pinv(double a, double p);
int or_discrete_gamma(int n_cat, double alpha, double *rates);

/* ---- diagd.ml / linear_utils.ml ---------------------------------------- */
int or_symm_eigs(const double *m, int n, double *l, double *u);
void or_b_of_exchangeable_pair(const double *r, const double *pi, int n, double *b);
int or_of_d_b(const double *d, const double *b, int n, double *x, double *l, double *xit);
int or_normed_of_exchangeable_pair(const double *r, const double *pi, int n,
                                   double *x, double *l, double *xit);
void or_to_exp(double *dst, const double *x, const double *l, const double *xit, int n,
               double bl);
void or_multi_exp(double *dst, const double *x, const double *l, const double *xit,
                  const double *rates, int n_rates, int n, double bl);

/* ---- tip encoding (nuc_models.ml, prot_models.ml, alignment.ml) -------- */
int or_lv_of_char(int n_states, char c, double *lv);
int or_informative(char c);
void or_mask_reindex(const uint8_t *mask, int n, int *out, int *n_out);
int or_emper_freq(int n_states, const char *seqs, int n_seq, int n_sites, double *freq);
int or_premask(int n_states, const char *ref_seqs, int n_ref, const char *query_seqs,
               int n_query, int n_sites, int discard_nonoverlapped, uint8_t *mask,
               uint8_t *query_keep, int *masklen);

/* ---- model + GLV ops (gmix_model.ml) ------------------------------------ */
typedef struct {
  int n_rates;
  int n_states;
  const double *statd;  /* [S] */
  const double *x;      /* [S][S] */
  const double *l;      /* [S]   */
  const double *xit;    /* [S][S] */
  const double *rates;  /* [R]   */
  /* Price-CAT (gcat_model.ml:33-95, 327-345): with n_cats > 0 the GLVs carry
   * no rate axis (n_rates == 1, rates = {1}) and site i evolves by the
   * matrix of rate cat_rates[site_cat[i]] (i indexes the pre-masked sites). */
  int n_cats;
  const double *cat_rates;  /* [n_cats] */
  const int *site_cat;      /* [N] */
} or_model;

void or_evolve_into(const or_model *m, double *dst_a, int64_t *dst_e,
                    const double *src_a, const int64_t *src_e, int n_sites, double bl);
double or_log_like3(const or_model *m, double *util, int n_sites,
                    const double *xa, const int64_t *xe, const double *ya,
                    const int64_t *ye, const double *za, const int64_t *ze);
double or_masked_logdot(double *util, int n_rates, int n_sites, int n_states,
                        const double *xa, const int64_t *xe, const double *ya,
                        const int64_t *ye, const uint16_t *mask);
double or_logdot(double *util, int n_rates, int n_sites, int n_states,
                 const double *xa, const int64_t *xe, const double *ya,
                 const int64_t *ye);
double or_slow_log_like3(const or_model *m, int n_sites,
                         const double *xa, const int64_t *xe, const double *ya,
                         const int64_t *ye, const double *za, const int64_t *ze);

/* ---- reference-tree precompute (like_stree.ml, glv_arr.ml) ------------- */
typedef struct {
  int n_nodes;             /* G; ids are post-order, root = G-1 */
  const int *child_start;  /* [G+1] CSR into children */
  const int *children;     /* children of each node in Newick order */
  const double *bl;        /* [G] branch lengths, root = 0 */
} or_tree;

/* leaf_seqs: [G][n_sites] characters (rows of internal nodes are ignored).
 * Arrays darr/parr/snodes are [G][R][N][S]; exponent arrays [G][N]. */
int or_calc_distal_and_proximal(const or_model *m, const or_tree *t,
                                const char *leaf_seqs, int n_sites,
                                double *darr, int64_t *darr_e,
                                double *parr, int64_t *parr_e,
                                double *util_arr, int64_t *util_arr_e);
void or_pull_exponents_arr(int min_twoexp, double *arr, int64_t *arr_e, int n_glvs,
                           int n_rates, int n_sites, int n_states);
void or_prep_supernodes(const or_model *m, const or_tree *t, int n_sites,
                        const double *darr, const int64_t *darr_e,
                        const double *parr, const int64_t *parr_e,
                        double *snodes, int64_t *snodes_e);

/* ---- Minimization (minimization.ml + GSL 2.x Brent restatement) -------- */
typedef double (*or_fun)(double x, void *ctx);
int or_brent(or_fun f, void *ctx, double raw_start, double left, double right,
             double tolerance, int max_iters, double *result);
/* The GSL minimiser's state (gsl_min_fminimizer + brent_state_t) and its
 * set / iterate steps, exposed so that tests can drive the raw iteration. */
typedef struct {
  double x_lower, x_upper, x_minimum, f_lower, f_upper, f_minimum;
  double v, w, d, e, f_v, f_w;
} or_gsl_brent;
int or_gsl_brent_set(or_gsl_brent *s, or_fun f, void *ctx, double x_minimum,
                     double x_lower, double x_upper);
int or_gsl_brent_set_with_values(or_gsl_brent *s, or_fun f, void *ctx, double x_minimum,
                                 double f_minimum, double x_lower, double f_lower,
                                 double x_upper, double f_upper);
int or_gsl_brent_iterate(or_gsl_brent *s, or_fun f, void *ctx);

/* ---- Integration (integration.ml:28: GSL 2.x gsl_integration_qng, the
 * QUADPACK QNG method) ------------------------------------------------------ */
enum { OR_QNG_OK = 0, OR_QNG_ETOL = 1, OR_QNG_EBADTOL = 2, OR_QNG_ABORT = 3 };
/* *stop (may be NULL) is read after every evaluation of f; a non-zero value
 * abandons the rule (an exception raised inside f). */
int or_qng(or_fun f, void *ctx, double a, double b, double epsabs, double epsrel,
           double *result, double *abserr, int *neval, const int *stop);

/* ---- pendant-length priors (core.ml:54-56, 86-92; pplacer_run.ml:380-389) */
enum { OR_PRIOR_FLAT_EXP = 0, OR_PRIOR_UNIFORM = 1, OR_PRIOR_INFORMATIVE = 2 };
double or_exponential_pdf(double x, double mu);

/* ---- placement core (core.ml, three_tax.ml, fig.ml) --------------------- */
typedef struct {
  double start_pend, max_pend, initial_tolerance, strike_box, keep_factor;
  int max_strikes, max_pitches, keep_at_most;
  /* -p (prefs.ml:198-203, 216-223): posterior probabilities of the kept
   * placements from their marginal likelihoods (Three_tax.calc_marg_prob) */
  int calc_pp;
  double pp_rel_err;
  int prior_kind;              /* OR_PRIOR_* */
  double prior_mean;           /* flat exponential prior mean */
  const double *prior_means;   /* informative: mean per edge id */
} or_prefs;

typedef struct {
  const or_model *model;
  const or_tree *tree;
  int n_sites;              /* N after pre-mask */
  const double *darr; const int64_t *darr_e;
  const double *parr; const int64_t *parr_e;
  const double *snodes; const int64_t *snodes_e;
  or_prefs prefs;
} or_core;

typedef struct {
  int loc;
  double ml_ratio, log_like, dist_bl, pend_bl;
  double post_prob, marginal_prob;  /* with calc_pp (placement.ml:62-65) */
} or_placement;

typedef struct {
  long n_logdots;     /* stage-1 scores computed */
  long n_like_calls;  /* stage-2 log_like3 evaluations (Three_tax.n_like_calls) */
  int n_pitched;      /* locations Brent-optimised in baseball */
  int n_skipped;      /* locations dropped on a GSL error */
  int n_max_iter;     /* ExceededMaxIter events */
  double t_ranking, t_ml;
} or_stats;

int or_stage1_scores(const or_core *c, const char *query, double *scores);
int or_rank(const double *scores, int n_edges, int *order);
int or_three_tax_optimize(const or_core *c, const char *query, int loc, double pend,
                          double distal, double tolerance, double *ll,
                          double *pend_out, double *dist_out, long *n_calls);
int or_process_query(const or_core *c, const char *query, or_placement *out,
                     int *n_out, or_stats *stats);
int or_process_queries(const or_core *c, const char *queries, int n_queries,
                       int n_threads, or_placement *out /* [Q][keep_at_most] */,
                       int *n_out, int *status, or_stats *stats);
void or_ll_normalized_prob(const double *ll, int n, double *out);
/* Three_tax.log_like (three_tax.ml:30-36) at (loc, pend, distal) */
int or_three_tax_log_like(const or_core *c, const char *query, int loc, double pend,
                          double distal, double *ll);
/* Three_tax.calc_marg_prob (three_tax.ml:169-215) for the query at loc with
 * the given branch lengths, with the prior of c->prefs at loc */
int or_calc_marg_prob(const or_core *c, const char *query, int loc, double pend,
                      double distal, double *marg, long *n_calls);

#ifdef __cplusplus
}
#endif
#endif
