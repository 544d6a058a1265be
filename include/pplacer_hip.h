/*
 * pplacer_hip.h -- C ABI of libpplacer_hip.so, the MI355X (gfx950) placement
 * path of pplacer.  Plain pointers and sizes only; every entry point returns 0
 * on success or a PP_E* code, and pp_last_error() gives the thread's message.
 *
 * Two groups of entry points:
 *
 *  1. Per-call twins of the C stubs that pplacer_src/linear.ml:13-89 binds
 *     (`external ... = "ten_masked_logdot_c"` etc.).  Same arguments, same
 *     arithmetic, host buffers in and out; the work runs on the GPU.  A
 *     maintainer re-points each `external` at a thin OCaml stub that unpacks
 *     the Bigarrays and calls the twin (INTEGRATION.md has the stub file).
 *
 *  2. A batched placement API replacing the per-query loop of
 *     Core.pplacer_core (pplacer_src/core.ml:76-411) as driven by
 *     Pplacer_run.run_placements (pplacer_src/pplacer_run.ml:167-656):
 *     model build, reference-tree precompute, stage-1 ranking of every
 *     (query, edge) pair, stage-2 baseball + Brent branch-length
 *     optimisation, keep/refine and like-weight ratios, all device-resident.
 *
 * Layout conventions follow pplacer's GLVs (pplacer_src/linear_c.c:5-12): a
 * GLV is double a[rate][site][state] (c_layout) plus one exponent per site.
 */
#ifndef PPLACER_HIP_H
#define PPLACER_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PP_OK = 0,
  PP_EINVAL = 1,     /* bad arguments or dimensions */
  PP_EDEVICE = 2,    /* HIP runtime error */
  PP_ENOMEM = 3,     /* device allocation failed */
  PP_ESTATE = 4,     /* call out of order (e.g. no model / tree yet) */
  PP_EQUERY = 5      /* a query failed the way core.ml fails it (see status) */
};

/* per-query status of the batched API (mirrors the exceptions core.ml lets
 * escape from process_query) */
enum {
  PP_Q_OK = 0,
  PP_Q_NO_INFORMATIVE = 1, /* "sequence has no informative sites" core.ml:157-158 */
  PP_Q_BAD_CHAR = 2,       /* "not a known base" nuc_models.ml:49-54 */
  PP_Q_NO_RESULTS = 3,     /* "empty results" core.ml:319-321 */
  PP_Q_FIND_START = 4,     /* Minimization.FindStartFailure */
  PP_Q_INVALID_START = 5,  /* Minimization.InvalidStartValues */
  PP_Q_ASSERT = 6,         /* assert(prox_bl >= 0.) three_tax.ml:43 */
  PP_Q_RANK_EXHAUSTED = 7, /* query too long for the GPU path: more masked sites than a
                              4-state stage-2 workgroup's LDS indexes (24 bytes per site
                              of 160 KB in the FP64 form, about 6,500); no counterpart in
                              the reference */
  PP_Q_MAXITER = 8,        /* Minimization.ExceededMaxIter outside the optimiser (-p) */
  PP_Q_GSL = 9             /* an uncaught Gsl_exn (-p: EBADFUNC, EBADTOL) */
};

/* Pendant-length priors for -p (Core.prior, core.ml:54-56) */
enum { PP_PRIOR_FLAT_EXP = 0, PP_PRIOR_UNIFORM = 1, PP_PRIOR_INFORMATIVE = 2 };

/* Precision forms of a placement pass.  PP_FP64: the reference method in
 * FP64, each probe's per-site logs summed in site order as the stubs do.
 * PP_FP32: stage 1 and the 4-state baseball phase on FP32 GLVs with per-site
 * power-of-two exponents and FP64 log sums; the refinement of the kept
 * placements (final_tolerance), their likelihoods, ratios and output in FP64.
 * 20-state models run stage 2 in FP64 in both forms. */
enum { PP_FP64 = 0, PP_FP32 = 1 };

const char *pp_last_error(void);

/* ======================================================================= */
/* 1. linear_c.c twins (host buffers; GPU compute)                          */
/* ======================================================================= */

/* linear_c.c:29-90 gemmish_c: dst[site][i] = sum_j a[i][j] * b[site][j] */
int pp_gemmish(double *dst, const double *a, const double *b, int n_sites, int n_states);

/* linear_c.c:92-146 dediagonalize: dst[i][j] = sum_k lambda[k] u[i][k] uit[j][k] */
int pp_dediagonalize(double *dst, const double *u, const double *lambda, const double *uit,
                     int n);

/* linear_c.c:367-424 ten_log_like3_c -> *out; util[n_sites] receives the
 * per-site sums as the stub leaves them */
int pp_ten_log_like3(const double *statd, const double *x, const double *y, const double *z,
                     double *util, int n_rates, int n_sites, int n_states, double *out);

/* linear_c.c:426-441 ten_pairwise_prod_c */
int pp_ten_pairwise_prod(double *dst, const double *x, const double *y, int n_rates,
                         int n_sites, int n_states);

/* linear_c.c:443-463 ten_statd_pairwise_prod_c */
int pp_ten_statd_pairwise_prod(const double *statd, double *dst, const double *a,
                               const double *b, int n_rates, int n_sites, int n_states);

/* linear_c.c:465-539 ten_masked_logdot_c -> *out */
int pp_ten_masked_logdot(const double *x, const double *y, const uint16_t *mask, double *util,
                         int n_rates, int n_sites, int n_states, double *out);

/* linear_c.c:541-605 ten_bounded_logdot_c -> *out */
int pp_ten_bounded_logdot(const double *x, const double *y, int first, int last, double *util,
                          int n_rates, int n_sites, int n_states, double *out);

/* Allocation-free variant of pp_ten_masked_logdot.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * stream capture, serving loops).  x, y ([n_rates][n_sites][n_states]), mask
 * ([n_sites]) and out (one double) are device pointers; the work is enqueued on
 * `stream` (a hipStream_t, NULL for the default stream) and *out is written
 * when the stream reaches it.  The per-site util vector lives in a static device
 * array of 262,144 doubles per GPU (2 MiB, held until the process ends), so
 * n_sites above that returns PP_EINVAL; calls that overlap on one GPU share that
 * array and race. */
int pp_ten_masked_logdot_noalloc(const double *x, const double *y, const uint16_t *mask,
                                 int n_rates, int n_sites, int n_states, double *out,
                                 void *stream);

/* The mat_* stubs of the Price-CAT model's rate-free [site][state] GLVs
 * (linear_c.c:148-340, bound in linear.ml:28-49): the same arithmetic with one
 * rate and no -n log(n_rates) term. */
/* linear_c.c:168-213 mat_log_like3_c -> *out */
int pp_mat_log_like3(const double *statd, const double *x, const double *y, const double *z,
                     int n_sites, int n_states, double *out);
/* linear_c.c:215-228 mat_pairwise_prod_c */
int pp_mat_pairwise_prod(double *dst, const double *x, const double *y, int n_sites, int n_states);
/* linear_c.c:230-248 mat_statd_pairwise_prod_c */
int pp_mat_statd_pairwise_prod(const double *statd, double *dst, const double *a, const double *b,
                               int n_sites, int n_states);
/* linear_c.c:250-293 mat_masked_logdot_c -> *out */
int pp_mat_masked_logdot(const double *x, const double *y, const uint16_t *mask, int n_sites,
                         int n_states, double *out);
/* linear_c.c:295-340 mat_bounded_logdot_c -> *out */
int pp_mat_bounded_logdot(const double *x, const double *y, int first, int last, int n_sites,
                          int n_states, double *out);

/* ======================================================================= */
/* 2. batched placement                                                     */
/* ======================================================================= */

typedef struct pp_ctx pp_ctx;

/* One context owns one GPU's resident model, tree GLVs and work buffers.
 * Create it after any fork() (pplacer forks children, pplacer_run.ml:628). */
int pp_ctx_create(int device, pp_ctx **out);
int pp_ctx_destroy(pp_ctx *ctx);

/* Glvm.Model build for the gamma mixture model (gmix_model.ml:25-37):
 * exch = symmetric exchangeabilities [S][S] (diagonal ignored), statd [S],
 * rates [R] (Gamma.discrete_gamma).  Diagonalises Q = B D in FP64
 * (diagd.ml:123-172). */
int pp_model_build(pp_ctx *ctx, int n_states, int n_rates, const double *exch,
                   const double *statd, const double *rates);
/* read back the diagonalisation: x [S][S], l [S], xit [S][S] */
int pp_model_get(pp_ctx *ctx, double *x, double *l, double *xit);

/* Price-CAT model (gcat_model.ml:33-95, refpkg.ml:39-63 "Price-CAT" block):
 * rate-free GLVs, each site evolving by the matrix of its category's rate.
 * The categories are those of the pre-masked sites (Gcat_model.mask_sites,
 * gcat_model.ml:429-439, after the 0-based shift of init_of_json :475) and
 * must be set before pp_tree_set / pp_tree_upload. */
int pp_model_build_cat(pp_ctx *ctx, int n_states, int n_cats, const double *exch,
                       const double *statd, const double *cat_rates);
int pp_site_categories(pp_ctx *ctx, int n_sites, const int *site_cat);

/* Gamma.discrete_gamma (gamma.ml:52-70) */
int pp_discrete_gamma(int n_cat, double alpha, double *rates);

/* Alignment.emper_freq (alignment.ml:186-217) over [n_seq][n_sites] chars */
int pp_emper_freq(int n_states, const char *seqs, int n_seq, int n_sites, double *freq);

/* Pplacer_run.premask (pplacer_run.ml:59-144): mask[N0] of columns kept,
 * query_keep[Q], *masklen */
int pp_premask(int n_states, const char *ref_seqs, int n_ref, const char *query_seqs,
               int n_query, int n_sites, int discard_nonoverlapped, uint8_t *mask,
               uint8_t *query_keep, int *masklen);

/* Reference tree + leaf sequences (already pre-masked, n_sites columns).
 * Node ids are pplacer's post-order ids (newick_parser.mly:49-62); root is
 * n_nodes-1 with bl 0.  child_start[n_nodes+1]/children[] list each node's
 * children in Newick order; leaf_seqs is [n_nodes][n_sites] (rows of internal
 * nodes ignored).  Runs Like_stree.calc_distal_and_proximal, the -10
 * exponent pull and Glv_arr.prep_supernodes (pplacer_run.ml:321-341) on the
 * GPU and keeps darr/parr/snodes resident. */
int pp_tree_set(pp_ctx *ctx, int n_nodes, const int *child_start, const int *children,
                const double *bl, int n_sites, const char *leaf_seqs);

/* Alternative to pp_tree_set for a host that computed the GLVs itself (the
 * OCaml host's darr/parr/snodes after prep_supernodes): arrays are
 * [G][R][N][S] doubles and [G][N] int64 exponents. */
int pp_tree_upload(pp_ctx *ctx, int n_nodes, const double *bl, int n_sites,
                   const double *darr, const int64_t *darr_e, const double *parr,
                   const int64_t *parr_e, const double *snodes, const int64_t *snodes_e);

/* Copy the resident GLVs back in pplacer's layout ([G][R][N][S], [G][N]). */
int pp_tree_download(pp_ctx *ctx, double *darr, int64_t *darr_e, double *parr,
                     int64_t *parr_e, double *snodes, int64_t *snodes_e);

/* Prefs that steer Core.pplacer_core (prefs.ml:65-121 defaults in brackets). */
typedef struct {
  double start_pend;        /* 0.1  */
  double max_pend;          /* 2.0  */
  double initial_tolerance; /* 0.01 */
  double strike_box;        /* 3.0  */
  double keep_factor;       /* 0.01 */
  int max_strikes;          /* 6 (0 = evaluate every edge) */
  int max_pitches;          /* 40 */
  int keep_at_most;         /* 7 */
  /* -p (prefs.ml:198-203): marginal likelihoods and posterior probabilities
   * of the kept placements.  -p also sets keep_at_most 20, keep_factor 0.001
   * and max_strikes 20; the caller sets those fields as the CLI does. */
  int calc_pp;              /* 0 */
  double pp_rel_err;        /* 0.01 (--pp-rel-err) */
} pp_prefs;

void pp_prefs_default(pp_prefs *p);

/* Placement.make_ml fields (placement.ml:50-60) */
typedef struct {
  int loc;          /* edge_num */
  double ml_ratio;  /* like_weight_ratio */
  double log_like;  /* likelihood */
  double dist_bl;   /* distal_length */
  double pend_bl;   /* pendant_length */
  double post_prob;      /* post_prob with calc_pp, else NaN (placement.ml:62-65) */
  double marginal_prob;  /* marginal_like with calc_pp, else NaN */
} pp_placement;

typedef struct {
  double t_ranking_ms;   /* stage 1 + rank ("ranking" timing, core.ml:212) */
  double t_ml_ms;        /* stage 2 ("ML calculation", core.ml:322) */
  int64_t n_logdots;     /* (query, edge) stage-1 scores */
  int64_t n_like_calls;  /* Model.log_like3 calls as the reference makes them
                            (Three_tax.n_like_calls), the GSL minimiser's set
                            re-evaluations included; the probes actually run are
                            counted by n_site_probes_* */
  int64_t n_pitched;     /* Brent-optimised (query, edge) pairs */
  /* per-kernel device time (HIP events on the library's stream) */
  double t_stage1_ms;    /* stage-1 scoring kernels (tile codes + scores) */
  double t_pitch_ms;     /* baseball + Brent kernels */
  double t_refine_ms;    /* keep + final refinement + finalize kernels */
  /* algorithmic work: sums over evaluations of the masked length M */
  int64_t n_site_pairs;       /* sum over (query, edge) of M  (stage 1) */
  int64_t n_site_probes_pend; /* sum over pendant-only probes of M (stage 2) */
  int64_t n_site_probes_dist; /* sum over probes that re-evolve prox/dist of M */
  int64_t n_launches;         /* kernel launches issued */
  double t_table_ms;          /* stage-1 log-util table build (part of the ranking time) */
  int64_t n_site_probes_pend_pitch; /* n_site_probes_pend of the baseball kernels alone */
  int64_t n_site_probes_dist_pitch; /* n_site_probes_dist of the baseball kernels alone */
  int64_t n_pitch_launches;         /* launches of the baseball kernel */
  double t_pp_ms;                   /* -p marginal likelihoods ("PP calculation", core.ml:380) */
  double t_pitch_kernel_ms;         /* sum over baseball launches of each launch's own
                                       duration (launches on concurrent streams overlap,
                                       so this can exceed t_pitch_ms) */
  /* FP32 form: queries that failed the prox_bl >= 0 assertion (three_tax.ml:41-45)
   * and were placed again in FP64, and the device time of that re-placement
   * (its ranking and ML time, also counted in t_ml_ms; the per-kernel times
   * and work counters above describe the FP32 pass alone) */
  int64_t n_fp64_replaced;
  double t_fp64_replace_ms;
} pp_stats;

/* Stage 1 for a batch: scores[q][e] = Glv.masked_logdot of the query (full
 * length, evolved by start_pend) against supernode e, for every non-root
 * edge e in 0..n_nodes-2 (core.ml:131-141,187-194).  queries is
 * [n_queries][n_sites] characters (upper case). */
int pp_stage1_scores(pp_ctx *ctx, int n_queries, const char *queries, double start_pend,
                     int precision, double *scores);

/* Stage-1 ranking: for each query the first k edges of the stable
 * descending sort of its scores (fig.ml:143-147; nan last, ties by edge id). */
int pp_rank_topk(pp_ctx *ctx, int n_queries, int n_edges, const double *scores, int k,
                 int *ranked);

/* Three_tax.optimize from a given start (core.ml:225-281): one job per
 * (query, loc, pend0, distal0); results are the last-probe state
 * (ll, pend, distal) as Three_tax.get_results reports it. */
int pp_optimize_batch(pp_ctx *ctx, int n_jobs, const char *queries, const int *query_of_job,
                      const int *locs, const double *pend0, const double *dist0,
                      double tolerance, double max_pend, int precision, double *ll,
                      double *pend, double *dist, int *n_calls, int *status);

/* The prior Core.pplacer_core is given (pplacer_run.ml:380-389): kind
 * PP_PRIOR_UNIFORM (means unused), PP_PRIOR_FLAT_EXP (means[0] = the mean),
 * PP_PRIOR_INFORMATIVE (means[n_nodes] = the mean of each edge id, from
 * pp_prior_mean_map).  pp_tree_set installs the reference's default, the flat
 * exponential with mean pp_tree_length / (n_nodes - 1). */
int pp_set_prior(pp_ctx *ctx, int kind, int n, const double *means);

/* Gtree.tree_length (gtree.ml:102-114) and Core.prior_mean_map with
 * top_segment_f = (+ lower) (core.ml:34-52) on a tree given as in
 * pp_tree_set; host-only.  means[n_nodes] (the root entry is unused). */
int pp_tree_length(int n_nodes, const int *child_start, const int *children, const double *bl,
                   double *length);
int pp_prior_mean_map(int n_nodes, const int *child_start, const int *children,
                      const double *bl, double lower, double *means);

/* Three_tax.calc_marg_prob (three_tax.ml:169-215) per job: the marginal
 * log-likelihood of the query at (loc, pend, distal) under the context's
 * prior, with relative tolerance rel_err and the pendant range capped at
 * max_pend; status is 0 or a Minimization / GSL failure code. */
int pp_marginal_batch(pp_ctx *ctx, int n_jobs, const char *queries, const int *query_of_job,
                      const int *locs, const double *pend, const double *dist,
                      double rel_err, double max_pend, double *marg, int *n_calls,
                      int *status);

/* Full Core.process_query for a batch of queries: stage 1, ranking,
 * baseball with Brent at initial_tolerance, keep, refine at 1e-5,
 * ll_normalized_prob, with calc_pp the marginal likelihoods and posterior
 * probabilities (core.ml:378-403), and the ml_ratio sort.  out is
 * [n_queries][max(1, keep_at_most)], n_out[q] placements each; status[q] is a
 * PP_Q_* code.  Returns PP_EQUERY, with every output filled, when some query
 * failed.  Equivalent to pp_queries_upload + pp_place_resident +
 * pp_results_download: the batch stays resident afterwards. */
int pp_place_batch(pp_ctx *ctx, int n_queries, const char *queries, const pp_prefs *prefs,
                   int precision, pp_placement *out, int *n_out, int *status,
                   pp_stats *stats);

/* Device-resident variant for throughput measurement and pipelines: the
 * queries are uploaded once (pp_queries_upload), placed by pp_place_resident
 * and fetched with pp_results_download.  In the 4-state FP32 form, the
 * queries of a pass that fail on the prox_bl assertion (PP_Q_ASSERT; FP32
 * likelihoods steer Brent along other paths) are placed again in the FP64
 * form inside pp_place_resident (one status download and a pass over those
 * queries alone; its work and time are added to stats), so both forms fail
 * exactly the queries the reference fails. */
int pp_queries_upload(pp_ctx *ctx, int n_queries, const char *queries);
int pp_place_resident(pp_ctx *ctx, const pp_prefs *prefs, int precision, pp_stats *stats);
int pp_results_download(pp_ctx *ctx, pp_placement *out, int *n_out, int *status);
int pp_synchronize(pp_ctx *ctx);

/* Launch shape of a placement pass; no counterpart in the reference (its
 * per-query loop has no launches).  For tuners: results do not depend on it.
 *   pitch_streams[f], refine_streams[f]: streams over which the 4-state
 *     baseball phase and the refinement phase spread their masked-length
 *     buckets, for form f (0 = PP_FP64, 1 = PP_FP32); 0 runs them one after
 *     another on the main stream; at most PP_MAX_BUCKET_STREAMS.  The -p
 *     phase and 20-state passes always run on the main stream.
 *   chunk: queries per device chunk (stage 1 of the next chunk overlaps
 *     stage 2 of the current one), 1 .. PP_MAX_CHUNK, further capped so that a
 *     chunk's scores ([chunk][n_edges] doubles) stay under 2 GiB.
 *   graph: 1 captures the pass's stage-1 and stage-2 launches (every chunk,
 *     every stream) into a HIP graph and replays it on the next passes with
 *     the same launch plan (resident queries, prefs, precision, shape and
 *     buffers); a pass with another plan captures again.  With graph = 1
 *     pp_stats holds the whole graph's device time in t_ml_ms and no
 *     per-phase or per-launch times.  0 (default) launches eagerly.
 * pp_launch_config_set returns PP_EINVAL, changing nothing, for a value out
 * of range. */
#define PP_MAX_BUCKET_STREAMS 8
#define PP_MAX_CHUNK 262144
typedef struct {
  int pitch_streams[2];
  int refine_streams[2];
  int chunk;
  int graph;
} pp_launch_config;
int pp_launch_config_get(pp_ctx *ctx, pp_launch_config *out);
int pp_launch_config_set(pp_ctx *ctx, const pp_launch_config *cfg);

/* ======================================================================= */
/* Device GLVs: Glvm.Model's Glv for the gamma mixture model               */
/* ======================================================================= */
/* Gmix_model.Model.Glv (gmix_model.ml:58-257, the Glv signature of
 * glvm.ml:28-45) with the GLV held on the GPU: a[n_rates][n_sites][n_states]
 * doubles and e[n_sites] pulled base-2 exponents.  An OCaml Glv.t becomes a
 * custom block around a pp_glv*, so Like_stree / Three_tax code that calls
 * Glv.* and Model.* keeps its data in HBM between calls.  Every entry runs on
 * the context's main stream in call order; those that return values to the
 * host wait for them.  The model entries (pp_gmix_*) use the context's model
 * (pp_model_build, gamma rates; a Price-CAT model gives PP_EINVAL) and need
 * GLVs of its rates and states. */
typedef struct pp_glv pp_glv;
/* OCaml's fpclass, in its order: Glv.fp_classify returns the worst class */
enum { PP_FP_NORMAL = 0, PP_FP_SUBNORMAL = 1, PP_FP_ZERO = 2, PP_FP_INFINITE = 3, PP_FP_NAN = 4 };

/* Glv.make (gmix_model.ml:76-79), contents zero.  A GLV lives until
 * pp_glv_free or until its context's pp_ctx_destroy, which frees the GLVs
 * still made on it (their handles are invalid afterwards). */
int pp_glv_make(pp_ctx *ctx, int n_rates, int n_sites, int n_states, pp_glv **out);
int pp_glv_free(pp_glv *g);
/* get_n_rates, get_n_sites, get_n_states (gmix_model.ml:63-69) */
int pp_glv_dims(const pp_glv *g, int *n_rates, int *n_sites, int *n_states);
/* Glv.mimic, Glv.copy, Glv.memcpy, Glv.set_unit (gmix_model.ml:82-99) */
int pp_glv_mimic(const pp_glv *x, pp_glv **out);
int pp_glv_copy(const pp_glv *x, pp_glv **out);
int pp_glv_memcpy(pp_glv *dst, const pp_glv *src);
int pp_glv_set_unit(pp_glv *g);
/* Glv.fp_classify (gmix_model.ml:103-104, tensor.ml:43-44) -> PP_FP_* */
int pp_glv_fp_classify(const pp_glv *g, int *cls);
/* the GLV's contents from / to host arrays shaped as its Bigarrays
 * (a [n_rates][n_sites][n_states], e [n_sites] OCaml ints) */
int pp_glv_set(pp_glv *g, const double *a, const int64_t *e);
int pp_glv_get(const pp_glv *g, double *a, int64_t *e);
/* Glv.mask_into (gmix_model.ml:129-141): site_mask[n_mask] (n_mask = src's
 * sites) keeps dst's number of sites, in order */
int pp_glv_mask_into(const uint8_t *site_mask, int n_mask, const pp_glv *src, pp_glv *dst);
/* Glv.prep_constant_rate_glv_from_lv_arr (gmix_model.ml:149-156):
 * lv_arr [n_sites][n_states] into every rate, exponents 0 */
int pp_glv_prep_constant_rate_from_lv_arr(pp_glv *g, const double *lv_arr);
/* Glv.perhaps_pull_exponent (gmix_model.ml:163-187) */
int pp_glv_perhaps_pull_exponent(int min_allowed_twoexp, pp_glv *g);
/* Glv.masked_logdot, bounded_logdot, logdot (gmix_model.ml:194-212), the
 * exponent terms included; the util workspace is the library's */
int pp_glv_masked_logdot(const pp_glv *x, const pp_glv *y, const uint16_t *mask, double *out);
int pp_glv_bounded_logdot(const pp_glv *x, const pp_glv *y, int start, int last, double *out);
int pp_glv_logdot(const pp_glv *x, const pp_glv *y, double *out);
/* Glv.pairwise_prod and listwise_prod (gmix_model.ml:215-233) */
int pp_glv_pairwise_prod(pp_glv *dst, const pp_glv *x, const pp_glv *y);
int pp_glv_listwise_prod(pp_glv *dst, const pp_glv *const *glvs, int n);
/* Glv.summarize_post (gmix_model.ml:242-256): u[n_sites][n_states] = the
 * per-site state sums over rates; the caller applies summarize_f */
int pp_glv_summarize_post(const pp_glv *g, double *u);

/* Model.log_like3 (gmix_model.ml:298-304) */
int pp_gmix_log_like3(pp_ctx *ctx, const pp_glv *x, const pp_glv *y, const pp_glv *z, double *out);
/* Model.site_log_like_arr3 (gmix_model.ml:323-345): out[n_sites];
 * PP_EQUERY ("Site %d has zero likelihood.") where a site's is not positive */
int pp_gmix_site_log_like_arr3(pp_ctx *ctx, const pp_glv *x, const pp_glv *y, const pp_glv *z,
                               double *out);
/* Model.slow_log_like3 (gmix_model.ml:347-348) */
int pp_gmix_slow_log_like3(pp_ctx *ctx, const pp_glv *x, const pp_glv *y, const pp_glv *z, double *out);
/* Model.evolve_into (gmix_model.ml:308-315): dst = src evolved for branch
 * length bl by the model's P(bl) of every rate (Diagd.multi_exp) */
int pp_gmix_evolve_into(pp_ctx *ctx, pp_glv *dst, const pp_glv *src, double bl);
/* Model.statd_pairwise_prod (gmix_model.ml:319-321) */
int pp_gmix_statd_pairwise_prod(pp_ctx *ctx, pp_glv *dst, const pp_glv *a, const pp_glv *b);

#ifdef __cplusplus
}
#endif
#endif
