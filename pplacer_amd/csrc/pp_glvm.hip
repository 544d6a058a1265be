// Device-resident GLVs of the gamma mixture model: the Glv submodule of
// Gmix_model.Model (gmix_model.ml:58-257, a Glvm.Model, glvm.ml:11-55) and the
// model functions that act on them (gmix_model.ml:296-351), with the GLV data
// living in HBM behind an opaque handle.  Included at the end of pp_capi.hip
// (it uses pp_ctx and the twin kernels).
//
// Every kernel goes to the context's main stream in call order, so a chain of
// evolve / product / pull calls runs without host round trips; the entries
// that return values to the host (logdots, log likelihoods, fp_classify,
// pp_glv_get, summarize_post) wait for their result.  The arithmetic is the
// stubs' and the OCaml's, operation for operation.

#include <cfloat>
#include <climits>

struct pp_glv {
  pp_ctx *ctx = nullptr;
  int R = 0, N = 0, S = 0;
  double *a = nullptr;  // [R][N][S]
  int *e = nullptr;     // [N] pulled base-2 exponents
};

namespace {
namespace gv {

__global__ void fill_kernel(double *a, long n, double v) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) a[t] = v;
}

__global__ void fill_int_kernel(int *e, long n, int v) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) e[t] = v;
}

// OCaml's fpclass order: FP_normal < FP_subnormal < FP_zero < FP_infinite < FP_nan
__global__ void classify_kernel(const double *a, long n, int *worst) {
  int w = PP_FP_NORMAL;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const double v = a[t];
    const int c = v != v                ? PP_FP_NAN
                  : __builtin_isinf(v)  ? PP_FP_INFINITE
                  : v == 0.0            ? PP_FP_ZERO
                  : fabs(v) < DBL_MIN   ? PP_FP_SUBNORMAL
                                        : PP_FP_NORMAL;
    w = c > w ? c : w;
  }
  if (w > PP_FP_NORMAL) atomicMax(worst, w);
}

// dst site j <- src site idx[j] (Glv.copy_site for every kept site)
__global__ void gather_sites_kernel(double *da, int *de, const double *sa, const int *se, const int *idx, int R,
                                    int Ns, int Nd, int S) {
  const long n = (long)R * Nd * S;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const int r = (int)(t / ((long)Nd * S));
    const long rem = t % ((long)Nd * S);
    const int j = (int)(rem / S), s = (int)(rem % S);
    da[t] = sa[((long)r * Ns + idx[j]) * S + s];
    if (r == 0 && s == 0) de[j] = se[idx[j]];
  }
}

// every rate's slice <- lv[site][state] (gmix_model.ml:149-156)
__global__ void constant_rate_kernel(double *a, const double *lv, int R, int N, int S) {
  const long n = (long)R * N * S, ns = (long)N * S;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    a[t] = lv[t % ns];
}

// perhaps_pull_exponent (gmix_model.ml:163-187), one lane per site
__global__ void pull_kernel(double *a, int *e, int R, int N, int S, int min_allowed_twoexp) {
  const int site = blockIdx.x * blockDim.x + threadIdx.x;
  if (site >= N) return;
  long max_twoexp = -4611686018427387903L;  // OCaml's -max_int
  for (int r = 0; r < R; r++) {
    const double *s = a + ((long)r * N + site) * S;
    double mx = s[0];  // Gsl.Vector.max: the first maximum under >
    for (int k = 1; k < S; k++)
      if (s[k] > mx) mx = s[k];
    int te = 0;
    (void)frexp(mx, &te);
    if (te > max_twoexp) max_twoexp = te;
  }
  if (max_twoexp < min_allowed_twoexp) {
    const double f = ldexp(0.5, (int)-max_twoexp + 1);  // of_twoexp (gstar_support.ml:44)
    for (int r = 0; r < R; r++) {
      double *s = a + ((long)r * N + site) * S;
      for (int k = 0; k < S; k++) s[k] *= f;
    }
    e[site] += (int)max_twoexp;
  }
}

__global__ void add_exponents_kernel(int *d, const int *x, const int *y, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) d[t] = x[t] + y[t];
}

// tensor_mul (gstar_support.ml:66-74): rate r's slice by P_r with gemmish's
// order (linear_c.c:29-90): dst[site][i] = sum over j of P_r[i][j] src[site][j]
__global__ void evolve_kernel(double *dst, const double *src, const double *P, int R, int N, int S) {
  const long n = (long)R * N * S;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const int r = (int)(t / ((long)N * S));
    const int i = (int)(t % S);
    const double *p = P + ((long)r * S + i) * S;
    const double *x = src + (t - i);
    double acc = 0;
    for (int j = 0; j < S; j++) acc = acc + p[j] * x[j];
    dst[t] = acc;
  }
}

// Sums of the exponents as the OCaml accumulates them, one float addition per
// site in site order (total_twoexp / masked_total_twoexp /
// bounded_total_twoexp, gstar_support.ml:47-70), combined with the tensor
// part: out = t + log 2 * ((sum x + sum y) + sum z) (z optional).
__global__ void finish_logdot_kernel(const double *t, const int *xe, const int *ye, const int *ze,
                                     const uint16_t *mask, int first, int last, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double tx = 0., ty = 0., tz = 0.;
  for (int i = first; i <= last; i++) {
    if (mask && !mask[i]) continue;
    tx = tx + (double)xe[i];
    ty = ty + (double)ye[i];
    if (ze) tz = tz + (double)ze[i];
  }
  *out = *t + PP_LOG2 * (ze ? (tx + ty) + tz : tx + ty);
}

// site_log_like_arr3 (gmix_model.ml:323-345); a site whose likelihood is not
// positive is reported in *bad (the smallest such site)
__global__ void site_log_like_kernel(const double *statd, const double *x, const double *y, const double *z,
                                     const int *xe, const int *ye, const int *ze, int R, int N, int S,
                                     double *out, int *bad) {
  const int site = blockIdx.x * blockDim.x + threadIdx.x;
  if (site >= N) return;
  double like = 0.;
  for (int r = 0; r < R; r++)
    for (int s = 0; s < S; s++) {
      const long o = ((long)r * N + site) * S + s;
      like = like + statd[s] * x[o] * y[o] * z[o];
    }
  if (!(like > 0.)) atomicMin(bad, site);
  out[site] = log(like / (double)R) + PP_LOG2 * (double)(xe[site] + ye[site] + ze[site]);
}

// summarize_post's per-site state sums (gmix_model.ml:242-256)
__global__ void state_sums_kernel(const double *a, int R, int N, int S, double *u) {
  const long n = (long)N * S;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    double acc = 0.;
    for (int r = 0; r < R; r++) acc = acc + a[(long)r * n + t];
    u[t] = acc;
  }
}

int grid_for(long n) { return (int)std::max<long>(1, std::min<long>((n + 255) / 256, 8192)); }

bool same_dims(const pp_glv *x, const pp_glv *y) { return x->R == y->R && x->N == y->N && x->S == y->S; }

int glv_alloc(pp_ctx *ctx, int R, int N, int S, pp_glv **out) {
  CHECK(ctx && out, PP_EINVAL, "null argument");
  CHECK(R > 0 && N >= 0 && S > 0 && S <= kMaxStates, PP_EINVAL, "bad GLV dimensions");
  HIPCHECK(hipSetDevice(ctx->device));
  pp_glv *g = new pp_glv();
  g->ctx = ctx;
  g->R = R;
  g->N = N;
  g->S = S;
  const size_t na = std::max<size_t>((size_t)R * N * S, 1), ne = std::max<size_t>((size_t)N, 1);
  if (hipMalloc(&g->a, na * sizeof(double)) != hipSuccess || hipMalloc(&g->e, ne * sizeof(int)) != hipSuccess) {
    (void)hipFree(g->a);
    delete g;
    set_error("device allocation of a GLV failed");
    return PP_ENOMEM;
  }
  ctx->glvs.insert(g);
  *out = g;
  return PP_OK;
}

// The context's model as Gmix_model's: gamma rates, with the GLV's shape.
int check_gmix(const pp_ctx *c, const pp_glv *g) {
  CHECK(c->model.built, PP_ESTATE, "no model: call pp_model_build first");
  CHECK(c->model.n_cats == 0, PP_EINVAL, "the context's model is Price-CAT; these entries are Gmix_model's");
  CHECK(g->R == c->model.R && g->S == c->model.S, PP_EINVAL, "GLV dimensions differ from the model's");
  return PP_OK;
}

// One double computed on the device, fetched after the stream reaches it.
int fetch_double(pp_ctx *c, const double *d, double *out) {
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(out, d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return PP_OK;
}

int logdot_impl(const pp_glv *x, const pp_glv *y, const uint16_t *mask, int first, int last, double *out) {
  CHECK(x && y && out, PP_EINVAL, "null argument");
  CHECK(x->ctx == y->ctx, PP_EINVAL, "GLVs of different contexts");
  CHECK(same_dims(x, y), PP_EINVAL, "GLV dimensions differ");
  CHECK(first >= 0 && first <= last && last < x->N, PP_EINVAL, "bad site interval");
  pp_ctx *c = x->ctx;
  HIPCHECK(hipSetDevice(c->device));
  const int n_used = last - first + 1;
  Scratch s;
  uint16_t *dm = mask ? s.up(mask, (size_t)x->N) : nullptr;
  double *du = s.up<double>(nullptr, n_used), *dt = s.up<double>(nullptr, 2);
  CHECK(du && dt && (dm || !mask), PP_ENOMEM, "device allocation failed");
  hipLaunchKernelGGL(util_kernel, dim3((n_used + 255) / 256), dim3(256), 0, c->stream, du, nullptr, x->a, y->a,
                     nullptr, dm, first, n_used, x->R, x->N, x->S);
  hipLaunchKernelGGL(logsum_kernel, dim3(1), dim3(64), 0, c->stream, du, dm, first, n_used, x->R, 1, dt);
  hipLaunchKernelGGL(gv::finish_logdot_kernel, dim3(1), dim3(1), 0, c->stream, dt, x->e, y->e, nullptr, dm, first,
                     last, dt + 1);
  return fetch_double(c, dt + 1, out);
}

}  // namespace gv
}  // namespace

extern "C" {

int pp_glv_make(pp_ctx *ctx, int n_rates, int n_sites, int n_states, pp_glv **out) {
  int st = gv::glv_alloc(ctx, n_rates, n_sites, n_states, out);
  if (st) return st;
  pp_glv *g = *out;
  hipLaunchKernelGGL(gv::fill_kernel, dim3(gv::grid_for((long)g->R * g->N * g->S)), dim3(256), 0, ctx->stream, g->a,
                     (long)g->R * g->N * g->S, 0.0);
  hipLaunchKernelGGL(gv::fill_int_kernel, dim3(gv::grid_for(g->N)), dim3(256), 0, ctx->stream, g->e, (long)g->N, 0);
  HIPCHECK(hipGetLastError());
  return PP_OK;
}

int pp_glv_free(pp_glv *g) {
  if (!g) return PP_OK;
  if (g->ctx) {
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);  // pending work may still read it
    g->ctx->glvs.erase(g);
  }
  (void)hipFree(g->a);
  (void)hipFree(g->e);
  delete g;
  return PP_OK;
}

int pp_glv_dims(const pp_glv *g, int *n_rates, int *n_sites, int *n_states) {
  CHECK(g, PP_EINVAL, "null GLV");
  if (n_rates) *n_rates = g->R;
  if (n_sites) *n_sites = g->N;
  if (n_states) *n_states = g->S;
  return PP_OK;
}

int pp_glv_memcpy(pp_glv *dst, const pp_glv *src) {
  CHECK(dst && src && dst->ctx == src->ctx, PP_EINVAL, "null GLV or GLVs of different contexts");
  CHECK(gv::same_dims(dst, src), PP_EINVAL, "GLV dimensions differ");
  pp_ctx *c = src->ctx;
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipMemcpyAsync(dst->a, src->a, sizeof(double) * src->R * src->N * src->S, hipMemcpyDeviceToDevice,
                          c->stream));
  HIPCHECK(hipMemcpyAsync(dst->e, src->e, sizeof(int) * src->N, hipMemcpyDeviceToDevice, c->stream));
  return PP_OK;
}

int pp_glv_mimic(const pp_glv *x, pp_glv **out) {
  CHECK(x, PP_EINVAL, "null GLV");
  return pp_glv_make(x->ctx, x->R, x->N, x->S, out);
}

int pp_glv_copy(const pp_glv *x, pp_glv **out) {
  CHECK(x, PP_EINVAL, "null GLV");
  int st = gv::glv_alloc(x->ctx, x->R, x->N, x->S, out);
  if (st) return st;
  st = pp_glv_memcpy(*out, x);
  if (st) {
    pp_glv_free(*out);
    *out = nullptr;
  }
  return st;
}

int pp_glv_set_unit(pp_glv *g) {
  CHECK(g, PP_EINVAL, "null GLV");
  pp_ctx *c = g->ctx;
  HIPCHECK(hipSetDevice(c->device));
  hipLaunchKernelGGL(gv::fill_kernel, dim3(gv::grid_for((long)g->R * g->N * g->S)), dim3(256), 0, c->stream, g->a,
                     (long)g->R * g->N * g->S, 1.0);
  hipLaunchKernelGGL(gv::fill_int_kernel, dim3(gv::grid_for(g->N)), dim3(256), 0, c->stream, g->e, (long)g->N, 0);
  HIPCHECK(hipGetLastError());
  return PP_OK;
}

int pp_glv_fp_classify(const pp_glv *g, int *cls) {
  CHECK(g && cls, PP_EINVAL, "null argument");
  pp_ctx *c = g->ctx;
  HIPCHECK(hipSetDevice(c->device));
  Scratch s;
  const int zero = PP_FP_NORMAL;
  int *dw = s.up(&zero, 1);
  CHECK(dw, PP_ENOMEM, "device allocation failed");
  const long n = (long)g->R * g->N * g->S;
  hipLaunchKernelGGL(gv::classify_kernel, dim3(gv::grid_for(n)), dim3(256), 0, c->stream, g->a, n, dw);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(cls, dw, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return PP_OK;
}

int pp_glv_set(pp_glv *g, const double *a, const int64_t *e) {
  CHECK(g && a && e, PP_EINVAL, "null argument");
  std::vector<int> e32(g->N);
  for (int i = 0; i < g->N; i++) {
    CHECK(e[i] >= INT32_MIN && e[i] <= INT32_MAX, PP_EINVAL, "exponent beyond 32 bits");
    e32[i] = (int)e[i];
  }
  pp_ctx *c = g->ctx;
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));  // earlier work on g is done before it is overwritten
  HIPCHECK(hipMemcpy(g->a, a, sizeof(double) * g->R * g->N * g->S, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(g->e, e32.data(), sizeof(int) * g->N, hipMemcpyHostToDevice));
  return PP_OK;
}

int pp_glv_get(const pp_glv *g, double *a, int64_t *e) {
  CHECK(g && a && e, PP_EINVAL, "null argument");
  pp_ctx *c = g->ctx;
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  std::vector<int> e32(std::max(g->N, 1));
  HIPCHECK(hipMemcpy(a, g->a, sizeof(double) * g->R * g->N * g->S, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(e32.data(), g->e, sizeof(int) * g->N, hipMemcpyDeviceToHost));
  for (int i = 0; i < g->N; i++) e[i] = e32[i];
  return PP_OK;
}

int pp_glv_mask_into(const uint8_t *site_mask, int n_mask, const pp_glv *src, pp_glv *dst) {
  CHECK(site_mask && src && dst && src->ctx == dst->ctx, PP_EINVAL, "null argument or GLVs of different contexts");
  CHECK(n_mask == src->N, PP_EINVAL, "the mask's length differs from src's number of sites");
  CHECK(src->R == dst->R && src->S == dst->S, PP_EINVAL, "GLV rates or states differ");
  std::vector<int> idx;
  for (int i = 0; i < n_mask; i++)
    if (site_mask[i]) idx.push_back(i);
  CHECK((int)idx.size() == dst->N, PP_EINVAL, "the mask's kept sites differ from dst's number of sites");
  if (dst->N == 0) return PP_OK;
  pp_ctx *c = src->ctx;
  HIPCHECK(hipSetDevice(c->device));
  Scratch s;
  HIPCHECK(hipStreamSynchronize(c->stream));  // the index list below is freed on return
  int *di = s.up(idx.data(), idx.size());
  CHECK(di, PP_ENOMEM, "device allocation failed");
  const long n = (long)dst->R * dst->N * dst->S;
  hipLaunchKernelGGL(gv::gather_sites_kernel, dim3(gv::grid_for(n)), dim3(256), 0, c->stream, dst->a, dst->e, src->a,
                     src->e, di, src->R, src->N, dst->N, src->S);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(c->stream));
  return PP_OK;
}

int pp_glv_prep_constant_rate_from_lv_arr(pp_glv *g, const double *lv_arr) {
  CHECK(g && lv_arr, PP_EINVAL, "null argument");
  CHECK(g->N > 0, PP_EINVAL, "empty lv array");
  pp_ctx *c = g->ctx;
  HIPCHECK(hipSetDevice(c->device));
  Scratch s;
  HIPCHECK(hipStreamSynchronize(c->stream));
  double *dl = s.up(lv_arr, (size_t)g->N * g->S);
  CHECK(dl, PP_ENOMEM, "device allocation failed");
  const long n = (long)g->R * g->N * g->S;
  hipLaunchKernelGGL(gv::constant_rate_kernel, dim3(gv::grid_for(n)), dim3(256), 0, c->stream, g->a, dl, g->R, g->N,
                     g->S);
  hipLaunchKernelGGL(gv::fill_int_kernel, dim3(gv::grid_for(g->N)), dim3(256), 0, c->stream, g->e, (long)g->N, 0);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(c->stream));
  return PP_OK;
}

int pp_glv_perhaps_pull_exponent(int min_allowed_twoexp, pp_glv *g) {
  CHECK(g, PP_EINVAL, "null GLV");
  pp_ctx *c = g->ctx;
  HIPCHECK(hipSetDevice(c->device));
  if (g->N == 0) return PP_OK;
  hipLaunchKernelGGL(gv::pull_kernel, dim3((g->N + 255) / 256), dim3(256), 0, c->stream, g->a, g->e, g->R, g->N, g->S,
                     min_allowed_twoexp);
  HIPCHECK(hipGetLastError());
  return PP_OK;
}

int pp_glv_masked_logdot(const pp_glv *x, const pp_glv *y, const uint16_t *mask, double *out) {
  CHECK(mask, PP_EINVAL, "mask is null");
  CHECK(x && x->N > 0, PP_EINVAL, "empty GLV");
  return gv::logdot_impl(x, y, mask, 0, x->N - 1, out);
}

int pp_glv_bounded_logdot(const pp_glv *x, const pp_glv *y, int start, int last, double *out) {
  return gv::logdot_impl(x, y, nullptr, start, last, out);
}

int pp_glv_logdot(const pp_glv *x, const pp_glv *y, double *out) {
  CHECK(x && x->N > 0, PP_EINVAL, "empty GLV");
  return gv::logdot_impl(x, y, nullptr, 0, x->N - 1, out);
}

int pp_glv_pairwise_prod(pp_glv *dst, const pp_glv *x, const pp_glv *y) {
  CHECK(dst && x && y && dst->ctx == x->ctx && x->ctx == y->ctx, PP_EINVAL, "null GLV or different contexts");
  CHECK(gv::same_dims(x, y) && gv::same_dims(dst, x), PP_EINVAL, "GLV dimensions differ");
  pp_ctx *c = x->ctx;
  HIPCHECK(hipSetDevice(c->device));
  const long n = (long)x->R * x->N * x->S;
  hipLaunchKernelGGL(gv::add_exponents_kernel, dim3((x->N + 255) / 256 + 1), dim3(256), 0, c->stream, dst->e, x->e,
                     y->e, x->N);
  if (n) hipLaunchKernelGGL(pairwise_kernel, dim3(gv::grid_for(n)), dim3(256), 0, c->stream, dst->a, x->a, y->a, n,
                            nullptr, x->S);
  HIPCHECK(hipGetLastError());
  return PP_OK;
}

int pp_glv_listwise_prod(pp_glv *dst, const pp_glv *const *glvs, int n) {
  CHECK(dst && glvs && n > 0, PP_EINVAL, "empty list");  // listwise_prod [] is assert(false)
  if (n == 1) return pp_glv_memcpy(dst, glvs[0]);
  int st = pp_glv_pairwise_prod(dst, glvs[0], glvs[1]);
  for (int i = 2; i < n && !st; i++) st = pp_glv_pairwise_prod(dst, dst, glvs[i]);
  return st;
}

int pp_glv_summarize_post(const pp_glv *g, double *u) {
  CHECK(g && u, PP_EINVAL, "null argument");
  pp_ctx *c = g->ctx;
  HIPCHECK(hipSetDevice(c->device));
  if (g->N == 0) return PP_OK;
  Scratch s;
  double *du = s.up<double>(nullptr, (size_t)g->N * g->S);
  CHECK(du, PP_ENOMEM, "device allocation failed");
  const long n = (long)g->N * g->S;
  hipLaunchKernelGGL(gv::state_sums_kernel, dim3(gv::grid_for(n)), dim3(256), 0, c->stream, g->a, g->R, g->N, g->S, du);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(u, du, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return PP_OK;
}

int pp_gmix_log_like3(pp_ctx *ctx, const pp_glv *x, const pp_glv *y, const pp_glv *z, double *out) {
  CHECK(ctx && x && y && z && out, PP_EINVAL, "null argument");
  CHECK(x->ctx == ctx && y->ctx == ctx && z->ctx == ctx, PP_EINVAL, "GLVs of another context");
  CHECK(gv::same_dims(x, y) && gv::same_dims(y, z), PP_EINVAL, "GLV dimensions differ");
  int st = gv::check_gmix(ctx, x);
  if (st) return st;
  CHECK(x->N > 0, PP_EINVAL, "empty GLVs");
  HIPCHECK(hipSetDevice(ctx->device));
  Scratch s;
  double *ds = s.up(ctx->model.statd.data(), (size_t)x->S);
  double *du = s.up<double>(nullptr, x->N), *dt = s.up<double>(nullptr, 2);
  CHECK(ds && du && dt, PP_ENOMEM, "device allocation failed");
  hipLaunchKernelGGL(util_kernel, dim3((x->N + 255) / 256), dim3(256), 0, ctx->stream, du, ds, x->a, y->a, z->a,
                     nullptr, 0, x->N, x->R, x->N, x->S);
  hipLaunchKernelGGL(logsum_kernel, dim3(1), dim3(64), 0, ctx->stream, du, nullptr, 0, x->N, x->R, 1, dt);
  hipLaunchKernelGGL(gv::finish_logdot_kernel, dim3(1), dim3(1), 0, ctx->stream, dt, x->e, y->e, z->e, nullptr, 0,
                     x->N - 1, dt + 1);
  return gv::fetch_double(ctx, dt + 1, out);
}

int pp_gmix_site_log_like_arr3(pp_ctx *ctx, const pp_glv *x, const pp_glv *y, const pp_glv *z, double *out) {
  CHECK(ctx && x && y && z && out, PP_EINVAL, "null argument");
  CHECK(x->ctx == ctx && y->ctx == ctx && z->ctx == ctx, PP_EINVAL, "GLVs of another context");
  CHECK(gv::same_dims(x, y) && gv::same_dims(y, z), PP_EINVAL, "GLV dimensions differ");
  int st = gv::check_gmix(ctx, x);
  if (st) return st;
  if (x->N == 0) return PP_OK;
  HIPCHECK(hipSetDevice(ctx->device));
  Scratch s;
  const int none = INT32_MAX;
  double *ds = s.up(ctx->model.statd.data(), (size_t)x->S), *dout = s.up<double>(nullptr, x->N);
  int *dbad = s.up(&none, 1);
  CHECK(ds && dout && dbad, PP_ENOMEM, "device allocation failed");
  hipLaunchKernelGGL(gv::site_log_like_kernel, dim3((x->N + 255) / 256), dim3(256), 0, ctx->stream, ds, x->a, y->a,
                     z->a, x->e, y->e, z->e, x->R, x->N, x->S, dout, dbad);
  HIPCHECK(hipGetLastError());
  int bad = none;
  HIPCHECK(hipMemcpyAsync(out, dout, sizeof(double) * x->N, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHECK(hipMemcpyAsync(&bad, dbad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  CHECK(bad == none, PP_EQUERY, "Site " + std::to_string(bad) + " has zero likelihood.");
  return PP_OK;
}

int pp_gmix_slow_log_like3(pp_ctx *ctx, const pp_glv *x, const pp_glv *y, const pp_glv *z, double *out) {
  CHECK(out && x, PP_EINVAL, "null argument");
  std::vector<double> v(std::max(x->N, 1));
  const int st = pp_gmix_site_log_like_arr3(ctx, x, y, z, v.data());
  if (st) return st;
  double tot = 0.;  // Array.fold_left ( +. ) 0., gmix_model.ml:347-348
  for (int i = 0; i < x->N; i++) tot = tot + v[i];
  *out = tot;
  return PP_OK;
}

int pp_gmix_evolve_into(pp_ctx *ctx, pp_glv *dst, const pp_glv *src, double bl) {
  CHECK(ctx && dst && src, PP_EINVAL, "null argument");
  CHECK(dst->ctx == ctx && src->ctx == ctx, PP_EINVAL, "GLVs of another context");
  CHECK(gv::same_dims(dst, src), PP_EINVAL, "GLV dimensions differ");
  int st = gv::check_gmix(ctx, src);
  if (st) return st;
  HIPCHECK(hipSetDevice(ctx->device));
  const int R = src->R, S = src->S;
  std::vector<double> P((size_t)R * S * S);
  multi_exp(ctx->model, bl, P.data());  // Diagd.multi_exp, the tensor of prep_tensor_for_bl
  Scratch s;
  HIPCHECK(hipStreamSynchronize(ctx->stream));  // the tensor below is freed on return
  double *dp = s.up(P.data(), P.size());
  CHECK(dp, PP_ENOMEM, "device allocation failed");
  HIPCHECK(hipMemcpyAsync(dst->e, src->e, sizeof(int) * src->N, hipMemcpyDeviceToDevice, ctx->stream));
  const long n = (long)R * src->N * S;
  if (n) hipLaunchKernelGGL(gv::evolve_kernel, dim3(gv::grid_for(n)), dim3(256), 0, ctx->stream, dst->a, src->a, dp, R,
                            src->N, S);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return PP_OK;
}

int pp_gmix_statd_pairwise_prod(pp_ctx *ctx, pp_glv *dst, const pp_glv *a, const pp_glv *b) {
  CHECK(ctx && dst && a && b, PP_EINVAL, "null argument");
  CHECK(dst->ctx == ctx && a->ctx == ctx && b->ctx == ctx, PP_EINVAL, "GLVs of another context");
  CHECK(gv::same_dims(a, b) && gv::same_dims(dst, a), PP_EINVAL, "GLV dimensions differ");
  int st = gv::check_gmix(ctx, a);
  if (st) return st;
  HIPCHECK(hipSetDevice(ctx->device));
  Scratch s;
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  double *ds = s.up(ctx->model.statd.data(), (size_t)a->S);
  CHECK(ds, PP_ENOMEM, "device allocation failed");
  const long n = (long)a->R * a->N * a->S;
  hipLaunchKernelGGL(gv::add_exponents_kernel, dim3((a->N + 255) / 256 + 1), dim3(256), 0, ctx->stream, dst->e, a->e,
                     b->e, a->N);
  if (n) hipLaunchKernelGGL(pairwise_kernel, dim3(gv::grid_for(n)), dim3(256), 0, ctx->stream, dst->a, a->a, b->a, n,
                            ds, a->S);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return PP_OK;
}

}  // extern "C"
