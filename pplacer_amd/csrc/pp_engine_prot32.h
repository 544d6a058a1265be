// FP32 form of the 20-state stage 2: evolves in the state basis on the matrix
// cores.
//
// The FP64 engine (pp_engine_prot.h) evolves in the eigenbasis of Q, which in
// FP32 would rebuild small GLV entries from eigen coordinates up to cond(X)
// larger and lose them.  This form evolves with the transition matrices
// themselves, as Gmix_model.evolve_into does (gmix_model.ml:308-315,
// linear_c.c:29-90): P_r(t)[i][j] = sum_k x[i][k] e_r[k] xit[j][k]
// (Diagd.multi_exp, diagd.ml:79-100) is formed in FP64 from the master's
// exponentials and rounded to FP32 once per changed length; every product
// after that (P, rows, pendant rows) is of non-negative FP32 numbers, so no
// cancellation can amplify rounding.
//
//   yd = P_d,r . dist_row,   yp = P_p,r . prox_row      (per site, per rate)
//   w_i = (pi_i yp_i) yd_i,  util = sum_r sum_i w_i z_r[c][i]
//   z_r[c][i] = sum_k x[i][k] e_pend,r[k] L[c][k]        (pendant row of code c)
//
// The two evolves are [64 pairs x 20] x [20 x 20] products per wave, issued as
// v_mfma_f32_4x4x1f32 (16 independent 4x4 blocks, K = 1, FP32 accumulation):
// block b = lanes 4b..4b+3 holds four pairs of one rate; its A operand is
// rows 4g..4g+3 of that rate's P (lane 4b+m holds row 4g+m), its B operand
// is each lane's own row, so the result D[b][v][n] = y_(lane 4b+n)[4g+v]
// lands in the lane that owns the pair.  Five blocks of rows (g) times twenty
// steps of j evolve one matrix for 64 pairs with no padding.
//
// Lane l of a worker wave: rate r = (l >> 2) mod R, site 4 ((l >> 2) / R) +
// (l & 3) of the wave's current tile of 64 / R masked sites; a site's R rates
// sit in lanes l, l ^ 4, l ^ 8, l ^ 12 of one DPP row.  The w of each pair
// (the distal probe's products) are kept for the pendant-only probes in a
// per-workgroup region of a global scratch buffer, [tile][g][lane] as float4;
// a tile is whichever wave takes it from the probe's shared counter.
//
// The transition tables are built by the worker waves before an extra barrier
// (C) that the master waits at too: 4x4 blocks of entries per lane, FP64
// fused multiply-adds over k, from transposed copies of x, xit and L in LDS.
//
// LDS: the shared header (ProtShared), the masked-site list [N], then
// xT[k][i], xitT[k][j], ltT[k][c] (FP64) and the FP32 tables Pp[R][i][j],
// Pd[R][i][j], Z[R][c][i].
// The gamma model only (R rates, all used at every site); Price-CAT keeps the
// FP64 engine, whose per-site categories would break the blocks' shared rate.
#pragma once

#include "pp_engine_prot.h"

namespace pp {

typedef float pp_f4 __attribute__((ext_vector_type(4)));

template <int R>
struct Prot32Layout {
  static constexpr int S = kProtS;
  static __host__ __device__ size_t off_xT(int N) { return ProtLayout::align(ProtLayout::bytes(N)); }
  static __host__ __device__ size_t off_xitT(int N) { return off_xT(N) + sizeof(double) * S * S; }
  static __host__ __device__ size_t off_ltT(int N) { return off_xitT(N) + sizeof(double) * S * S; }
  static __host__ __device__ size_t off_P(int N) { return off_ltT(N) + sizeof(double) * S * kMaxCodes; }
  static __host__ __device__ size_t off_Z(int N) { return off_P(N) + sizeof(float) * 2 * R * S * S; }
  static __host__ __device__ size_t bytes(int N) { return off_Z(N) + sizeof(float) * R * kMaxCodes * S; }
  // masked sites of one 64-pair tile
  static constexpr int kTileSites = 64 / R;
  // floats of per-workgroup scratch for n_tiles tiles (w of every pair)
  static __host__ __device__ size_t scratch_floats(int n_tiles) { return (size_t)n_tiles * 64 * S; }
  static __device__ double *xT(char *s, int N) { return reinterpret_cast<double *>(s + off_xT(N)); }
  static __device__ double *xitT(char *s, int N) { return reinterpret_cast<double *>(s + off_xitT(N)); }
  static __device__ double *ltT(char *s, int N) { return reinterpret_cast<double *>(s + off_ltT(N)); }
  static __device__ float *Pp(char *s, int N) { return reinterpret_cast<float *>(s + off_P(N)); }
  static __device__ float *Pd(char *s, int N) { return reinterpret_cast<float *>(s + off_P(N)) + R * S * S; }
  static __device__ float *Z(char *s, int N) { return reinterpret_cast<float *>(s + off_Z(N)); }
};

// Model header plus the transposed FP64 copies the table build reads.
template <int R>
__device__ __forceinline__ void prot32_copy_model(char *smem, const DevModel *dm, int N) {
  constexpr int S = kProtS;
  ProtShared *sh = reinterpret_cast<ProtShared *>(smem);
  prot_copy_model(sh, dm);
  if (threadIdx.x == 0) sh->n_codes = dm->n_codes;
  double *xT = Prot32Layout<R>::xT(smem, N), *xitT = Prot32Layout<R>::xitT(smem, N);
  double *ltT = Prot32Layout<R>::ltT(smem, N);
  for (int t = threadIdx.x; t < S * S; t += blockDim.x) {
    const int k = t / S, i = t % S;
    xT[t] = dm->x[i * S + k];
    xitT[t] = dm->xit[i * S + k];
  }
  __syncthreads();  // sh->lt is written by prot_copy_model
  for (int t = threadIdx.x; t < S * kMaxCodes; t += blockDim.x) {
    const int k = t / kMaxCodes, c = t % kMaxCodes;
    ltT[t] = c < dm->n_codes ? sh->lt[c * S + k] : 0.0;
  }
}

// One thread's share (threads tid, tid + nthr, ...) of the FP32 transition
// tables a probe needs (sh->tflags): 4x4 blocks of P_p,r and P_d,r [i][j]
// (2 R 25 blocks), of Z[r][c][i] (R ceil(n_codes / 4) 5 blocks), each entry a
// 20-term FP64 sum over k of x[i][k] e_r[k] v[k] rounded once to FP32.
template <int R>
__device__ void prot32_tables(char *smem, int N, int tid, int nthr) {
  constexpr int S = kProtS;
  using L = Prot32Layout<R>;
  ProtShared *sh = reinterpret_cast<ProtShared *>(smem);
  const int fl = sh->tflags, n_codes = sh->n_codes;
  const double *xT = L::xT(smem, N), *xitT = L::xitT(smem, N), *ltT = L::ltT(smem, N);
  const int ncb = (n_codes + 3) >> 2;
  const int nd = (fl & 1) ? 2 * R * 25 : 0;
  const int np = (fl & 2) ? R * ncb * 5 : 0;
  for (int b = tid; b < nd + np; b += nthr) {
    const double *e, *V;
    int ib, vb, vs, r, which = 0;
    if (b < nd) {
      which = b / (R * 25);
      const int rem = b % (R * 25);
      r = rem / 25;
      ib = (rem % 25) / 5;
      vb = rem % 5;
      V = xitT;
      vs = S;
      e = (which ? sh->ed : sh->ep) + r * S;
    } else {
      const int rem = b - nd;
      r = rem / (ncb * 5);
      vb = (rem % (ncb * 5)) / 5;
      ib = rem % 5;
      V = ltT;
      vs = kMaxCodes;
      e = sh->en + r * S;
    }
    double acc[4][4];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int s = 0; s < 4; s++) acc[q][s] = 0.0;
#pragma unroll 4
    for (int k = 0; k < S; k++) {
      const double ek = e[k];
      double u[4], v[4];
#pragma unroll
      for (int q = 0; q < 4; q++) u[q] = xT[k * S + 4 * ib + q] * ek;
#pragma unroll
      for (int s = 0; s < 4; s++) v[s] = V[k * vs + 4 * vb + s];
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int s = 0; s < 4; s++) acc[q][s] = fma(u[q], v[s], acc[q][s]);
    }
    if (b < nd) {
      float *P = (which ? L::Pd(smem, N) : L::Pp(smem, N)) + r * S * S;
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int s = 0; s < 4; s++) P[(4 * ib + q) * S + 4 * vb + s] = (float)acc[q][s];
    } else {
      float *Zr = L::Z(smem, N) + r * kMaxCodes * S;
#pragma unroll
      for (int s = 0; s < 4; s++)
        if (4 * vb + s < n_codes)
#pragma unroll
          for (int q = 0; q < 4; q++) Zr[(4 * vb + s) * S + 4 * ib + q] = (float)acc[q][s];
    }
  }
}

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// y = P . row for the wave's 64 pairs: arow = row 4g + (lane & 3) of the
// lane's rate's P (g = 0..4 at arow + 4 g S), row = the lane's own row.
// Five accumulators, twenty K = 1 steps each.
__device__ __forceinline__ void prot32_evolve(const float *arow, const pp_f4 *row, pp_f4 (&y)[5]) {
  constexpr int S = kProtS;
  pp_f4 b[5];
#pragma unroll
  for (int q = 0; q < 5; q++) {
    b[q] = row[q];
    y[q] = (pp_f4){0.f, 0.f, 0.f, 0.f};
  }
#pragma unroll
  for (int jq = 0; jq < 5; jq++) {
    pp_f4 a[5];
#pragma unroll
    for (int g = 0; g < 5; g++) a[g] = *reinterpret_cast<const pp_f4 *>(arow + 4 * g * S + 4 * jq);
#pragma unroll
    for (int jj = 0; jj < 4; jj++)
#pragma unroll
      for (int g = 0; g < 5; g++) y[g] = __builtin_amdgcn_mfma_f32_4x4x1f32(a[g][jj], b[jq][jj], y[g], 0, 0, 0);
  }
}

// Worker waves of the FP32 form.  wsc: this workgroup's scratch region.
// The waves take the tiles of a probe from a shared counter (sh->tile_next,
// cleared by the master with each probe command), so the matrix cores of the
// CU's SIMDs share the work as they drain it.
template <int R>
__device__ __forceinline__ void prot32_worker(char *smem, float *wsc, int N, const uint8_t *codes,
                                              const float *darr, const float *parr, const int *darr_e,
                                              const int *parr_e) {
  constexpr int S = kProtS, RS = R * S;
  using L = Prot32Layout<R>;
  constexpr int ST = L::kTileSites;
  static_assert(R == 1 || R == 4, "the rates of a site are a rotation cycle of a DPP row");
  ProtShared *sh = reinterpret_cast<ProtShared *>(smem);
  const int *sidx = ProtLayout::sidx(smem);
  const int lane = threadIdx.x & 63;
  const int wv = wave_id() - 1;
  const int nw = (int)(blockDim.x >> 6) - 1;
  const int r = (lane >> 2) & (R - 1);
  const int ls = ((lane >> 2) / R) * 4 + (lane & 3);
  const int mrow = lane & 3;
  const float *Z = L::Z(smem, N);
  // A operands: row 4g + mrow of the lane's rate's P, four j at a time
  const float *app = L::Pp(smem, N) + (r * S + mrow) * S, *apd = L::Pd(smem, N) + (r * S + mrow) * S;
  pp_f4 *ws = reinterpret_cast<pp_f4 *>(wsc);
  // The command stream is QUERY (GATHER PROBE*)* ... EXIT (pp_engine_mw.h);
  // probes carry one more barrier (C) after the table build.
  __syncthreads();  // A
  int op = sh->op;
  while (op == CMD_QUERY) {
    const int M = sh->M;
    const int n_tiles = (M + ST - 1) / ST;
    const uint8_t *qc = codes + (size_t)sh->q * N;
    __syncthreads();  // B
    __syncthreads();  // A
    op = sh->op;
    while (op == CMD_GATHER) {
      const int loc = sh->loc;
      long long es = 0;
      for (int T = wv; T < n_tiles; T += nw) {
        const int m = T * ST + ls;
        if (m < M && r == 0) {
          const size_t row = (size_t)loc * N + sidx[m];
          es += (long long)darr_e[row] + (long long)parr_e[row];
        }
      }
      es = wave_sum_ll(es);
      if (lane == 0) sh->ipart[wv] = es;
      __syncthreads();  // B
      __syncthreads();  // A
      op = sh->op;
      while (op == CMD_PROBE || op == CMD_PROBE_DIST) {
        prot32_tables<R>(smem, N, threadIdx.x - 64, blockDim.x - 64);
        __syncthreads();  // C: the tables are in place
        const bool nd = op == CMD_PROBE_DIST;
        const size_t row0 = (size_t)sh->loc * N;
        double lsum = 0.0;
        for (;;) {
          int T = 0;
          if (lane == 0) T = __hip_atomic_fetch_add(&sh->tile_next, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          T = __builtin_amdgcn_readfirstlane(T);
          if (T >= n_tiles) break;
          const int m = T * ST + ls;
          const bool act = m < M;
          // inactive lanes evolve the tile's first site (any valid row)
          const int site = sidx[act ? m : T * ST];
          pp_f4 w[5];
          if (nd) {
            pp_f4 yd[5], yp[5];
            prot32_evolve(apd, reinterpret_cast<const pp_f4 *>(darr + (row0 + site) * RS + r * S), yd);
            prot32_evolve(app, reinterpret_cast<const pp_f4 *>(parr + (row0 + site) * RS + r * S), yp);
#pragma unroll
            for (int g = 0; g < 5; g++) {
#pragma unroll
              for (int v = 0; v < 4; v++) w[g][v] = ((float)sh->statd[4 * g + v] * yp[g][v]) * yd[g][v];
              ws[(size_t)(T * 5 + g) * 64 + lane] = w[g];
            }
          } else {
#pragma unroll
            for (int g = 0; g < 5; g++) w[g] = ws[(size_t)(T * 5 + g) * 64 + lane];
          }
          const pp_f4 *zr = reinterpret_cast<const pp_f4 *>(Z + (r * kMaxCodes + qc[site]) * S);
          float util = 0.f;
#pragma unroll
          for (int g = 0; g < 5; g++) {
            const pp_f4 z = zr[g];
#pragma unroll
            for (int v = 0; v < 4; v++) util = (g == 0 && v == 0) ? w[0][0] * z[0] : fmaf(w[g][v], z[v], util);
          }
          if (R == 4) {  // lanes l, l + 4, l + 8, l + 12 of the row
            util += dpp_f32<0x124>(util);  // row_ror:4
            util += dpp_f32<0x128>(util);  // row_ror:8
          }
          if (act && r == 0) lsum += (double)__logf(util);
        }
        lsum = wave_sum(lsum);
        if (lane == 0) sh->part[wv] = lsum;
        __syncthreads();  // B
        __syncthreads();  // A
        op = sh->op;
      }
    }
  }
  // op == CMD_EXIT
}

}  // namespace pp
