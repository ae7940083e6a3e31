// Stages after the pair correlation, all on the GPU:
//   S1  peak-significance histogram/CDF and the logistic weight fit (FP64)
//   S2  weighted-L1 shift solve per axis (FP64): IRLS warm start, then exact
//       network-simplex descent over spanning trees of the frame-position graph
//   S3  per-frame shift parameters, translations, error statistics, FRC crops
//   K3  Lanczos/integer shift + imageSum + even/odd half sums
//   K4  tiled Hann FRC (row FFTs, column FFTs with FP64 accumulation, ring means)
//   KD  dose filter (aligned-row FFTs, per-column filtered accumulation, inverse)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <map>
#include <vector>

#include "zr_common.h"
#include "zr_launch.h"
#include "zr_stages.h"

// ---------------------------------------------------------------- constants
// Deterministic logistic-fit schedule (same constants as the oracle).
constexpr int LOGI_G0 = 32, LOGI_G1 = 32, LOGI_G2 = 32;
constexpr int LOGI_ROUNDS = 30, LOGI_MAX_ITERS = 400;
constexpr double LOGI_K0 = 0.1, LOGI_K1 = 20.0, LOGI_N0 = 0.05, LOGI_N1 = 5.0;
constexpr int MAXN = 64;          // frames handled by the one-wave solver
constexpr int DOSE_CHUNK = 8;     // frames per row/column pass of the direct dose path
constexpr int IRLS_ITERS = 2;
constexpr int MAX_PIVOTS = 4000;

struct FrameShift {
  int iy, ix, conv, pad;
  float k[9];        // 3x3 Lanczos kernel, k[a][b] = u[a] v[b]
  float u[3], v[3];  // its separable factors (rows, columns)
};

struct SolveInfo {           // device-resident scalars of the solve
  int nps, nbins, status_x, status_y;
  int pivots_x, pivots_y, tiles_y, tiles_x;
  int crop[4];
  double logi[3];
  double dpf;
};

// ------------------------------------------------------------------ S1
ZR_DEV float f32_sub(float a, float b) { return __fsub_rn(a, b); }
ZR_DEV float f32_mul(float a, float b) { return __fmul_rn(a, b); }
ZR_DEV float f32_add(float a, float b) { return __fadd_rn(a, b); }
ZR_DEV float f32_div(float a, float b) { return __fdiv_rn(a, b); }

// Histogram of positive peak significances (numpy.histogram with an integer
// bin count on float32 data: float32 edges from linspace, index by scaling,
// then the +-1 edge corrections), CDF in float32.
__global__ __launch_bounds__(256) void k_ps_hist(const float* __restrict__ peaksig, int nn, float* __restrict__ hs,
                                                 float* __restrict__ cdf, float* __restrict__ pdf, SolveInfo* info,
                                                 int maxbins) {
  __shared__ float smin[4], smax[4];
  __shared__ int scnt[4];
  __shared__ unsigned counts[2048];
  const int tid = threadIdx.x;
  float mn = INFINITY, mx = -INFINITY;
  int c = 0;
  for (int i = tid; i < nn; i += 256) {
    float v = peaksig[i];
    if (v > 0.f) { mn = fminf(mn, v); mx = fmaxf(mx, v); ++c; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    mn = fminf(mn, __shfl_xor(mn, o, 64));
    mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    c += __shfl_xor(c, o, 64);
  }
  if ((tid & 63) == 0) { smin[tid >> 6] = mn; smax[tid >> 6] = mx; scnt[tid >> 6] = c; }
  __syncthreads();
  mn = fminf(fminf(smin[0], smin[1]), fminf(smin[2], smin[3]));
  mx = fmaxf(fmaxf(smax[0], smax[1]), fmaxf(smax[2], smax[3]));
  c = scnt[0] + scnt[1] + scnt[2] + scnt[3];
  int nb = (int)((double)c / 7.0);
  if (nb < 1 || nb > maxbins || nb > 2048) {
    if (tid == 0) { info->nps = c; info->nbins = 0; }
    return;
  }
  float first = mn, last = mx;
  if (first == last) { first = f32_sub(first, 0.5f); last = f32_add(last, 0.5f); }
  const float delta = f32_sub(last, first);
  const float step = f32_div(delta, (float)nb);
  for (int i = tid; i < nb; i += 256) counts[i] = 0;
  __syncthreads();
  auto edge = [&](int i) -> float { return i == nb ? last : f32_add(f32_mul((float)i, step), first); };
  for (int i = tid; i < nn; i += 256) {
    float v = peaksig[i];
    if (!(v > 0.f)) continue;
    float fi = f32_mul(f32_div(f32_sub(v, first), delta), (float)nb);
    int k = (int)fi;
    if (k == nb) k -= 1;
    if (v < edge(k)) k -= 1;
    if (v >= edge(k + 1) && k != nb - 1) k += 1;
    atomicAdd(&counts[k], 1u);
  }
  __syncthreads();
  if (tid == 0) {
    float run = 0.f;
    for (int i = 0; i < nb; ++i) { pdf[i] = (float)counts[i]; run = f32_add(run, pdf[i]); cdf[i] = run; }
    for (int i = 0; i < nb; ++i) { cdf[i] = f32_div(cdf[i], run); hs[i] = edge(i); }
    info->nps = c;
    info->nbins = nb;
    info->logi[0] = mn;      // thr bounds of the fit: data min/max
    info->logi[1] = mx;
  }
}

// 1 - (1 + e^{-K (thr - h)})^{-nu} (zorro/zorro.py:1344), the power taken as
// exp(-nu log1p(.)): the log term depends on (thr, K) only, so a pattern-search
// step shares it between the candidates that differ in nu
ZR_DEV double logistic_lse(double h, double thr, double K) { return log1p(exp(-K * (-h + thr))); }
ZR_DEV double logistic_from_lse(double lse, double nu) { return 1.0 - exp(-nu * lse); }
ZR_DEV double logistic_d(double h, double thr, double K, double nu) {
  return logistic_from_lse(logistic_lse(h, thr, K), nu);
}

ZR_DEV double logi_obj(const float* hs, const float* cdf, int nb, double thr, double K, double nu) {
  double s = 0.0;
  for (int i = 0; i < nb; ++i) s += fabs((double)cdf[i] - logistic_d((double)hs[i], thr, K, nu));
  return s;
}

__global__ __launch_bounds__(256) void k_logi_grid(const float* __restrict__ hs, const float* __restrict__ cdf,
                                                  const SolveInfo* info, double* __restrict__ obj) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = LOGI_G0 * LOGI_G1 * LOGI_G2;
  if (i >= total || info->nbins < 1) return;
  const int a = i / (LOGI_G1 * LOGI_G2), b = (i / LOGI_G2) % LOGI_G1, c = i % LOGI_G2;
  const double lo = info->logi[0], hi = info->logi[1];
  const double thr = lo + (hi - lo) * (double)a / (LOGI_G0 - 1);
  const double lk0 = log(LOGI_K0), lk1 = log(LOGI_K1), ln0 = log(LOGI_N0), ln1 = log(LOGI_N1);
  const double lk = lk0 + (lk1 - lk0) * (double)b / (LOGI_G1 - 1);
  const double ln = ln0 + (ln1 - ln0) * (double)c / (LOGI_G2 - 1);
  obj[i] = logi_obj(hs, cdf, info->nbins, thr, exp(lk), exp(ln));
}

// argmin of the grid (first index) then the 3x3x3 pattern search.  One block of
// LOGI_NT: the log terms of the nine (thr, K) pairs of a step, then every
// (candidate, bin) term, are evaluated in parallel (K = e^lk and nu = e^ln once
// per candidate), each candidate's objective is then summed over bins in index
// order.  Dynamic LDS: logi_refine_lds_bytes(), for up to LOGI_MAXB bins.
constexpr int LOGI_NT = 1024, LOGI_MAXB = 290;
constexpr size_t logi_refine_lds_bytes() { return sizeof(double) * (27 + 9) * LOGI_MAXB; }
__global__ __launch_bounds__(LOGI_NT) void k_logi_refine(const float* __restrict__ hs, const float* __restrict__ cdf,
                                                        const double* __restrict__ obj, SolveInfo* info) {
  extern __shared__ double logi_dyn[];
  double* terms = logi_dyn;                   // [27][nb]
  double* lse = logi_dyn + 27 * LOGI_MAXB;    // [9][nb]: (thr index o0 + 1) * 3 + (K index o1 + 1)
  __shared__ double cval[27], ccand[27][5];
  __shared__ double rv[LOGI_NT / 64];
  __shared__ int ri[LOGI_NT / 64];
  const int tid = threadIdx.x;
  if (info->nbins < 1) return;
  const int total = LOGI_G0 * LOGI_G1 * LOGI_G2;
  double bv = INFINITY;
  int bi = 0x7fffffff;
  for (int i = tid; i < total; i += LOGI_NT) {
    double v = obj[i];
    if (v < bv || (v == bv && i < bi)) { bv = v; bi = i; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    double ov = __shfl_xor(bv, o, 64);
    int oi = __shfl_xor(bi, o, 64);
    if (ov < bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
  }
  if ((tid & 63) == 0) { rv[tid >> 6] = bv; ri[tid >> 6] = bi; }
  __syncthreads();
  bv = rv[0]; bi = ri[0];
  for (int k = 1; k < LOGI_NT / 64; ++k)
    if (rv[k] < bv || (rv[k] == bv && ri[k] < bi)) { bv = rv[k]; bi = ri[k]; }
  const double lo = info->logi[0], hi = info->logi[1];
  const double lk0 = log(LOGI_K0), lk1 = log(LOGI_K1), ln0 = log(LOGI_N0), ln1 = log(LOGI_N1);
  double cur[3];
  {
    int a = bi / (LOGI_G1 * LOGI_G2), b = (bi / LOGI_G2) % LOGI_G1, c = bi % LOGI_G2;
    cur[0] = lo + (hi - lo) * (double)a / (LOGI_G0 - 1);
    cur[1] = lk0 + (lk1 - lk0) * (double)b / (LOGI_G1 - 1);
    cur[2] = ln0 + (ln1 - ln0) * (double)c / (LOGI_G2 - 1);
  }
  double fcur = bv;
  double step[3] = {(hi - lo) / (LOGI_G0 - 1), (lk1 - lk0) / (LOGI_G1 - 1), (ln1 - ln0) / (LOGI_G2 - 1)};
  const double lob[3] = {lo, lk0, ln0}, hib[3] = {hi, lk1, ln1};
  const int nb = info->nbins;
  int halvings = 0;
  for (int it = 0; it < LOGI_MAX_ITERS; ++it) {
    if (tid < 27) {
      int o0 = tid / 9 - 1, o1 = (tid / 3) % 3 - 1, o2 = tid % 3 - 1;
      ccand[tid][0] = fmin(fmax(cur[0] + o0 * step[0], lob[0]), hib[0]);
      ccand[tid][1] = fmin(fmax(cur[1] + o1 * step[1], lob[1]), hib[1]);
      ccand[tid][2] = fmin(fmax(cur[2] + o2 * step[2], lob[2]), hib[2]);
      ccand[tid][3] = exp(ccand[tid][1]);
      ccand[tid][4] = exp(ccand[tid][2]);
    }
    __syncthreads();
    for (int e = tid; e < 9 * nb; e += LOGI_NT) {
      const int a = e / nb, i = e % nb, c = 3 * a + 1;   // a candidate with this (thr, K)
      lse[e] = logistic_lse((double)hs[i], ccand[c][0], ccand[c][3]);
    }
    __syncthreads();
    for (int e = tid; e < 27 * nb; e += LOGI_NT) {
      int c = e / nb, i = e % nb;
      terms[e] = fabs((double)cdf[i] - logistic_from_lse(lse[(c / 3) * nb + i], ccand[c][4]));
    }
    __syncthreads();
    if (tid < 27) {
      double sum = 0.0;
      for (int i = 0; i < nb; ++i) sum += terms[tid * nb + i];
      cval[tid] = sum;
    }
    __syncthreads();
    double mv = cval[0];
    int mi = 0;
    for (int c = 1; c < 27; ++c)
      if (cval[c] < mv) { mv = cval[c]; mi = c; }
    if (mv < fcur) {
      cur[0] = ccand[mi][0]; cur[1] = ccand[mi][1]; cur[2] = ccand[mi][2];
      fcur = mv;
    } else {
      step[0] *= 0.5; step[1] *= 0.5; step[2] *= 0.5;
      if (++halvings >= LOGI_ROUNDS) break;
    }
    __syncthreads();
  }
  if (tid == 0) {
    info->logi[0] = cur[0];
    info->logi[1] = exp(cur[1]);
    info->logi[2] = exp(cur[2]);
  }
}

// ------------------------------------------------------------------ S2
struct SolveArgs {
  const float* shifts;    // [N][N][2]
  const float* corr;      // [N][N]
  const float* peaksig;   // [N][N]
  const int2* edges;      // [M] graph edge (i, j): the equation p_j - p_i = b
  const int2* meas;       // [M] trimat entry (I, J) holding b and the peak significance
  const int2* dups;       // [M] (first edge with the same (i, j), next such edge or -1)
  int N, M;
  int weight_mode;
  double logi_fixed[3];
  double bound;
  double* weights;        // [M]   (unclipped)
  double* resid;          // [M][2]
  double* pos;            // [2][N]  positions p (p_0 = 0)
  SolveInfo* info;
};

// One 256-thread block per axis (blockIdx.x).  Solves min sum_e w_e |p_J - p_I - b_e|
// over frame positions p (p_0 = 0) exactly:
//   1. IRLS warm start on the weighted position Laplacian (Cholesky, node 0
//      removed);
//   2. Kruskal spanning tree over |residual| ascending (ties: edge index);
//   3. dual network simplex with bound flipping.  Every non-tree edge e sits at
//      a flow bound f_e = s_e w_e, s_e = sign(residual) (for a zero residual the
//      remembered side).  The tree edge above subtree S carries the flow that
//      balances the bounds crossing S; where that exceeds its weight, S moves
//      (the directional slope is negative), passing breakpoints of crossing
//      edges (each flips its side) until the slope turns non-negative; the last
//      one enters the tree.  When every tree flow fits its weight the flows
//      certify optimality, also at degenerate vertices.  A long run of
//      zero-length pivots switches to Bland's rule (smallest edge index,
//      single breakpoint), which cannot cycle.
// Edges live in registers (thread t owns edges t + 256 k); frame-level work runs
// in wave 0 with lane v standing for frame v.  Positions are recomputed at the
// end by summing measured shifts along the final tree, so the vertex is exact.
constexpr int L1_NT = 256;
constexpr int L1_EPT = 8;                 // edges per thread: M <= 2048
static_assert(L1_EPT < 16, "a thread's breakpoint count fits the four-bit prefix sum");
constexpr int L1_LDC = MAXN + 1;          // matrix row stride
constexpr int L1_BLAND = 3;               // zero-length pivots (x N) before Bland's rule

ZR_DEV double lane_bcast_d(double v, int src) {   // src uniform
  const unsigned long long u = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, src);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), src);
  return __longlong_as_double(((unsigned long long)hi << 32) | lo);
}
ZR_DEV unsigned long long shfl_u64(unsigned long long v, int src) {
  const unsigned lo = __shfl((unsigned)v, src, 64);
  const unsigned hi = __shfl((unsigned)(v >> 32), src, 64);
  return ((unsigned long long)hi << 32) | lo;
}
ZR_DEV bool lex_less(double d1, int e1, double d2, int e2) { return d1 < d2 || (d1 == d2 && e1 < e2); }
// Lexicographic (d, e) minimum over the wave (all 64 lanes active) with a
// payload pl that follows the winner, in every lane: within rows of 16 lanes
// by DPP exchanges (lanes 1 and 2 apart, then the half-row and row mirrors),
// then the four row minima in row order.  DPP moves instead of LDS permutes:
// the solver's reductions are single-wave latency chains.
template <int CTRL> ZR_DEV void lexmin_dpp(double& d, int& e, int& pl) {
  const unsigned long long u = __double_as_longlong(d);
  const unsigned lo = __builtin_amdgcn_update_dpp((unsigned)u, (unsigned)u, CTRL, 0xF, 0xF, false);
  const unsigned hi = __builtin_amdgcn_update_dpp((unsigned)(u >> 32), (unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
  const int oe = __builtin_amdgcn_update_dpp(e, e, CTRL, 0xF, 0xF, false);
  const int op = __builtin_amdgcn_update_dpp(pl, pl, CTRL, 0xF, 0xF, false);
  const double od = __longlong_as_double(((unsigned long long)hi << 32) | lo);
  if (lex_less(od, oe, d, e)) { d = od; e = oe; pl = op; }
}
ZR_DEV void wave_lexmin(double& d, int& e, int& pl) {
  lexmin_dpp<0xB1>(d, e, pl);    // quad_perm [1, 0, 3, 2]
  lexmin_dpp<0x4E>(d, e, pl);    // quad_perm [2, 3, 0, 1]
  lexmin_dpp<0x141>(d, e, pl);   // row_half_mirror
  lexmin_dpp<0x140>(d, e, pl);   // row_mirror
  double bd = lane_bcast_d(d, 0);
  int be = __builtin_amdgcn_readlane(e, 0), bp = __builtin_amdgcn_readlane(pl, 0);
#pragma unroll
  for (int r = 1; r < 4; ++r) {
    const double od = lane_bcast_d(d, 16 * r);
    const int oe = __builtin_amdgcn_readlane(e, 16 * r), op = __builtin_amdgcn_readlane(pl, 16 * r);
    if (lex_less(od, oe, bd, be)) { bd = od; be = oe; bp = op; }
  }
  d = bd;
  e = be;
  pl = bp;
}

// dynamic LDS of k_l1_solve: matrix region (Cholesky / slope sums / sort),
// then b[M], w[M], breakpoints[M] (double), endpoints[M], tree flags[M],
// breakpoint edges[M], next parallel edge[M] (int)
inline __host__ __device__ size_t l1_dyn_bytes(int M) {
  return sizeof(double) * ((size_t)MAXN * L1_LDC + 3 * (size_t)M) + 4 * sizeof(int) * (size_t)M;
}

__global__ __launch_bounds__(L1_NT) void k_l1_solve(SolveArgs a) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int axis = blockIdx.x;            // 0: y (index 0), 1: x (index 1)
  const int N = a.N, M = a.M, n1 = N - 1;
  constexpr int LDC = L1_LDC;
  extern __shared__ double dyn[];
  double* Lm = dyn;
  double* b_s = Lm + MAXN * LDC;
  double* w_s = b_s + M;
  double* cd_s = w_s + M;                         // breakpoint list: distances
  int* es_s = reinterpret_cast<int*>(cd_s + M);   // i | j << 8
  int* tfl = es_s + M;
  int* ce_s = tfl + M;                            // breakpoint list: edges
  int* dn_s = ce_s + M;                           // next edge with the same (i, j), or -1
  __shared__ double pvs[MAXN], idiag[MAXN];
  __shared__ int t_par[MAXN], t_ped[MAXN];
  __shared__ unsigned long long t_sub[MAXN];
  __shared__ int s_nc, s_enter;
  __shared__ double s_last;
  __shared__ unsigned long long sub_s[MAXN];
  __shared__ int par_s[MAXN], ped_s[MAXN], adj_o[MAXN + 1], adj_c[MAXN], adj[2 * MAXN];
  __shared__ double red_d[2][4], red_w[2][4];
  __shared__ int red_i[2][4];
  __shared__ int s_key, s_v, s_used;
  __shared__ double s_best;
  __shared__ double avs_s[MAXN];     // net bound flow into each frame (the slope sums read it by broadcast)

  // ---- data: each thread keeps its edges in registers
  int eij[L1_EPT];
  double eb[L1_EPT], ew[L1_EPT], er[L1_EPT];
  unsigned tmask = 0u, smask = 0u;      // bit k: tree edge / side +1
  unsigned dfirst = 0u;                 // bit k: first edge of its (i, j) (it sums the parallel ones)
#pragma unroll
  for (int k = 0; k < L1_EPT; ++k) {
    const int e = tid + k * L1_NT;
    eij[k] = -1; eb[k] = 0.0; ew[k] = 0.0; er[k] = 0.0;
    if (e < M) {
      const int2 ed = a.edges[e], ms = a.meas[e];
      const double bb = (double)a.shifts[((size_t)ms.x * N + ms.y) * 2 + axis];
      const double ps = (double)a.peaksig[ms.x * N + ms.y];
      double we;
      switch (a.weight_mode) {
        case 0: we = logistic_d(ps, a.info->logi[0], a.info->logi[1], a.info->logi[2]); break;
        case 1: we = logistic_d(ps, a.logi_fixed[0], a.logi_fixed[1], a.logi_fixed[2]); break;
        case 3: we = 1.0; break;
        default: we = ps; break;
      }
      eij[k] = ed.x | (ed.y << 8);
      eb[k] = bb;
      ew[k] = we;
      const int2 dp = a.dups[e];        // the parallel-edge chain, read once into LDS
      if (dp.x == e) dfirst |= 1u << k;
      dn_s[e] = dp.y;
    }
  }
  if (a.weight_mode == 4) {   // 'norm': (ps - min)/(max - min)
    double mn = INFINITY, mx = -INFINITY;
#pragma unroll
    for (int k = 0; k < L1_EPT; ++k)
      if (eij[k] >= 0) { mn = fmin(mn, ew[k]); mx = fmax(mx, ew[k]); }
    for (int o = 32; o > 0; o >>= 1) { mn = fmin(mn, __shfl_xor(mn, o, 64)); mx = fmax(mx, __shfl_xor(mx, o, 64)); }
    if (lane == 0) { red_d[0][wid] = mn; red_w[0][wid] = mx; }
    __syncthreads();
    mn = fmin(fmin(red_d[0][0], red_d[0][1]), fmin(red_d[0][2], red_d[0][3]));
    mx = fmax(fmax(red_w[0][0], red_w[0][1]), fmax(red_w[0][2], red_w[0][3]));
    __syncthreads();
#pragma unroll
    for (int k = 0; k < L1_EPT; ++k)
      if (eij[k] >= 0) ew[k] = (ew[k] - mn) / (mx - mn);
  }
#pragma unroll
  for (int k = 0; k < L1_EPT; ++k) {
    const int e = tid + k * L1_NT;
    if (eij[k] >= 0) {
      es_s[e] = eij[k];
      b_s[e] = eb[k];
      w_s[e] = ew[k];
      tfl[e] = 0;
      if (axis == 0) a.weights[e] = fmax(ew[k], 1e-6);   // errorDict['weights']: clipped to >= 1e-6
    }
  }
  if (tid < MAXN) pvs[tid] = 0.0;
  __syncthreads();
  // ---- IRLS warm start on the position Laplacian (node 0 fixed).  Edge (i < j)
  // stores -d in the lower triangle (j, i) and -d * b in the upper (i, j).
  double eps = 1.0;
  for (int it = 0; it < IRLS_ITERS; ++it) {
    for (int q = tid; q < N * LDC; q += L1_NT) Lm[q] = 0.0;
    __syncthreads();
    // per-edge weights staged in cd_s; the first edge of each (i, j) sums its
    // parallel edges in edge order (the equations can repeat a frame span)
#pragma unroll
    for (int k = 0; k < L1_EPT; ++k) {
      if (eij[k] < 0) continue;
      const int i = eij[k] & 255, j = eij[k] >> 8;
      const double re = pvs[j] - pvs[i] - eb[k];
      cd_s[tid + k * L1_NT] = fmax(ew[k], 1e-12) / fmax(fabs(re), eps);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < L1_EPT; ++k) {
      const int e = tid + k * L1_NT;
      if (eij[k] < 0 || !((dfirst >> k) & 1u)) continue;
      double d = cd_s[e], db = cd_s[e] * b_s[e];
      for (int q = dn_s[e]; q >= 0; q = dn_s[q]) { d += cd_s[q]; db += cd_s[q] * b_s[q]; }
      const int i = eij[k] & 255, j = eij[k] >> 8;
      Lm[j * LDC + i] = -d;
      Lm[i * LDC + j] = -db;
    }
    __syncthreads();
    double gv = 0.0;     // wave 0, lane k: right-hand side of frame k
    if (tid < N) {
      double sd = 0.0;
      for (int j = 0; j < N; ++j) {
        if (j < tid) { sd -= Lm[tid * LDC + j]; gv -= Lm[j * LDC + tid]; }
        else if (j > tid) { sd -= Lm[j * LDC + tid]; gv += Lm[tid * LDC + j]; }
      }
      Lm[tid * LDC + tid] = sd;
    }
    __syncthreads();
    // Cholesky (lower) of rows/cols 1..N-1; the factor's diagonal goes to diag[]
    for (int k = 1; k < N; ++k) {
      const double ik = 1.0 / sqrt(fmax(Lm[k * LDC + k], 1e-300));
      if (tid == 0) idiag[k] = ik;
      for (int i = k + 1 + tid; i < N; i += L1_NT) Lm[i * LDC + k] *= ik;
      __syncthreads();
      const int nt = N - 1 - k;
      const int T = nt * (nt + 1) / 2;
      for (int q = tid; q < T; q += L1_NT) {
        int rr = (int)((sqrtf(8.f * q + 1.f) - 1.f) * 0.5f);
        while ((rr + 1) * (rr + 2) / 2 <= q) ++rr;
        while (rr * (rr + 1) / 2 > q) --rr;
        const int i = k + 1 + rr, j = k + 1 + (q - rr * (rr + 1) / 2);
        Lm[i * LDC + j] -= Lm[i * LDC + k] * Lm[j * LDC + k];
      }
      __syncthreads();
    }
    // L y = g, then L^T p = y (wave 0; lane i holds y_i, then p_i)
    if (wid == 0) {
      for (int j = 1; j < N; ++j) {
        const double yj = lane_bcast_d(gv, j) * idiag[j];
        if (lane == j) gv = yj;
        else if (lane > j && lane < N) gv -= Lm[lane * LDC + j] * yj;
      }
      for (int i = N - 1; i >= 1; --i) {
        const double pi = lane_bcast_d(gv, i) * idiag[i];
        if (lane == i) gv = pi;
        else if (lane >= 1 && lane < i) gv -= Lm[i * LDC + lane] * pi;
      }
      if (lane < N) pvs[lane] = (lane == 0) ? 0.0 : gv;
    }
    __syncthreads();
    eps = fmax(eps * 0.5, 1e-9);
  }
  // initial sides from the warm-start residuals
#pragma unroll
  for (int k = 0; k < L1_EPT; ++k) {
    if (eij[k] < 0) continue;
    const int i = eij[k] & 255, j = eij[k] >> 8;
    if (pvs[j] - pvs[i] - eb[k] >= 0.0) smask |= 1u << k;
  }
  // ---- initial tree: Kruskal over |residual| ascending (ties: edge index)
  {
    int P2 = 1;
    while (P2 < M) P2 <<= 1;
    double* key = Lm;
    int* val = reinterpret_cast<int*>(Lm + P2);
    for (int i = tid; i < P2; i += L1_NT) {
      if (i < M) {
        const int ij = es_s[i];
        key[i] = fabs(pvs[ij >> 8] - pvs[ij & 255] - b_s[i]);
        val[i] = i;
      } else {
        key[i] = INFINITY;
        val[i] = 0x7fffffff;
      }
    }
    __syncthreads();
    for (int kk = 2; kk <= P2; kk <<= 1)
      for (int jj = kk >> 1; jj > 0; jj >>= 1) {
        for (int q = tid; q < P2 / 2; q += L1_NT) {
          const int i = ((q & ~(jj - 1)) << 1) | (q & (jj - 1));
          const int l = i | jj;
          const bool up = (i & kk) == 0;
          const double ki = key[i], kl = key[l];
          const int vi = val[i], vl = val[l];
          const bool gt = (ki > kl) || (ki == kl && vi > vl);
          if (gt == up) { key[i] = kl; key[l] = ki; val[i] = vl; val[l] = vi; }
        }
        __syncthreads();
      }
    if (wid == 0) {   // union-find as per-lane component labels
      int label = lane, used = 0;
      for (int t0 = 0; t0 < M && used < n1; t0 += 64) {
        const int myv = (t0 + lane < M) ? val[t0 + lane] : 0;
        const int myij = es_s[myv];
        const int cnt = min(64, M - t0);
        for (int q = 0; q < cnt && used < n1; ++q) {
          const int e = __builtin_amdgcn_readlane(myv, q);
          const int ij = __builtin_amdgcn_readlane(myij, q);
          const int lx = __builtin_amdgcn_readlane(label, ij & 255);
          const int ly = __builtin_amdgcn_readlane(label, ij >> 8);
          if (lx != ly) {
            label = (label == ly) ? lx : label;
            if (lane == 0) tfl[e] = 1;
            ++used;
          }
        }
      }
      if (lane == 0) s_used = used;
    }
    __syncthreads();
  }
  const bool connected = (s_used == n1);
#pragma unroll
  for (int k = 0; k < L1_EPT; ++k)
    if (eij[k] >= 0 && tfl[tid + k * L1_NT]) tmask |= 1u << k;

  // tree from node 0: adjacency lists, BFS parents, exact positions (tree sums)
  auto build_tree = [&]() {
    if (tid < MAXN) adj_c[tid] = 0;
    __syncthreads();
    for (int e = tid; e < M; e += L1_NT)
      if (tfl[e]) { atomicAdd(&adj_c[es_s[e] & 255], 1); atomicAdd(&adj_c[es_s[e] >> 8], 1); }
    __syncthreads();
    if (wid == 0) {
      const int deg = (lane < N) ? adj_c[lane] : 0;
      int off = deg;
      for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(off, o, 64);
        if (lane >= o) off += t;
      }
      off -= deg;
      if (lane < N) { adj_o[lane] = off; adj_c[lane] = off; }
      if (lane == N - 1) adj_o[N] = off + deg;
    }
    __syncthreads();
    for (int e = tid; e < M; e += L1_NT)
      if (tfl[e]) {
        const int p0 = atomicAdd(&adj_c[es_s[e] & 255], 1);
        if (p0 < 2 * MAXN) adj[p0] = e;
        const int p1 = atomicAdd(&adj_c[es_s[e] >> 8], 1);
        if (p1 < 2 * MAXN) adj[p1] = e;
      }
    __syncthreads();
    if (wid == 0) {
      if (lane < N) { par_s[lane] = -1; ped_s[lane] = -1; }
      if (lane == 0) pvs[0] = 0.0;
      unsigned long long seen = 1ull;
      for (int round = 0; round < N; ++round) {
        bool got = false;
        int np = -1, ne = -1;
        double npv = 0.0;
        if (lane < N && !((seen >> lane) & 1ull)) {
          const int qe = min(adj_o[lane + 1], 2 * MAXN);
          for (int q = adj_o[lane]; q < qe; ++q) {
            const int e = adj[q];
            const int ij = es_s[e];
            const int u = ((ij & 255) == lane) ? (ij >> 8) : (ij & 255);
            if ((seen >> u) & 1ull) {
              np = u; ne = e;
              npv = ((ij & 255) == u) ? pvs[u] + b_s[e] : pvs[u] - b_s[e];
              got = true;
              break;
            }
          }
        }
        const unsigned long long add = __ballot(got);
        if (got) { par_s[lane] = np; ped_s[lane] = ne; pvs[lane] = npv; }
        if (add == 0ull) break;
        seen |= add;
      }
    }
    __syncthreads();
  };
  // subtree masks sub[v] = {v} + descendants (wave 0): ancestor sets by
  // pointer doubling, then one ballot per frame
  auto subtree_masks = [&]() {
    if (wid == 0) {
      const int par = (lane < N) ? par_s[lane] : -1;
      unsigned long long anc = (par >= 0) ? (1ull << par) : 0ull;
      int jmp = par;
#pragma unroll
      for (int s = 0; s < 6; ++s) {
        const int src = jmp >= 0 ? jmp : lane;
        const unsigned long long anc_j = shfl_u64(anc, src);
        const int jmp_j = __shfl(jmp, src, 64);
        if (jmp >= 0) { anc |= anc_j; jmp = jmp_j; }
      }
      for (int v = 0; v < N; ++v) {
        const unsigned long long m = __ballot(lane < N && ((anc >> v) & 1ull));
        if (lane == v) sub_s[v] = m | (1ull << v);
      }
    }
    __syncthreads();
  };
  build_tree();
  subtree_masks();
  if (!connected && tid == 0) { if (axis == 0) a.info->status_y = -1; else a.info->status_x = -1; }

  // ---- dual network simplex
  double wsum = 0.0;
#pragma unroll
  for (int k = 0; k < L1_EPT; ++k) wsum += ew[k];
  wsum = wave_sum_d(wsum);
  if (lane == 0) red_d[0][wid] = wsum;
  __syncthreads();
  wsum = (red_d[0][0] + red_d[0][1]) + (red_d[0][2] + red_d[0][3]);
  const double tol = -1e-12 * (1.0 + wsum);
  for (int q = tid; q < N * LDC; q += L1_NT) Lm[q] = 0.0;   // frame pairs without an edge stay 0
  __syncthreads();
  int piv = 0, streak = 0;
  bool optimal = false;
  for (; connected && piv <= MAX_PIVOTS; ++piv) {
    const bool bland = streak > L1_BLAND * N;
    // flow bounds of the non-tree edges: C[i][j] = +w s (into j), C[j][i] = -w s,
    // summed over the parallel edges of (i, j) in edge order by the first of them
    // (tree edges carry no bound flow); per-edge values staged in cd_s
    if (tid == 0) s_nc = 0;
#pragma unroll
    for (int k = 0; k < L1_EPT; ++k) {
      if (eij[k] < 0) continue;
      const int e = tid + k * L1_NT;
      if ((tmask >> k) & 1u) {
        er[k] = 0.0;
        cd_s[e] = 0.0;
      } else {
        const int i = eij[k] & 255, j = eij[k] >> 8;
        const double r = pvs[j] - pvs[i] - eb[k];
        er[k] = r;
        const double s = r > 0.0 ? 1.0 : (r < 0.0 ? -1.0 : (((smask >> k) & 1u) ? 1.0 : -1.0));
        cd_s[e] = ew[k] * s;
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < L1_EPT; ++k) {
      const int e = tid + k * L1_NT;
      if (eij[k] < 0 || !((dfirst >> k) & 1u)) continue;
      double f = cd_s[e];
      for (int q = dn_s[e]; q >= 0; q = dn_s[q]) f += cd_s[q];
      const int i = eij[k] & 255, j = eij[k] >> 8;
      Lm[i * LDC + j] = f;
      Lm[j * LDC + i] = -f;
    }
    __syncthreads();
    // slopes of moving each subtree by +-delta: w(tree edge) +- (bound flow into the subtree)
    if (wid == 0) {
      double av = 0.0;
      if (lane < N)
        for (int x = 0; x < N; ++x) av += Lm[x * LDC + lane];
      const unsigned long long sb = (lane < N) ? sub_s[lane] : 0ull;
      // A = sum of av over the lane's subtree: av by LDS broadcast (zero past
      // N), four partial sums over u mod 4
      avs_s[lane] = lane < N ? av : 0.0;
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      double A;
      {
        double a4[4] = {0.0, 0.0, 0.0, 0.0};
        const unsigned slo = (unsigned)sb, shi = (unsigned)(sb >> 32);
#pragma unroll
        for (int u = 0; u < MAXN; ++u) {
          const double au = avs_s[u];
          const bool in = ((u < 32 ? slo >> u : shi >> (u - 32)) & 1u) != 0;
          a4[u & 3] += in ? au : 0.0;
        }
        A = (a4[0] + a4[1]) + (a4[2] + a4[3]);
      }
      double bval = INFINITY;
      int bkey = 0x7fffffff;
      if (lane >= 1 && lane < N) {
        const int pe = ped_s[lane];
        const double wt = w_s[pe];
        const double sp = wt + A, sm = wt - A;
        const bool up = !(sm < sp);
        const double v = up ? sp : sm;
        if (v < tol) {
          if (bland) { bval = 0.0; bkey = 2 * pe + (up ? 0 : 1); }
          else { bval = v; bkey = 2 * lane + (up ? 0 : 1); }
        }
      }
      int bvv = lane;
      wave_lexmin(bval, bkey, bvv);
      if (lane == 0) { s_key = bkey; s_v = bvv; }
      if (bkey != 0x7fffffff && lane == bvv) {   // slope of the chosen move (also under Bland's rule)
        const double wt = w_s[ped_s[lane]];
        s_best = (bkey & 1) ? wt - A : wt + A;
      }
    }
    __syncthreads();
    if (s_key == 0x7fffffff) { optimal = true; break; }
    if (piv == MAX_PIVOTS) break;
    const int bv = s_v;
    const double bs = (s_key & 1) ? -1.0 : 1.0;
    const unsigned long long S = sub_s[bv];
    const int le = ped_s[bv];
    // breakpoints: crossing non-tree edges moving towards their other bound
    {
      int nmine = 0;
      double dk[L1_EPT];
      int ekk[L1_EPT];
#pragma unroll
      for (int k = 0; k < L1_EPT; ++k) {
        if (eij[k] < 0 || ((tmask >> k) & 1u)) continue;
        const int i = eij[k] & 255, j = eij[k] >> 8;
        const bool iin = (S >> i) & 1ull, jin = (S >> j) & 1ull;
        if (iin == jin) continue;
        const double dir = (jin ? 1.0 : -1.0) * bs;
        const double r = er[k];
        const double s = r > 0.0 ? 1.0 : (r < 0.0 ? -1.0 : (((smask >> k) & 1u) ? 1.0 : -1.0));
        if (dir * s >= 0.0) continue;
        dk[nmine] = fabs(r);
        ekk[nmine] = tid + k * L1_NT;
        ++nmine;
      }
      // wave prefix sum of the counts (nmine <= L1_EPT = 8: four bits), one
      // ballot and lane-masked bit count per bit
      int before = 0, tot = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const unsigned long long m = __ballot((nmine >> b) & 1);
        before += (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u)) << b;
        tot += __popcll(m) << b;
      }
      int base = 0;
      if (lane == 0 && tot > 0) base = atomicAdd(&s_nc, tot);
      base = __builtin_amdgcn_readfirstlane(base) + before;   // every lane is active here: lane 0's value
      for (int q = 0; q < nmine; ++q) { cd_s[base + q] = dk[q]; ce_s[base + q] = ekk[q]; }
    }
    __syncthreads();
    // line search (wave 0): pass breakpoints in (distance, edge) order until the
    // slope turns non-negative; the last one passed enters
    if (wid == 0) {
      const int nc = s_nc;
      double slope = s_best, last = 0.0;
      int last_e = -1, enter = -1;
      for (int guard = 0; guard <= nc; ++guard) {
        double md = INFINITY;
        int me = 0x7fffffff;
        for (int t = lane; t < nc; t += 64) {
          const double d = cd_s[t];
          const int e = ce_s[t];
          if (lex_less(last, last_e, d, e) && lex_less(d, e, md, me)) { md = d; me = e; }
        }
        {
          int unused = 0;
          wave_lexmin(md, me, unused);
        }
        if (me == 0x7fffffff) break;
        slope += 2.0 * w_s[me];
        last = md;
        last_e = me;
        enter = me;
        if (slope >= 0.0 || bland) break;
      }
      if (lane == 0) { s_enter = enter; s_last = last; }
    }
    __syncthreads();
    const int enter = s_enter;
    const double last = s_last;
    if (enter < 0) break;   // cannot improve (numerically flat)
    streak = (last == 0.0) ? streak + 1 : 0;
    // passed breakpoints flip sides; the entering edge joins the tree, the
    // leaving edge takes the bound on the side the subtree moved to
#pragma unroll
    for (int k = 0; k < L1_EPT; ++k) {
      if (eij[k] < 0) continue;
      const int e = tid + k * L1_NT;
      const int i = eij[k] & 255, j = eij[k] >> 8;
      const bool jin = (S >> j) & 1ull;
      if (e == le) {
        tmask &= ~(1u << k);
        const double sd = jin ? bs : -bs;
        smask = sd > 0.0 ? (smask | (1u << k)) : (smask & ~(1u << k));
        continue;
      }
      if ((tmask >> k) & 1u) continue;
      if (e == enter) { tmask |= 1u << k; continue; }
      const bool iin = (S >> i) & 1ull;
      if (iin == jin) continue;
      const double dir = (jin ? 1.0 : -1.0) * bs;
      const double r = er[k];
      const double s = r > 0.0 ? 1.0 : (r < 0.0 ? -1.0 : (((smask >> k) & 1u) ? 1.0 : -1.0));
      if (dir * s >= 0.0) continue;
      if (!lex_less(last, enter, fabs(r), e)) smask = (s < 0.0) ? (smask | (1u << k)) : (smask & ~(1u << k));
    }
    // tree update (wave 0, lane v = frame v).  S moves by bs * last and hangs
    // from s_out through the entering edge; the path s_in .. bv reverses.
    //   path node v (ancestor-or-self of s_in inside S): new parent = its child
    //   on the path (s_out for s_in), new subtree = S minus that child's old subtree;
    //   outside S: old ancestors of bv lose S, s_out and its ancestors gain S.
    if (wid == 0) {
      const int ij = es_s[enter];
      const int s_in = ((S >> (ij & 255)) & 1ull) ? (ij & 255) : (ij >> 8);
      const int s_out = (s_in == (ij & 255)) ? (ij >> 8) : (ij & 255);
      const bool live = lane < N;
      const int op = live ? par_s[lane] : -1, oe = live ? ped_s[lane] : -1;
      const unsigned long long osub = live ? sub_s[lane] : 0ull;
      const bool inS = live && ((S >> lane) & 1ull);
      const bool onpath = inS && ((osub >> s_in) & 1ull);
      if (onpath && lane != bv) { t_par[op] = lane; t_ped[op] = oe; t_sub[op] = osub; }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      int np = op, ne = oe;
      unsigned long long ns = osub;
      if (onpath) {
        if (lane == s_in) { np = s_out; ne = enter; ns = S; }
        else { np = t_par[lane]; ne = t_ped[lane]; ns = S & ~t_sub[lane]; }
      } else if (live && !inS) {
        if ((osub >> bv) & 1ull) ns &= ~S;
        if ((osub >> s_out) & 1ull) ns |= S;
      }
      if (live) {
        par_s[lane] = np;
        ped_s[lane] = ne;
        sub_s[lane] = ns;
        if (inS) pvs[lane] += bs * last;
      }
    }
    __syncthreads();
  }
  // exact positions by tree sums along the final tree
#pragma unroll
  for (int k = 0; k < L1_EPT; ++k)
    if (eij[k] >= 0) tfl[tid + k * L1_NT] = (tmask >> k) & 1u;
  __syncthreads();
  build_tree();
  if (tid == 0) {
    if (axis == 0) { a.info->pivots_y = piv; if (a.info->status_y == 0) a.info->status_y = optimal ? 1 : 2; }
    else { a.info->pivots_x = piv; if (a.info->status_x == 0) a.info->status_x = optimal ? 1 : 2; }
  }
  for (int k = tid; k < N; k += L1_NT) a.pos[axis * N + k] = pvs[k];
#pragma unroll
  for (int k = 0; k < L1_EPT; ++k) {
    if (eij[k] < 0) continue;
    const int e = tid + k * L1_NT;
    const int i = eij[k] & 255, j = eij[k] >> 8;
    a.resid[e * 2 + axis] = pvs[j] - pvs[i] - eb[k];
  }
}

// ------------------------------------------------------------------ S3
ZR_DEV double sinc_d(double x) {
  double y = M_PI * (x == 0.0 ? 1.0e-20 : x);
  return sin(y) / y;
}

struct PostArgs {
  const double* pos;      // [2][N]
  const double* weights;  // [M]
  const double* resid;    // [M][2]
  const int2* meas;       // [M] trimat entry of each equation
  int N, M;
  int centroid, lanczos, subpix;
  int H, W, box;
  double overlap;
  double* trans_raw;      // [N][2]
  double* trans;          // [N][2] (centred)
  double* rel;            // [N-1][2]
  double* err_stats;      // [2]
  double* err_frames;     // [N][2]  (errorY, errorX)
  FrameShift* fs;
  SolveInfo* info;
};

// Modes without a solver: triMode 'first' takes translations = -shiftsTriMat[0, :]
// (zorro/zorro.py:1628-1629), 'refine' -shiftsTriMat[K, K] (:1639-1643).
// Positions p = -translations.
__global__ __launch_bounds__(64) void k_direct_positions(const float* __restrict__ shifts, int N, int refine,
                                                         double* __restrict__ pos, SolveInfo* info) {
  for (int k = threadIdx.x; k < N; k += 64) {
    const size_t e = refine ? ((size_t)k * N + k) : (size_t)k;
    pos[k] = (double)shifts[e * 2 + 0];
    pos[N + k] = (double)shifts[e * 2 + 1];
  }
  if (threadIdx.x == 0) { info->status_x = 1; info->status_y = 1; }
}

constexpr int POST_NT = 256;
__global__ __launch_bounds__(POST_NT) void k_solution_post(PostArgs a) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int N = a.N, M = a.M;
  __shared__ double cen[2];
  if (tid == 0) {
    double s0 = 0.0, s1 = 0.0;
    for (int k = 0; k < N; ++k) {
      double ty = -a.pos[k], tx = -a.pos[N + k];
      a.trans_raw[2 * k] = ty;
      a.trans_raw[2 * k + 1] = tx;
      s0 += ty;
      s1 += tx;
      if (k + 1 < N) {
        a.rel[2 * k] = a.pos[k + 1] - a.pos[k];
        a.rel[2 * k + 1] = a.pos[N + k + 1] - a.pos[N + k];
      }
    }
    cen[0] = a.centroid ? s0 / N : 0.0;
    cen[1] = a.centroid ? s1 / N : 0.0;
  }
  if (wid == 0) {
    // error statistics (weights clipped to >= 1e-6)
    double sw = 0, swe = 0, swe2 = 0;
    for (int e = lane; e < M; e += 64) {
      double w = fmax(a.weights[e], 1e-6);
      double ey = a.resid[2 * e], ex = a.resid[2 * e + 1];
      double en = sqrt(ey * ey + ex * ex);
      sw += w; swe += w * en; swe2 += w * en * en;
    }
    sw = wave_sum_d(sw); swe = wave_sum_d(swe); swe2 = wave_sum_d(swe2);
    if (lane == 0 && M > 0) {
      double mean = swe / sw;
      a.err_stats[0] = mean;
      a.err_stats[1] = sqrt(swe2 / sw - mean * mean);
    }
  }
  __syncthreads();
  for (int k = tid; k < N; k += POST_NT) {
    double ty = a.trans_raw[2 * k] - cen[0], tx = a.trans_raw[2 * k + 1] - cen[1];
    a.trans[2 * k] = ty;
    a.trans[2 * k + 1] = tx;
    FrameShift f;
    f.iy = (int)rint(ty);
    f.ix = (int)rint(tx);
    f.pad = 0;
    double ry = ty - floor(ty), rx = tx - floor(tx);
    f.conv = (a.subpix > 1 && a.lanczos && !(fabs(ry) <= 1e-8 && fabs(rx) <= 1e-8)) ? 1 : 0;
    if (f.conv) {
      double xr[3], yr[3], kk[9], s = 0.0;
      for (int i = 0; i < 3; ++i) {
        xr[i] = (2.0 / 3.0) * ((-1.5 + i) + 1.0 - rx);
        yr[i] = (2.0 / 3.0) * ((-1.5 + i) + 1.0 - ry);
      }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          kk[i * 3 + j] = sinc_d(xr[j] * 2.0) * sinc_d(xr[j]) * sinc_d(yr[i] * 2.0) * sinc_d(yr[i]);
          s += kk[i * 3 + j];
        }
      for (int i = 0; i < 9; ++i) f.k[i] = (float)(kk[i] / s);
      double su = 0.0, sv = 0.0, uu[3], vv[3];
      for (int i = 0; i < 3; ++i) {
        uu[i] = sinc_d(yr[i] * 2.0) * sinc_d(yr[i]);
        vv[i] = sinc_d(xr[i] * 2.0) * sinc_d(xr[i]);
        su += uu[i];
        sv += vv[i];
      }
      for (int i = 0; i < 3; ++i) { f.u[i] = (float)(uu[i] / su); f.v[i] = (float)(vv[i] / sv); }
    } else {
      for (int i = 0; i < 9; ++i) f.k[i] = 0.f;
      for (int i = 0; i < 3; ++i) { f.u[i] = 0.f; f.v[i] = 0.f; }
    }
    a.fs[k] = f;
  }
  // per-frame error estimates with the reference's (J, J-1) indexing:
  // padded[i][j] holds the equation (i, j) for frame pair (i, j), i < j.  One
  // wave per frame, the equations split over its lanes (fixed reduction order)
  for (int k = wid; k < N; k += POST_NT / 64) {
    double nx = 0, ny = 0, dn = 0;
    const int kc = (k - 1 + N) % N;
    for (int e = lane; e < M; e += 64) {
      const int2 ed = a.meas[e];
      const bool row = (ed.x == k);                     // padded[k][:]
      const bool col = (ed.y == kc);                    // padded[:][k-1] (k=0 -> last column)
      if (row || col) {
        const double w = fmax(a.weights[e], 1e-6);
        const double c = (row ? 1.0 : 0.0) + (col ? 1.0 : 0.0);
        nx += c * w * fabs(a.resid[2 * e + 1]);
        ny += c * w * fabs(a.resid[2 * e]);
        dn += c * w;
      }
    }
    nx = wave_sum_d(nx);
    ny = wave_sum_d(ny);
    dn = wave_sum_d(dn);
    if (lane == 0) {
      a.err_frames[2 * k] = dn > 0.0 ? ny / dn : 0.0;
      a.err_frames[2 * k + 1] = dn > 0.0 ? nx / dn : 0.0;
    }
  }
  __syncthreads();
  if (tid == 0) {
    // FRC crop limits (zorro/zorro.py:2264-2274) and tile counts
    double mny = INFINITY, mxy = -INFINITY, mnx = INFINITY, mxx = -INFINITY;
    bool any = false;
    for (int k = 0; k < N; ++k) {
      double ty = a.trans[2 * k], tx = a.trans[2 * k + 1];
      mny = fmin(mny, ty); mxy = fmax(mxy, ty); mnx = fmin(mnx, tx); mxx = fmax(mxx, tx);
      if (ty != 0.0 || tx != 0.0) any = true;
    }
    int c0, c1, c2, c3;
    if (!any) { c0 = c1 = c2 = c3 = 5; }
    else {
      c2 = (int)(-fmin(floor(mny), 5.0));
      c3 = (int)(-fmin(floor(mnx), 5.0));
      c0 = (int)fmax(ceil(mxy), 5.0);
      c1 = (int)fmax(ceil(mxx), 5.0);
    }
    a.info->crop[0] = c0; a.info->crop[1] = c1; a.info->crop[2] = c2; a.info->crop[3] = c3;
    double ov = a.overlap > 0.8 ? 0.8 : a.overlap;
    a.info->tiles_x = (int)floor((double)(a.W - c1 - c3 - a.box) / a.box / (1.0 - ov));
    a.info->tiles_y = (int)floor((double)(a.H - c0 - c2 - a.box) / a.box / (1.0 - ov));
    if (a.info->tiles_x < 0) a.info->tiles_x = 0;
    if (a.info->tiles_y < 0) a.info->tiles_y = 0;
  }
}

// ------------------------------------------------------------------ K3
ZR_DEV int wrapi(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}
ZR_DEV int refl1(int i, int n) { return i < 0 ? -1 - i : (i >= n ? 2 * n - 1 - i : i); }
// numpy 'symmetric' pad index for i in [0, 2n) (and beyond, periodically)
ZR_DEV int sym_idx(int i, int n) {
  int p = 2 * n;
  i %= p;
  if (i < 0) i += p;
  return i < n ? i : p - 1 - i;
}

// Aligned frame = roll by (iy, ix) then 3x3 convolution with reflect edges:
//   out[y,x] = sum_ab K[a][b] g[refl(y+1-a), refl(x+1-b)],  g[i,j] = f[(i-iy) mod H, (j-ix) mod W]
// RowSrc holds the frame as a buffer resource and the byte offsets of the three
// source rows of one output row (computed once per row; wave-uniform when the
// row is, so that they live in scalar registers).
struct RowSrc {
  __amdgpu_buffer_rsrc_t rsrc;
  int ro[3];
};
// wrap v into [0, n) for v in (-n, 2n)
ZR_DEV int wrap1(int v, int n) {
  v += (v < 0) ? n : 0;
  return v - ((v >= n) ? n : 0);
}
ZR_DEV RowSrc row_sources(const float* __restrict__ f, const FrameShift& s, int y, int H, int W) {
  RowSrc rs;
  rs.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)f, (short)0, (int)((size_t)H * W * sizeof(float)), 0x00020000);
  const int iy = s.iy;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int yy = s.conv ? wrap1(refl1(y + 1 - a, H) - iy, H) : wrap1(y - iy, H);
    rs.ro[a] = __builtin_amdgcn_readfirstlane(yy * W * (int)sizeof(float));
  }
  return rs;
}
ZR_DEV float rs_load(const RowSrc& rs, int a, int col) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs.rsrc, col * 4, rs.ro[a], 0));
}
ZR_DEV float aligned_px_row(const RowSrc& rs, const FrameShift& s, int x, int W) {
  if (!s.conv) return rs_load(rs, 0, wrap1(x - s.ix, W));
  const int xm = wrap1(refl1(x + 1, W) - s.ix, W);
  const int x0 = wrap1(x - s.ix, W);
  const int xp = wrap1(refl1(x - 1, W) - s.ix, W);
  float acc = 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a)
    acc += s.k[a * 3 + 0] * rs_load(rs, a, xm) + s.k[a * 3 + 1] * rs_load(rs, a, x0) + s.k[a * 3 + 2] * rs_load(rs, a, xp);
  return acc;
}
// integer shifts reduced once per frame so |shift| < size (the wrap helpers need it)
ZR_DEV FrameShift reduce_shift(FrameShift s, int H, int W) {
  s.iy %= H;
  s.ix %= W;
  return s;
}

// Shift + sums: a thread owns one column x and SS_R consecutive rows; lanes of a
// wave own consecutive columns, so every load of the 3 x (SS_R + 2) source
// neighbourhood is one contiguous 256-byte row segment.  Frames are summed in
// order into imageSum and the even/odd half sums.  The 3x3 Lanczos kernel is
// the outer product u[a] v[b] of its row and column factors, so each source row
// is filtered once along x (h = sum_b v[b] g[b]) and each output row takes the
// three filtered rows it needs (sum_a u[a] h).  The source rows of a wave are
// wave-uniform (rolled, reflected row indices in scalar registers) and reach
// the loads as the scalar offsets of a per-frame buffer resource, so the vector
// work per frame is the column indices, 30 loads and the filter.  A frame with
// an integer shift takes u = v = (0, 1, 0), which gives exactly the rolled
// pixel.  The per-frame shifts are reduced once per block into LDS.
// Tile of a block: SS_WX waves side by side (64 SS_WX columns) times 4 / SS_WX
// waves stacked, each wave SS_R rows.
constexpr int SS_R = 16, SS_WX = 4;
template <int R, int WX>
__global__ __launch_bounds__(256) void k_shift_sum(const float* __restrict__ frames, const FrameShift* __restrict__ fs,
                                                  int N, int H, int W, float* __restrict__ img_sum,
                                                  float* __restrict__ even_sum, float* __restrict__ odd_sum,
                                                  float* __restrict__ aligned, double* __restrict__ rowsum,
                                                  int gx) {
  constexpr int WY = 4 / WX;
  static_assert(WX * WY == 4, "four waves per block");
  __shared__ FrameShift sfs[MAXN];
  __shared__ double red[4];
  for (int j = threadIdx.x; j < N; j += 256) {
    FrameShift f = reduce_shift(fs[j], H, W);
    if (!f.conv) {
#pragma unroll
      for (int i = 0; i < 3; ++i) f.u[i] = f.v[i] = (i == 1) ? 1.f : 0.f;
      f.conv = 1;
    }
    sfs[j] = f;
  }
  __syncthreads();
  // 1-D grid of gx x gy tiles, remapped so that neighbouring tiles (which share
  // halo rows and columns) run on the same XCD and meet in its L2
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int bx = lin % gx, by = lin / gx;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int x = (bx * WX + wave % WX) * 64 + (threadIdx.x & 63);
  const int xs = x < W ? x : W - 1;           // lanes past the right edge read a valid column, store nothing
  const int y0 = (by * WY + wave / WX) * R;   // wave-uniform
  const unsigned frame_bytes = (unsigned)((size_t)H * W * sizeof(float));
  double part = 0.0;
  if (y0 < H) {
    float st[R], se[R], so[R];
#pragma unroll
    for (int i = 0; i < R; ++i) { st[i] = 0.f; se[i] = 0.f; so[i] = 0.f; }
    for (int J = 0; J < N; ++J) {
      const FrameShift& s = sfs[J];
      const int iy = __builtin_amdgcn_readfirstlane(s.iy), ix = __builtin_amdgcn_readfirstlane(s.ix);
      const float u0 = s.u[0], u1 = s.u[1], u2 = s.u[2], v0 = s.v[0], v1 = s.v[1], v2 = s.v[2];
      const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(frames + (size_t)J * H * W), (short)0, (int)frame_bytes, 0x00020000);
      // byte offsets of the source columns of x+1, x, x-1 (reflected at the frame edge, then rolled)
      const int c0 = wrap1(refl1(xs + 1, W) - ix, W) * 4;
      const int c1 = wrap1(xs - ix, W) * 4;
      const int c2 = wrap1(refl1(xs - 1, W) - ix, W) * 4;
      float h[R + 2];
#pragma unroll
      for (int r = 0; r < R + 2; ++r) {      // source row of output row y0 - 1 + r
        const int yy = min(y0 - 1 + r, H);
        const int ro = __builtin_amdgcn_readfirstlane(wrap1(refl1(yy, H) - iy, H) * W * 4);
        const float g0 = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, c0, ro, 0));
        const float g1 = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, c1, ro, 0));
        const float g2 = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, c2, ro, 0));
        h[r] = v0 * g0 + v1 * g1 + v2 * g2;
      }
#pragma unroll
      for (int i = 0; i < R; ++i) {
        // out[y] = sum_a u[a] h[y+1-a]; row y+1-a sits at h[i+2-a]
        const float v = u0 * h[i + 2] + u1 * h[i + 1] + u2 * h[i];
        st[i] += v;
        if (J & 1) so[i] += v; else se[i] += v;
        if (aligned && x < W && y0 + i < H) aligned[((size_t)J * H + y0 + i) * W + x] = v;
      }
    }
    if (x < W) {
#pragma unroll
      for (int i = 0; i < R; ++i) {
        if (y0 + i < H) {
          const size_t o = (size_t)(y0 + i) * W + x;
          img_sum[o] = st[i];
          even_sum[o] = se[i];
          odd_sum[o] = so[i];
          part += st[i];
        }
      }
    }
  }
  part = wave_sum_d(part);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  if (threadIdx.x == 0) rowsum[lin] = red[0] + red[1] + red[2] + red[3];
}

// ------------------------------------------------------------------ K4
// Row pass over FRC tiles: Hann-windowed tile rows -> real FFT (box) -> kx-major.
// Hann apodisation of a box (zorro/zorro_util.py:62-104), r <= 1 inside the
// inscribed ellipse: hann[y][x] = 0.5 (1 + cos(pi r)), tabulated once per plan.
__global__ __launch_bounds__(256) void k_frc_hann(int box, float* __restrict__ hann) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= box * box) return;
  const int yl = e / box, xl = e % box;
  const double yv = (double)(yl - box / 2) / (box / 2), xv = (double)(xl - box / 2) / (box / 2);
  const double rr = xv * xv + yv * yv;
  hann[e] = rr <= 1.0 ? (float)(0.5 * (1.0 + cos(M_PI * sqrt(rr)))) : 0.f;
}

// Rows per k_frc_rows block: one group of FRC_TPR(BOX) threads per tile row
// (16 complex values per thread), 256 threads per block.
constexpr int frc_tpr(int box) { return box / 32; }
constexpr int frc_rows_per_block(int box) { return 256 / frc_tpr(box); }

template <int BOX>
__global__ __launch_bounds__(256) void k_frc_rows(const float* __restrict__ even, const float* __restrict__ odd,
                                                 const SolveInfo* __restrict__ info, int W, int tmax_x, double step,
                                                 const float2* __restrict__ twL, const float2* __restrict__ tw2L,
                                                 const float* __restrict__ hann, int tile0, float2* __restrict__ T) {
  constexpr int L = BOX / 2, NT = 256, TPR = frc_tpr(BOX), RB = frc_rows_per_block(BOX), PER = L / TPR;
  static_assert(BOX % RB == 0 && L % TPR == 0, "tile rows must split evenly");
  __shared__ float2 buf[RB][zr_fft_lds(L)];
  const int tid = threadIdx.x;
  const int g = tid / TPR, gt = tid - g * TPR;   // row group, thread in group
  constexpr int nrb = BOX / RB;
  int blk = blockIdx.x;
  const int rb = blk % nrb;
  blk /= nrb;
  const int par = blk % 2;
  const int lt = blk / 2;                  // tile within the launch's range (T holds only those)
  const int tile = tile0 + lt;
  const int ty = tile / tmax_x, tx = tile % tmax_x;
  if (ty >= info->tiles_y || tx >= info->tiles_x) return;   // block-uniform
  const int oy = (int)(ty * BOX * step + info->crop[0]);
  const int ox = (int)(tx * BOX * step + info->crop[1]);
  const float* img = par ? odd : even;
  const int yl = rb * RB + g;
  const float2* row = reinterpret_cast<const float2*>(img + (size_t)(oy + yl) * W + ox);
  const float2* hw = reinterpret_cast<const float2*>(hann + (size_t)yl * BOX);
  const bool even_x = ((((size_t)(oy + yl) * W + ox) & 1) == 0);   // float2 loads need an even offset
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int n = gt + i * TPR;
    const float2 h = hw[n];
    float2 v;
    if (even_x) {
      v = row[n];
    } else {
      const float* r1 = img + (size_t)(oy + yl) * W + ox;
      v = make_float2(r1[2 * n], r1[2 * n + 1]);
    }
    buf[g][n] = make_float2(h.x * v.x, h.y * v.y);
  }
  __shared__ float2 twl_[zr_ldstw_size(L)];
  const LdsTw tws_ = zr_ldstw_fill(twl_, twL, L, threadIdx.x, blockDim.x);
  __syncthreads();
  wg_fft<L, -1, TPR>(buf[g], tws_, gt);
  // real-FFT post-processing; T[tile][par][k][y] takes RB consecutive rows per k
  const int Wh = L + 1;
  for (int k = tid; k < Wh; k += NT) {
    const float2 w = tw2L[k];
    float2* dst = T + (((size_t)lt * 2 + par) * Wh + k) * BOX + rb * RB;
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const float2 zk = buf[r][k % L];
      const float2 zn = c_conj(buf[r][(L - k) % L]);
      const float2 e = c_scale(c_add(zk, zn), 0.5f);
      const float2 o = c_scale(c_sub(zk, zn), 0.5f);
      const float2 Ok = make_float2(o.y, -o.x);
      dst[r] = c_add(e, c_mul(w, Ok));
    }
  }
}

// Column pass: per kx, FFT the even and odd columns of the tiles of one tile
// group and accumulate Re(E conj(O)) / sqrt(|O|^2 |E|^2) in FP64 into that
// group's partial plane (k_frc_reduce adds the groups in a fixed order).
// Group g owns the tiles t = g (mod FRC_GROUPS) and adds them in ascending
// order; a launch covers the tiles [tile0, tile1) held in T, and with
// `accumulate` it continues the group's sum from `part`, so that any split of
// the tiles into consecutive ranges gives the same bits as one launch.
constexpr int FRC_GROUPS = 16;
// Tiles transformed together per k_frc_cols block (even and odd column of each,
// one group of BOX/16 threads per transform); the ratios are added in the
// tiles' order, as one tile at a time would add them.
constexpr int FRC_TB = 4;
constexpr int frc_cols_nt(int box) { return 2 * FRC_TB * (box / 16); }
template <int BOX>
__global__ __launch_bounds__(frc_cols_nt(BOX)) void k_frc_cols(const float2* __restrict__ T,
                                                              const SolveInfo* __restrict__ info, int tmax_x,
                                                              int tile0, int tile1, int accumulate,
                                                              const float2* __restrict__ twB,
                                                              double* __restrict__ part) {
  constexpr int TPC = BOX / 16, NT = frc_cols_nt(BOX), PER = BOX / NT;
  static_assert(BOX % NT == 0, "column must split evenly over the block");
  __shared__ float2 buf[2 * FRC_TB][zr_fft_lds(BOX)];     // [2 j + parity]
  __shared__ int live[FRC_TB];
  const int tid = threadIdx.x;
  const int q = tid / TPC, gt = tid - q * TPC;  // transform group: tile slot q / 2, parity q % 2
  const int Wh = BOX / 2 + 1;
  const int k = blockIdx.x % Wh, g = blockIdx.x / Wh;
  double* dst = part + (size_t)g * BOX * Wh;
  double acc[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) acc[i] = accumulate ? dst[(size_t)(tid + i * NT) * Wh + k] : 0.0;
  __shared__ float2 twl_[zr_ldstw_size(BOX)];
  const LdsTw tws_ = zr_ldstw_fill(twl_, twB, BOX, threadIdx.x, blockDim.x);
  // the group's first tile in the range
  const int first = tile0 + (((g - tile0) % FRC_GROUPS) + FRC_GROUPS) % FRC_GROUPS;
  for (int base = first; base < tile1; base += FRC_GROUPS * FRC_TB) {
    {
      const int tile = base + (q >> 1) * FRC_GROUPS;
      const int ty = tile / tmax_x, tx = tile % tmax_x;
      const bool ok = tile < tile1 && ty < info->tiles_y && tx < info->tiles_x;
      if (gt == 0 && (q & 1) == 0) live[q >> 1] = ok ? 1 : 0;
      const float2* col = T + (((size_t)(tile - tile0) * 2 + (q & 1)) * Wh + k) * BOX;
#pragma unroll
      for (int i = 0; i < BOX / TPC; ++i) {
        float2 v = make_float2(0.f, 0.f);
        if (ok) v = col[gt + i * TPC];
        buf[q][gt + i * TPC] = v;
      }
    }
    __syncthreads();
    wg_fft<BOX, -1, TPC>(buf[q], tws_, gt);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int n = tid + i * NT;
      for (int j = 0; j < FRC_TB; ++j) {
        if (!live[j]) continue;
        const float2 E = buf[2 * j][n], O = buf[2 * j + 1][n];
        const float num = E.x * O.x + E.y * O.y;
        const float den = sqrtf((O.x * O.x + O.y * O.y) * (E.x * E.x + E.y * E.y));
        acc[i] += (double)(num / den);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < PER; ++i) dst[(size_t)(tid + i * NT) * Wh + k] = acc[i];
}

__global__ __launch_bounds__(256) void k_frc_reduce(const double* __restrict__ part, size_t plane,
                                                    double* __restrict__ frc2d) {
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < plane; e += (size_t)gridDim.x * 256) {
    double a = 0.0;
    for (int g = 0; g < FRC_GROUPS; ++g) a += part[(size_t)g * plane + e];
    frc2d[e] = a;
  }
}

// Ring means of the normalised, fftshifted 2D FRC (bilinear radial weights).
// One block per output ring k < box/2; reads the half-plane via Hermitian symmetry.
__global__ __launch_bounds__(256) void k_frc_rings(const double* __restrict__ frc2d, int box, double* __restrict__ frc) {
  const int k = blockIdx.x;
  const int half = box / 2, Wh = half + 1;
  const double norm = frc2d[0];
  double num = 0.0, den = 0.0;
  // only pixels with k - 1 <= r < k + 1 carry weight to ring k: per row yy the
  // columns |xx| in [sqrt(max(0, (k-1)^2 - yy^2)), sqrt((k+1)^2 - yy^2)], widened
  // by one on each side and tested exactly
  for (int yy = -(k + 1) + (int)threadIdx.x; yy <= k + 1; yy += 256) {
    if (yy < -half || yy >= half) continue;
    const double y2 = (double)yy * yy;
    const double lo2 = (double)(k - 1) * (k - 1) - y2, hi2 = (double)(k + 1) * (k + 1) - y2;
    if (hi2 < 0.0) continue;
    const int xlo = lo2 > 0.0 ? max(0, (int)floor(sqrt(lo2)) - 1) : 0;
    const int xhi = (int)ceil(sqrt(hi2)) + 1;
    for (int ax = xlo; ax <= xhi; ++ax) {
      for (int sgn = 0; sgn < 2; ++sgn) {
        if (sgn == 1 && ax == 0) continue;
        const int xx = sgn ? -ax : ax;
        if (xx < -half || xx >= half) continue;
        double r = sqrt(y2 + (double)xx * xx);
        double rf = floor(r);
        double rem = r - rf;
        int ri = (int)rf;
        double wgt = 0.0;
        if (ri == k) wgt += 1.0 - rem;
        if (ri + 1 == k) wgt += rem;
        if (wgt == 0.0) continue;
        // fftshifted index (yy+half, xx+half) <-> FFT frequency (yy, xx)
        int fy = (yy + box) % box, fx = (xx + box) % box;
        double v;
        if (fx <= half) v = frc2d[(size_t)fy * Wh + fx];
        else v = frc2d[(size_t)((box - fy) % box) * Wh + (box - fx)];
        num += wgt * (v / norm);
        den += wgt;
      }
    }
  }
  __shared__ double sn[4], sd[4];
  num = wave_sum_d(num);
  den = wave_sum_d(den);
  if ((threadIdx.x & 63) == 0) { sn[threadIdx.x >> 6] = num; sd[threadIdx.x >> 6] = den; }
  __syncthreads();
  if (threadIdx.x == 0) frc[k] = (sn[0] + sn[1] + sn[2] + sn[3]) / (sd[0] + sd[1] + sd[2] + sd[3]);
}

// Column kx of frame J is past the frame's cut-off everywhere when its
// smallest frequency (ky = 0, computed as the column pass computes q) is:
// the dose filter is exactly 0 on the whole column, which then adds nothing.
ZR_DEV bool dose_column_dead(int kx, double ipx, double thresq, float cut_skip) {
  const double fx = kx;
  const float q0 = (float)sqrt(fx * fx * ipx * ipx);
  return q0 / (float)thresq > cut_skip;
}

// ------------------------------------------------------------------ KD
// Direct dose path, row pass over a chunk of frames: the aligned frame (roll +
// 3x3 Lanczos kernel), symmetric-padded to Hp x WP, real FFT of each padded
// row, written transposed in column pairs T[fb][k/2][Y][k%2] so that a thread
// stores RB rows x 2 columns (16 RB contiguous bytes).  One block = RB padded
// rows of one frame, one group of TPR threads per row (two rows at 5760, where
// four would leave one block per CU).
constexpr int dose_rows_rb(int wp) { return wp == 5760 ? 2 : (wp >= 3072 ? 4 : 8); }
constexpr int dose_rows_tpr(int wp) {
  return (wp >= 2048 && (wp / 2) % 128 == 0) ? 128 : ((wp / 2) % 192 == 0 ? 192 : 64);
}
template <int WP>
__global__ __launch_bounds__(dose_rows_rb(WP) * dose_rows_tpr(WP)) void k_dose_rows(
    const float* __restrict__ frames, const FrameShift* __restrict__ fs, int j0, int H, int W, int Hp,
    const float2* __restrict__ twL, const float2* __restrict__ tw2L, const double* __restrict__ thresq,
    double ipx, float cut_skip, float2* __restrict__ T) {
  constexpr int L = WP / 2, RB = dose_rows_rb(WP), TPR = dose_rows_tpr(WP), NT = RB * TPR, PER = L / TPR;
  static_assert(L % TPR == 0, "row length must split evenly over the row group");
  extern __shared__ float2 kd_lds[];
  float2 (*buf)[zr_fft_lds(L)] = reinterpret_cast<float2 (*)[zr_fft_lds(L)]>(kd_lds);
  const int tid = threadIdx.x;
  const int g = tid / TPR, gt = tid - g * TPR;
  const int nrb = Hp / RB;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int fb = blk / nrb, y0 = (blk % nrb) * RB;
  const int J = j0 + fb;
  const int Y = y0 + g;                        // padded row
  const FrameShift sh = reduce_shift(fs[J], H, W);
  const RowSrc rs = row_sources(frames + (size_t)J * H * W, sh, sym_idx(Y, H), H, W);
  // the padded row as WP floats in natural order (complex element n holds
  // pixels 2n, 2n+1), lanes on consecutive pixels so that every source-row
  // load of a wave is one contiguous segment
  float* rowf = reinterpret_cast<float*>(buf[g]);
#pragma unroll
  for (int i = 0; i < 2 * PER; ++i) {
    const int X = gt + i * TPR;
    rowf[X] = aligned_px_row(rs, sh, X < W ? X : (X < 2 * W ? 2 * W - 1 - X : sym_idx(X, W)), W);
  }
  __shared__ float2 twl_[zr_ldstw_size(L)];
  const LdsTw tws_ = zr_ldstw_fill(twl_, twL, L, threadIdx.x, blockDim.x);
  __syncthreads();
  wg_fft<L, -1, TPR>(buf[g], tws_, gt);
  const int Wh = L + 1, KP = (Wh + 1) / 2;
  const double tq = thresq[J];
  for (int kp = tid; kp < KP; kp += NT) {
    // a pair of columns past the frame's cut-off is never read by the column
    // pass (the second column's frequencies are higher than the first's)
    if (dose_column_dead(2 * kp, ipx, tq, cut_skip)) continue;
    float2 out[2][RB];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int k = 2 * kp + h;
      const bool live = k < Wh;
      const int kk = live ? k : 0;
      const float2 w = tw2L[kk];
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const float2 zk = buf[r][kk % L];
        const float2 zn = c_conj(buf[r][(L - kk) % L]);
        const float2 e = c_scale(c_add(zk, zn), 0.5f);
        const float2 o = c_scale(c_sub(zk, zn), 0.5f);
        const float2 Ok = make_float2(o.y, -o.x);
        out[h][r] = live ? c_add(e, c_mul(w, Ok)) : make_float2(0.f, 0.f);
      }
    }
    float4* dst = reinterpret_cast<float4*>(T + (((size_t)fb * KP + kp) * Hp + y0) * 2);
#pragma unroll
    for (int r = 0; r < RB; ++r) dst[r] = make_float4(out[0][r].x, out[0][r].y, out[1][r].x, out[1][r].y);
  }
}

// (q/qc)^order: repeated squaring for power-of-two orders (the default 32),
// powf otherwise
ZR_DEV float cutoff_pow(float x, float ord, int sq) {
  if (sq < 0) return powf(x, ord);
  for (int i = 0; i < sq; ++i) x *= x;
  return x;
}

struct DoseArgs {
  int N, H, W, Hp, Wp;
  double A, B, C, order, missing, vs, ipx, ipy;
  const double* rowsum;
  int nrows;
  double* accum;    // [N+1]
  double* thresq;   // [N]
  float* fjc;       // [N][2]: float(1 / thresq), float(-accum[J+1] / 2) (the spectral pass's per-frame constants)
  SolveInfo* info;
  int order_sq;     // order = 2^order_sq, or -1
  int normalize;    // doseFilter(normalize=True): scale by 1 / (sum over frames of the filter + 1)
  // (q/qc) above which the cut-off filter exp(-(q/qc)^order) is exactly zero in
  // float32 ((q/qc)^order >= 200; expf underflows below -104): those
  // frequencies of a frame add nothing to the filtered sum
  float cut_skip;
};


// dose per frame from mean(imageSum)/N, accumulated dose, per-frame cutoff
// radius: the last radial index whose optimal exposure is closer to the dose at
// the end of the frame than at its start.  One block of 256 threads.
__global__ __launch_bounds__(256) void k_dose_params(DoseArgs a) {
  __shared__ double red[4];
  const int tid = threadIdx.x;
  double s = 0.0;
  for (int i = tid; i < a.nrows; i += 256) s += a.rowsum[i];
  s = wave_sum_d(s);
  if ((tid & 63) == 0) red[tid >> 6] = s;
  __syncthreads();
  if (tid == 0) {
    const double tot = red[0] + red[1] + red[2] + red[3];
    const double dpf = tot / ((double)a.H * a.W) / a.N;
    const double missing = dpf * a.missing;
    a.info->dpf = dpf;
    a.accum[0] = missing;
    double run = 0.0;
    for (int j = 1; j <= a.N; ++j) { run += dpf; a.accum[j] = run + missing; }
  }
  // optimal exposure at radial index k < Hp/2, one pow per index into the
  // dynamic LDS table (the launch sizes it for Hp/2 doubles); index 0 takes
  // the square of index 1 (zorro/zorro.py:2422)
  extern __shared__ double opt_tab[];
  const int nq = a.Hp / 2;
  const double qs = sqrt(a.ipx * a.ipy);
  for (int k = tid; k < nq; k += 256) {
    const double o1 = 2.51284 * a.vs * (a.A * pow(qs, a.B) + a.C);
    opt_tab[k] = k == 0 ? o1 * o1 : 2.51284 * a.vs * (a.A * pow(k * qs, a.B) + a.C);
  }
  __syncthreads();
  // one wave per frame: the last radial index closer to the frame's end dose
  const int lane = tid & 63, wid = tid >> 6;
  for (int j = wid; j < a.N; j += 4) {
    const double d1 = a.accum[j + 1], d0 = a.accum[j];
    int last = -1;
    for (int k = lane; k < nq; k += 64) {
      const double o = opt_tab[k];
      if (fabs(d1 - o) < fabs(d0 - o)) last = k;
    }
    for (int o = 32; o > 0; o >>= 1) last = max(last, __shfl_xor(last, o, 64));
    if (lane == 0) {
      const double tq = last >= 0 ? last * qs : NAN;
      a.thresq[j] = tq;
      a.fjc[2 * j] = (float)(1.0 / tq);
      a.fjc[2 * j + 1] = (float)(-0.5 * a.accum[j + 1]);
    }
  }
}

// Column pass: per kx, accumulate FFT_y(T_J[kx]) * exp(-D_J/(2 Nc(q))) * exp(-(q/qc_J)^order)
// over the chunk's frames [j0, j1) into the running sum G[kx][0..HP) (frame
// order, so a split into chunks gives the bits of one pass over all frames);
// the last chunk applies the dosenorm scale and the inverse transform along y.
// T holds the chunk in the row pass's column-pair layout.
// threads of one column block: HP must split evenly over them
constexpr int dose_cols_nt(int hp) { return hp % 256 == 0 ? 256 : 384; }
template <int HP>
__global__ __launch_bounds__(dose_cols_nt(HP)) void k_dose_cols(const float2* __restrict__ T, DoseArgs a, int j0,
                                                               int j1, const float2* __restrict__ twH,
                                                               float2* __restrict__ G) {
  constexpr int NT = dose_cols_nt(HP), PER = HP / NT;
  static_assert(HP % NT == 0, "column length must split evenly over the block");
  __shared__ float2 buf[zr_fft_lds(HP)];
  const int tid = threadIdx.x;
  const int kx = xcd_remap(blockIdx.x, gridDim.x);
  const int Wh = a.Wp / 2 + 1, KP = (Wh + 1) / 2;
  float invc[PER], q[PER];
  float2 acc[PER];
  float2* gcol = G + (size_t)kx * HP;
  for (int i = 0; i < PER; ++i) {
    const int ky = tid + i * NT;
    const double fy = ky < HP / 2 ? ky : ky - HP;
    const double fx = kx;     // kx = Wp/2 is the Nyquist column (|f| identical)
    auto crit = [&](double yy, double xx) {
      double qq = sqrt(xx * xx * a.ipx * a.ipx + yy * yy * a.ipy * a.ipy);
      return a.vs * (a.A * pow(qq, a.B) + a.C);
    };
    double qq = sqrt(fx * fx * a.ipx * a.ipx + fy * fy * a.ipy * a.ipy);
    double c = crit(fy, fx);
    if (ky == HP / 2 && 2 * kx == a.Wp) {   // Nyquist corner takes its left neighbour squared
      double cl = crit(-HP / 2, a.Wp / 2 - 1);
      c = cl * cl;
    }
    invc[i] = (float)(1.0 / c);   // q = 0 -> c = inf -> 0
    q[i] = (float)qq;
    acc[i] = j0 > 0 ? gcol[ky] : make_float2(0.f, 0.f);
  }
  const float ord = (float)a.order;
  __shared__ float2 twl_[zr_ldstw_size(HP)];
  const LdsTw tws_ = zr_ldstw_fill(twl_, twH, HP, threadIdx.x, blockDim.x);
  for (int J = j0; J < j1; ++J) {
    // frame J's filter is exactly 0 on this whole column: it adds +0 (the row
    // pass wrote nothing there), so its load and transform are skipped
    if (dose_column_dead(kx, a.ipx, a.thresq[J], a.cut_skip)) continue;
    const float2* src = T + (((size_t)(J - j0) * KP + kx / 2) * HP) * 2 + (kx & 1);
    for (int i = 0; i < PER; ++i) buf[tid + i * NT] = src[2 * (tid + i * NT)];
    __syncthreads();
    wg_fft<HP, -1, NT>(buf, tws_, tid);
    const float mhd = (float)(-0.5 * a.accum[J + 1]);
    const float tq = (float)a.thresq[J];
    for (int i = 0; i < PER; ++i) {
      float filt = expf(mhd * invc[i]);
      float thr = expf(-cutoff_pow(q[i] / tq, ord, a.order_sq));
      float2 v = buf[tid + i * NT];
      float g = thr * filt;
      acc[i].x += v.x * g;
      acc[i].y += v.y * g;
    }
    __syncthreads();
  }
  if (j1 < a.N) {   // more chunks follow: keep the running sum
    for (int i = 0; i < PER; ++i) gcol[tid + i * NT] = acc[i];
    return;
  }
  if (a.normalize) {
    // filterMag = 1 / (sum of the filters + 1), float32 (zorro/zorro.py:2480-2486); the
    // filters depend only on the frequency, so their sum is formed here in frame order
    for (int i = 0; i < PER; ++i) {
      float gsum = 0.f;
      for (int J = 0; J < a.N; ++J) {
        const float filt = expf((float)(-0.5 * a.accum[J + 1]) * invc[i]);
        const float thr = expf(-cutoff_pow(q[i] / (float)a.thresq[J], ord, a.order_sq));
        gsum += thr * filt;
      }
      acc[i] = c_scale(acc[i], 1.f / (gsum + 1.f));
    }
  }
  for (int i = 0; i < PER; ++i) buf[tid + i * NT] = acc[i];
  __syncthreads();
  wg_fft<HP, 1, NT>(buf, tws_, tid);
  for (int y = tid; y < a.H; y += NT) gcol[y] = buf[y];
}

// Final rows: c2r of the accumulated half spectrum (raw), modulus, crop to W.
template <int WP>
__global__ __launch_bounds__(256) void k_dose_out(const float2* __restrict__ G, int Gs, int H, int W,
                                                 const float2* __restrict__ twL, const float2* __restrict__ tw2L,
                                                 float* __restrict__ filt_sum) {
  constexpr int L = WP / 2, NT = 256;
  __shared__ float2 buf[zr_fft_lds(L)];
  const int y = xcd_remap(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  for (int k = tid; k < L; k += NT) {
    float2 xk = G[(size_t)k * Gs + y];
    float2 xn = c_conj(G[(size_t)(L - k) * Gs + y]);
    float2 e = c_scale(c_add(xk, xn), 0.5f);
    float2 d = c_scale(c_sub(xk, xn), 0.5f);
    float2 w = tw2L[k];
    float2 o = c_mul(d, make_float2(w.x, -w.y));    // * exp(+2 pi i k / Wp)
    buf[k] = make_float2(e.x - o.y, e.y + o.x);      // E + i O
  }
  __shared__ float2 twl_[zr_ldstw_size(L)];
  const LdsTw tws_ = zr_ldstw_fill(twl_, twL, L, threadIdx.x, blockDim.x);
  __syncthreads();
  wg_fft<L, 1, NT>(buf, tws_, tid);
  for (int n = tid; n < L; n += NT) {
    float2 z = buf[n];
    int x0 = 2 * n, x1 = 2 * n + 1;
    if (x0 < W) filt_sum[(size_t)y * W + x0] = fabsf(2.f * z.x);
    if (x1 < W) filt_sum[(size_t)y * W + x1] = fabsf(2.f * z.y);
  }
}

// ---------------------------------------------------------------- KD (spectral)
// For unpadded frames the dose filter is applied to spectra built from the
// forward spectra F_J the correlation already computed:
//   FFT(aligned_J)(k) = F_J(k) * Py_J(ky) * Px_J(kx) + Ebar_J(k)
// Py/Px carry the circular roll (phase ramp) and the separable 3x3 Lanczos
// kernel as a circular convolution; Ebar_J is the DFT of the difference between
// the reference's reflect boundaries and circular ones, which is non-zero only
// on the four border lines of the frame:
//   Ebar(ky,kx) = R0(kx) + RH(kx) e^{-2 pi i ky (H-1)/H} + C0(ky) + CW(ky) e^{-2 pi i kx (W-1)/W}
// (R0/RH: rows 0 and H-1, C0/CW: columns 0 and W-1 without the corners).

// Border difference value at border pixel (y, x).
ZR_DEV float border_diff(const float* __restrict__ f, const FrameShift& s, int y, int x, int H, int W) {
  float acc = 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int yr = y + 1 - a;
    const int yref = wrap1(refl1(yr, H) - s.iy, H), ycir = wrap1(wrap1(yr, H) - s.iy, H);
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const int xr = x + 1 - b;
      const int xref = wrap1(refl1(xr, W) - s.ix, W), xcir = wrap1(wrap1(xr, W) - s.ix, W);
      acc += s.k[a * 3 + b] * (f[(size_t)yref * W + xref] - f[(size_t)ycir * W + xcir]);
    }
  }
  return acc;
}

// One block per (frame, border line): line 0 = row 0, 1 = row H-1 (half spectra
// over kx), 2 = column 0, 3 = column W-1 (corners excluded; all ky).  Blocks of
// lines 0 and 2 also write the frame's Px and Py factors.
template <int LEN>
__global__ __launch_bounds__(256) void k_dose_border(const float* __restrict__ frames, const FrameShift* __restrict__ fs,
                                                    int H, int W, const float2* __restrict__ tw,
                                                    const float2* __restrict__ twH, const float2* __restrict__ twW,
                                                    float2* __restrict__ R0, float2* __restrict__ RH,
                                                    float2* __restrict__ C0, float2* __restrict__ CW,
                                                    float2* __restrict__ Py, float2* __restrict__ Px) {
  constexpr int NT = 256;
  __shared__ float2 buf[zr_fft_lds(LEN)];
  const int J = blockIdx.x / 4, line = blockIdx.x % 4;
  const FrameShift s = reduce_shift(fs[J], H, W);
  const float* f = frames + (size_t)J * H * W;
  const int tid = threadIdx.x;
  const int Whalf = W / 2 + 1;
  if (line < 2) {   // rows: Px on line 0
    const int y = line == 0 ? 0 : H - 1;
    for (int x = tid; x < W; x += NT) buf[x] = make_float2(s.conv ? border_diff(f, s, y, x, H, W) : 0.f, 0.f);
    __syncthreads();
    wg_fft<LEN, -1, NT>(buf, tw, tid);
    float2* dst = (line == 0 ? R0 : RH) + (size_t)J * Whalf;
    for (int k = tid; k < Whalf; k += NT) dst[k] = buf[k];
    if (line == 0) {
      for (int k = tid; k < Whalf; k += NT) {
        // e^{-2 pi i k ix / W} * sum_b v_b e^{-2 pi i k (b-1)/W}
        const int ph = (int)(((long)k * ((s.ix % W + W) % W)) % W);
        float2 ramp = twW[ph];
        float2 kh = make_float2(1.f, 0.f);
        if (s.conv) {
          float2 e1 = twW[k % W];                 // e^{-2 pi i k / W}
          float2 em = make_float2(e1.x, -e1.y);   // e^{+2 pi i k / W}
          kh = make_float2(s.v[0] * em.x + s.v[1] + s.v[2] * e1.x, s.v[0] * em.y + s.v[2] * e1.y);
        }
        Px[(size_t)J * Whalf + k] = c_mul(ramp, kh);
      }
    }
  } else {          // columns: Py on line 2
    const int x = line == 2 ? 0 : W - 1;
    for (int y = tid; y < H; y += NT) {
      float v = 0.f;
      if (s.conv && y > 0 && y < H - 1) v = border_diff(f, s, y, x, H, W);
      buf[y] = make_float2(v, 0.f);
    }
    __syncthreads();
    wg_fft<LEN, -1, NT>(buf, tw, tid);
    float2* dst = (line == 2 ? C0 : CW) + (size_t)J * H;
    for (int k = tid; k < H; k += NT) dst[k] = buf[k];
    if (line == 2) {
      for (int k = tid; k < H; k += NT) {
        const int ph = (int)(((long)k * ((s.iy % H + H) % H)) % H);
        float2 ramp = twH[ph];
        float2 kh = make_float2(1.f, 0.f);
        if (s.conv) {
          float2 e1 = twH[k % H];
          float2 em = make_float2(e1.x, -e1.y);
          kh = make_float2(s.u[0] * em.x + s.u[1] + s.u[2] * e1.x, s.u[0] * em.y + s.u[2] * e1.y);
        }
        Py[(size_t)J * H + k] = c_mul(ramp, kh);
      }
    }
  }
}

// Spectral dose path, step 1: the filtered sum of the aligned frame spectra at
// every (kx, ky) of the half spectrum, kx-major into G (the column inverse
// transforms of step 2 run in place there).  Every (kx, ky) is independent, so
// the work is tiled finely over the chip instead of one column per block (a
// 2049-column half spectrum over 256 CUs would leave one column for a round of
// its own): a block covers DA_NT rows (one per thread) of DA_U consecutive
// columns.  A row's border terms (py, c0, cw of frame J) are loaded once for
// the block's DA_U columns; the column terms (px, r0, rh) are block-uniform.
constexpr int DA_NT = 256, DA_U = 4;
template <int HP>
__global__ __launch_bounds__(DA_NT) void k_dose_accum(const float2* __restrict__ S, int Wk, DoseArgs a,
                                                      const float2* __restrict__ R0, const float2* __restrict__ RH,
                                                      const float2* __restrict__ C0, const float2* __restrict__ CW,
                                                      const float2* __restrict__ Py, const float2* __restrict__ Px,
                                                      const float2* __restrict__ twH, const float2* __restrict__ twW,
                                                      float2* __restrict__ G) {
  static_assert(HP % DA_NT == 0, "rows must split evenly over the blocks");
  constexpr int NRB = HP / DA_NT;
  const int Whalf = a.Wp / 2 + 1;
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int ky = (lin % NRB) * DA_NT + threadIdx.x;
  const int kx0 = (lin / NRB) * DA_U;
  auto crit = [&](double yy, double xx) {
    const double qq = sqrt(xx * xx * a.ipx * a.ipx + yy * yy * a.ipy * a.ipy);
    return a.vs * (a.A * pow(qq, a.B) + a.C);
  };
  float invc[DA_U], q[DA_U], gsum[DA_U];
  float2 acc[DA_U], ex[DA_U];
  int kxs[DA_U];
  const double fy = ky < HP / 2 ? ky : ky - HP;
#pragma unroll
  for (int u = 0; u < DA_U; ++u) {
    const int kx = min(kx0 + u, Whalf - 1);     // columns past the end repeat the last one, unwritten
    kxs[u] = kx;
    const double fx = kx;
    const double qq = sqrt(fx * fx * a.ipx * a.ipx + fy * fy * a.ipy * a.ipy);
    double c = crit(fy, fx);
    if (ky == HP / 2 && 2 * kx == a.Wp) {        // Nyquist corner takes its left neighbour squared
      const double cl = crit(-HP / 2, a.Wp / 2 - 1);
      c = cl * cl;
    }
    invc[u] = (float)(1.0 / c);
    q[u] = (float)qq;
    acc[u] = make_float2(0.f, 0.f);
    gsum[u] = 0.f;
    ex[u] = twW[(int)(((long)kx * (a.W - 1)) % a.W)];   // e^{-2 pi i kx (W-1)/W}
  }
  const float2 ey = twH[(int)(((long)ky * (a.H - 1)) % a.H)];   // e^{-2 pi i ky (H-1)/H}
  const float ord = (float)a.order;
  for (int J = 0; J < a.N; ++J) {
    const float itq = a.fjc[2 * J];
    // frequencies past frame J's cut-off (its filter is exactly 0 there) add
    // +0 to the sums: their loads and products are skipped, with the same bits
    float x[DA_U];
    bool live[DA_U], any = false;
#pragma unroll
    for (int u = 0; u < DA_U; ++u) {
      x[u] = q[u] * itq;
      live[u] = !(x[u] > a.cut_skip);
      any = any || live[u];
    }
    if (!any) continue;
    const float2 py = Py[(size_t)J * a.H + ky], c0 = C0[(size_t)J * a.H + ky], cw = CW[(size_t)J * a.H + ky];
    const float mhd = a.fjc[2 * J + 1];
    float2 F[DA_U];
#pragma unroll
    for (int u = 0; u < DA_U; ++u) F[u] = live[u] ? S[((size_t)J * Wk + kxs[u]) * HP + ky] : make_float2(0.f, 0.f);
#pragma unroll
    for (int u = 0; u < DA_U; ++u) {
      if (!live[u]) continue;
      const float2 px = Px[(size_t)J * Whalf + kxs[u]];
      const float2 r0 = R0[(size_t)J * Whalf + kxs[u]], rh = RH[(size_t)J * Whalf + kxs[u]];
      float2 sp = c_mul(c_mul(F[u], py), px);
      sp = c_add(sp, c_add(r0, c_mul(rh, ey)));
      sp = c_add(sp, c_add(c0, c_mul(cw, ex[u])));
      // exp(-D/(2 Nc(q))) exp(-(q/qc)^order) as one exponential (both arguments <= 0)
      const float gval = expf(fmaf(mhd, invc[u], -cutoff_pow(x[u], ord, a.order_sq)));
      gsum[u] += gval;
      acc[u].x += sp.x * gval;
      acc[u].y += sp.y * gval;
    }
  }
#pragma unroll
  for (int u = 0; u < DA_U; ++u) {
    // normalize: filterMag = 1 / (sum of the filters + 1), float32 (zorro/zorro.py:2480-2486)
    const float2 v = a.normalize ? c_scale(acc[u], 1.f / (gsum[u] + 1.f)) : acc[u];
    if (kx0 + u < Whalf) G[(size_t)(kx0 + u) * HP + ky] = v;
  }
}

// Spectral dose path, step 2: inverse transform along y of every column of G,
// in place (frames square and unpadded: H = HP).
template <int HP>
__global__ __launch_bounds__(256) void k_dose_cols_ifft(const float2* __restrict__ twH, float2* __restrict__ G) {
  constexpr int NT = 256, PER = HP / NT;
  static_assert(HP % NT == 0, "column must split evenly over the block");
  __shared__ float2 buf[zr_fft_lds(HP)];
  const int tid = threadIdx.x;
  float2* col = G + (size_t)xcd_remap(blockIdx.x, gridDim.x) * HP;
  float2 v[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) v[i] = col[tid + i * NT];
#pragma unroll
  for (int i = 0; i < PER; ++i) buf[tid + i * NT] = v[i];
  __shared__ float2 twl_[zr_ldstw_size(HP)];
  const LdsTw tws_ = zr_ldstw_fill(twl_, twH, HP, threadIdx.x, blockDim.x);
  __syncthreads();
  wg_fft<HP, 1, NT>(buf, tws_, tid);
#pragma unroll
  for (int i = 0; i < PER; ++i) col[tid + i * NT] = buf[tid + i * NT];
}

// Standalone tiledFRC: forward twiddle tables of the tile transforms
// (e^{-2 pi i k/n} from the angle in FP64) and the tile grid, written on the
// device so that the call needs no host-to-device copy.
__global__ __launch_bounds__(256) void k_frc_setup(int box, float2* __restrict__ twB, float2* __restrict__ twB2,
                                                   SolveInfo* __restrict__ info, int tiles_y, int tiles_x, int c0,
                                                   int c1, int c2, int c3) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e < box) {
    const double a = -2.0 * M_PI * (double)e / (double)box;
    twB[e] = make_float2((float)cos(a), (float)sin(a));
  }
  if (e < box / 2) {
    const double a = -2.0 * M_PI * (double)e / (double)(box / 2);
    twB2[e] = make_float2((float)cos(a), (float)sin(a));
  }
  if (e == 0) {
    info->tiles_y = tiles_y;
    info->tiles_x = tiles_x;
    info->crop[0] = c0; info->crop[1] = c1; info->crop[2] = c2; info->crop[3] = c3;
  }
}

// ---------------------------------------------------------------- Fourier shifts
// shiftMethod 'fourier' (zorro/zorro.py:1748-1772, with the shift theorem the
// reference intends): aligned_J = crop(Re(R[F_J exp(-2 pi i (fy ty + fx tx))]) /
// (Hp Wp)), F_J the spectrum of the zero-padded frame, fy = k/Hp over the
// fftshifted range (Nyquist at -1/2).  The real part of the inverse equals the
// inverse of the Hermitian part of the product, F_J (ramp(k) + conj(ramp(-k)))/2
// (F_J is Hermitian), which a real-output inverse then takes exactly.
ZR_DEV double signed_freq(int k, int n) { return (double)(k < n / 2 ? k : k - n) / (double)n; }

// columns: one block per (frame of the chunk, kx); the frame's half-spectrum
// column (plus mean x mask spectrum on the masked path, whose frames were
// centred before their transforms), times the Hermitian ramp, inverse along y;
// rows y < H into Gf[frame][kx][y]
template <int HP>
__global__ __launch_bounds__(dose_cols_nt(HP)) void k_fshift_cols(const float2* __restrict__ S, int Wk, int Hs,
                                                                  int mask_slot, const double* __restrict__ mu,
                                                                  double npix, const double* __restrict__ trans,
                                                                  int Wp, int H, int j0,
                                                                  const float2* __restrict__ twH,
                                                                  float2* __restrict__ Gf) {
  constexpr int NT = dose_cols_nt(HP), PER = HP / NT;
  __shared__ float2 buf[zr_fft_lds(HP)];
  const int tid = threadIdx.x;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int jl = blk / Wk, kx = blk % Wk, J = j0 + jl;
  const double ty = trans[2 * J], tx = trans[2 * J + 1];
  const float m = mask_slot >= 0 ? (float)(mu[2 * J] / npix) : 0.f;
  const float2* col = S + ((size_t)J * Wk + kx) * Hs;
  const float2* mcol = S + ((size_t)(mask_slot >= 0 ? mask_slot : 0) * Wk + kx) * Hs;
  const double fx = signed_freq(kx, Wp), fx2 = signed_freq((Wp - kx) % Wp, Wp);
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int ky = tid + i * NT;
    float2 v = col[ky];
    if (mask_slot >= 0) {
      const float2 mk = mcol[ky];
      v = make_float2(v.x + m * mk.x, v.y + m * mk.y);
    }
    // phases in cycles, reduced to [-1/2, 1/2] before the trigonometry
    double p1 = -(signed_freq(ky, HP) * ty + fx * tx);
    double p2 = signed_freq((HP - ky) % HP, HP) * ty + fx2 * tx;
    p1 -= rint(p1);
    p2 -= rint(p2);
    double s1, c1, s2, c2;
    sincos(2.0 * M_PI * p1, &s1, &c1);
    sincos(2.0 * M_PI * p2, &s2, &c2);
    const float2 r = make_float2((float)(0.5 * (c1 + c2)), (float)(0.5 * (s1 + s2)));
    buf[ky] = c_mul(v, r);
  }
  __shared__ float2 twl_[zr_ldstw_size(HP)];
  const LdsTw tws_ = zr_ldstw_fill(twl_, twH, HP, threadIdx.x, blockDim.x);
  __syncthreads();
  wg_fft<HP, 1, NT>(buf, tws_, tid);
  float2* dst = Gf + ((size_t)jl * Wk + kx) * H;
  for (int y = tid; y < H; y += NT) dst[y] = buf[y];
}

// rows: one block per (frame of the chunk, y < H): real-output inverse along x
// from the half spectrum (columns 0..WP/2), scaled by 1 / (Hp WP), cropped to W
template <int WP>
__global__ __launch_bounds__(256) void k_fshift_rows(const float2* __restrict__ Gf, int Wk, int H, int W, int j0,
                                                     float scale, const float2* __restrict__ twL,
                                                     const float2* __restrict__ tw2L, float* __restrict__ aligned) {
  constexpr int L = WP / 2, NT = 256;
  __shared__ float2 buf[zr_fft_lds(L)];
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int jl = blk / H, y = blk % H, J = j0 + jl;
  const int tid = threadIdx.x;
  const float2* g = Gf + (size_t)jl * Wk * H + y;
  for (int k = tid; k < L; k += NT) {
    const float2 xk = g[(size_t)k * H];
    const float2 xn = c_conj(g[(size_t)(L - k) * H]);
    const float2 e = c_scale(c_add(xk, xn), 0.5f);
    const float2 d = c_scale(c_sub(xk, xn), 0.5f);
    const float2 w = tw2L[k];
    const float2 o = c_mul(d, make_float2(w.x, -w.y));    // * exp(+2 pi i k / WP)
    buf[k] = make_float2(e.x - o.y, e.y + o.x);            // E + i O
  }
  __shared__ float2 twl_[zr_ldstw_size(L)];
  const LdsTw tws_ = zr_ldstw_fill(twl_, twL, L, threadIdx.x, blockDim.x);
  __syncthreads();
  wg_fft<L, 1, NT>(buf, tws_, tid);
  float* out = aligned + ((size_t)J * H + y) * W;
  for (int n = tid; n < L; n += NT) {
    const float2 z = buf[n];
    const int x0 = 2 * n, x1 = 2 * n + 1;
    if (x0 < W) out[x0] = 2.f * z.x * scale;
    if (x1 < W) out[x1] = 2.f * z.y * scale;
  }
}

// ---------------------------------------------------------------- gausslp
// filterMode 'gausslp' (zorro/zorro.py:1707-1711):
// scipy.ndimage.gaussian_filter(img, 3.0) of a float32 image, i.e. for axis 0
// and then axis 1 a correlation with w[j] = exp(-j^2 / 18) / sum_j (|j| <= 12:
// truncate 4), 'reflect' edges (d c b a | a b c d | d c b a), sums in FP64 with
// the symmetric tap pairs added first as scipy's correlate1d does, and each
// pass rounded to float32 (scipy writes every axis into the float32 output).
constexpr int GLP_R = 12;
struct GaussTaps { double w[GLP_R + 1]; };   // w[0] centre, w[j] = w[-j]

ZR_DEV int refl_index(int i, int n) {
  const int p = 2 * n;
  i %= p;
  if (i < 0) i += p;
  return i < n ? i : p - 1 - i;
}

// axis 0 (along y): one thread per pixel, a wave reads 64 consecutive x of each row
__global__ __launch_bounds__(256) void k_gauss_axis0(const float* __restrict__ src, int H, int W, GaussTaps t,
                                                     float* __restrict__ dst) {
  const size_t e = (size_t)xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (e >= (size_t)H * W) return;
  const int y = (int)(e / W), x = (int)(e % W);
  double acc = t.w[0] * (double)src[(size_t)y * W + x];
  for (int j = 1; j <= GLP_R; ++j)
    acc += ((double)src[(size_t)refl_index(y + j, H) * W + x] + (double)src[(size_t)refl_index(y - j, H) * W + x]) *
           t.w[j];
  dst[e] = (float)acc;
}

// axis 1 (along x): one block per row segment of 256 pixels, staged in LDS with
// its 12-pixel halo on each side
__global__ __launch_bounds__(256) void k_gauss_axis1(const float* __restrict__ src, int H, int W, GaussTaps t,
                                                     float* __restrict__ dst) {
  __shared__ float row[256 + 2 * GLP_R];
  const int segs = (W + 255) / 256;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int y = b / segs, x0 = (b % segs) * 256;
  const float* r = src + (size_t)y * W;
  for (int i = threadIdx.x; i < 256 + 2 * GLP_R; i += 256) row[i] = r[refl_index(x0 - GLP_R + i, W)];
  __syncthreads();
  const int x = x0 + threadIdx.x;
  if (x >= W) return;
  const float* c = row + GLP_R + threadIdx.x;
  double acc = t.w[0] * (double)c[0];
  for (int j = 1; j <= GLP_R; ++j) acc += ((double)c[j] + (double)c[-j]) * t.w[j];
  dst[(size_t)y * W + x] = (float)acc;
}

// ---------------------------------------------------- filterMode 'background'
// backgroundEstimate (zorro/zorro_util.py:497-549, applied at zorro/zorro.py:
// 1702-1706): the sum low-passed by gaussian_filter(sigma 4, truncate 4, reflect),
// sampled every 128 pixels (interp2_nn at integer positions), cropped by 1 at
// the start and 2 at the end of each axis, fitted by a 2D Gaussian
// g(p; x, y) = p0 + p3 exp(-(x - p1)^2/p4^2 - (y - p2)^2/p5^2) under an L1 error
// with Powell's method from p = (mean, 1, 1, mean, mean*std, mean*std), and the
// fitted surface, evaluated at ((x - W/2)/128, (y - H/2)/128), subtracted.  The
// fit's grid coordinates are x = kx - nx/2 and y = ky - nx/2 (the reference
// centres y by the x count, ycrop.shape[1]).
constexpr int BG_BIN = 128, BG_R = 16, BG_NT = 256;
constexpr int BG_MAXG = 80;          // samples per axis: frames up to 10240 pixels
struct BgTaps { double w[BG_R + 1]; };
struct BgFit { double p[6]; double f; int nfev; int status; };   // status 0: converged

// axis-0 pass of the sigma-4 Gaussian at the sample rows y = 128 ky only
__global__ __launch_bounds__(256) void k_bg_rows(const float* __restrict__ src, int H, int W, int ny, BgTaps t,
                                                 float* __restrict__ g0) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)ny * W) return;
  const int ky = (int)(e / W), x = (int)(e % W), y = ky * BG_BIN;
  double acc = t.w[0] * (double)src[(size_t)y * W + x];
  for (int j = 1; j <= BG_R; ++j)
    acc += ((double)src[(size_t)refl_index(y + j, H) * W + x] + (double)src[(size_t)refl_index(y - j, H) * W + x]) *
           t.w[j];
  g0[e] = (float)acc;
}

// axis-1 pass at the sample columns x = 128 kx of those rows: nn[ky][kx]
__global__ __launch_bounds__(256) void k_bg_sample(const float* __restrict__ g0, int W, int ny, int nx, BgTaps t,
                                                   float* __restrict__ nn) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= ny * nx) return;
  const int ky = e / nx, x = (e % nx) * BG_BIN;
  const float* r = g0 + (size_t)ky * W;
  double acc = t.w[0] * (double)r[x];
  for (int j = 1; j <= BG_R; ++j) acc += ((double)r[refl_index(x + j, W)] + (double)r[refl_index(x - j, W)]) * t.w[j];
  nn[e] = (float)acc;
}

// The L1 fit on one block.  Every thread runs the same scalar control flow of
// Powell's direction-set method (Powell 1964) with Brent's bracketing and
// line minimisation (Brent 1973), the steps scipy's minimize(method='Powell')
// takes with its defaults (xtol = ftol = 1e-4, line tolerance 100 xtol, at most
// 6000 iterations and 6000 evaluations); each objective evaluation is summed
// over the samples by the whole block in a fixed order.
struct BgObj {
  const float* c;      // cropped samples [cy][cx], LDS
  int cx, n;
  float x0, y0;        // grid coordinates of the first sample (1 - nx/2 on both axes)
  double* red;         // LDS [BG_NT / 64]
  int nfev;
  ZR_DEV double operator()(const double* p) {
    const int tid = threadIdx.x;
    const double p4 = p[4] * p[4], p5 = p[5] * p[5];
    double acc = 0.0;
    for (int i = tid; i < n; i += BG_NT) {
      const double xc = (double)(x0 + (float)(i % cx)) - p[1], yc = (double)(y0 + (float)(i / cx)) - p[2];
      acc += fabs((double)c[i] - (p[0] + p[3] * exp(-xc * xc / p4 - yc * yc / p5)));
    }
    acc = wave_sum_d(acc);
    __syncthreads();                      // the previous evaluation's reads of red are done
    if ((tid & 63) == 0) red[tid >> 6] = acc;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < BG_NT / 64; ++k) s += red[k];
    ++nfev;
    return s;
  }
};

struct BgLine {          // f(x0 + a d) along one direction
  BgObj* f;
  double x0[6], d[6];
  ZR_DEV double operator()(double a) {
    double q[6];
    for (int i = 0; i < 6; ++i) q[i] = x0[i] + a * d[i];
    return (*f)(q);
  }
};

// Brent's minimisation of line(a) from the bracket grown from (0, 1); returns
// the minimiser and its value.
ZR_DEV void bg_brent(BgLine& line, double tol, double* amin, double* fmin) {
  const double gold = 1.618034, vsmall = 1e-21, grow = 110.0;
  double xa = 0.0, xb = 1.0;
  double fa = line(xa), fb = line(xb);
  if (fa < fb) { double t = xa; xa = xb; xb = t; t = fa; fa = fb; fb = t; }
  double xc = xb + gold * (xb - xa), fc = line(xc);
  for (int it = 0; fc < fb && it <= 1000; ++it) {
    const double t1 = (xb - xa) * (fb - fc), t2 = (xb - xc) * (fb - fa);
    const double val = t2 - t1;
    const double denom = fabs(val) < vsmall ? 2.0 * vsmall : 2.0 * val;
    double w = xb - ((xb - xc) * t2 - (xb - xa) * t1) / denom, fw;
    const double wlim = xb + grow * (xc - xb);
    if ((w - xc) * (xb - w) > 0.0) {
      fw = line(w);
      if (fw < fc) { xa = xb; xb = w; fa = fb; fb = fw; break; }
      if (fw > fb) { xc = w; fc = fw; break; }
      w = xc + gold * (xc - xb);
      fw = line(w);
    } else if ((w - wlim) * (wlim - xc) >= 0.0) {
      w = wlim;
      fw = line(w);
    } else if ((w - wlim) * (xc - w) > 0.0) {
      fw = line(w);
      if (fw < fc) {
        xb = xc; xc = w; w = xc + gold * (xc - xb);
        fb = fc; fc = fw; fw = line(w);
      }
    } else {
      w = xc + gold * (xc - xb);
      fw = line(w);
    }
    xa = xb; xb = xc; xc = w;
    fa = fb; fb = fc; fc = fw;
  }
  const double mintol = 1e-11, cg = 0.3819660;
  double x = xb, w = xb, v = xb, fx = fb, fw = fb, fv = fb;
  double a = xa < xc ? xa : xc, b = xa < xc ? xc : xa;
  double deltax = 0.0, rat = 0.0;
  for (int it = 0; it < 500; ++it) {
    const double tol1 = tol * fabs(x) + mintol, tol2 = 2.0 * tol1;
    const double xmid = 0.5 * (a + b);
    if (fabs(x - xmid) < (tol2 - 0.5 * (b - a))) break;
    if (fabs(deltax) <= tol1) {
      deltax = (x >= xmid) ? a - x : b - x;
      rat = cg * deltax;
    } else {
      double t1 = (x - w) * (fx - fv), t2 = (x - v) * (fx - fw);
      double pp = (x - v) * t2 - (x - w) * t1;
      t2 = 2.0 * (t2 - t1);
      if (t2 > 0.0) pp = -pp;
      t2 = fabs(t2);
      const double dx_temp = deltax;
      deltax = rat;
      if (pp > t2 * (a - x) && pp < t2 * (b - x) && fabs(pp) < fabs(0.5 * t2 * dx_temp)) {
        rat = pp / t2;
        const double u = x + rat;
        if ((u - a) < tol2 || (b - u) < tol2) rat = (xmid - x >= 0.0) ? tol1 : -tol1;
      } else {
        deltax = (x >= xmid) ? a - x : b - x;
        rat = cg * deltax;
      }
    }
    const double u = fabs(rat) < tol1 ? (rat >= 0.0 ? x + tol1 : x - tol1) : x + rat;
    const double fu = line(u);
    if (fu > fx) {
      if (u < x) a = u; else b = u;
      if (fu <= fw || w == x) { v = w; w = u; fv = fw; fw = fu; }
      else if (fu <= fv || v == x || v == w) { v = u; fv = fu; }
    } else {
      if (u >= x) a = x; else b = x;
      v = w; w = x; x = u;
      fv = fw; fw = fx; fx = fu;
    }
  }
  *amin = x;
  *fmin = fx;
}

__global__ __launch_bounds__(BG_NT) void k_bg_fit(const float* __restrict__ nn, int ny, int nx, BgFit* __restrict__ out) {
  __shared__ float c[BG_MAXG * BG_MAXG];
  __shared__ double red[BG_NT / 64];
  const int tid = threadIdx.x;
  const int cy = ny - 3, cx = nx - 3, n = cy * cx;
  for (int i = tid; i < n; i += BG_NT) c[i] = nn[(1 + i / cx) * nx + 1 + i % cx];
  __syncthreads();
  // start: mean and population std of the cropped samples (float32 results, as numpy gives them)
  double s1 = 0.0, s2 = 0.0;
  for (int i = tid; i < n; i += BG_NT) s1 += c[i];
  s1 = wave_sum_d(s1);
  if ((tid & 63) == 0) red[tid >> 6] = s1;
  __syncthreads();
  double mean = 0.0;
  for (int k = 0; k < BG_NT / 64; ++k) mean += red[k];
  mean /= n;
  __syncthreads();
  for (int i = tid; i < n; i += BG_NT) s2 += ((double)c[i] - mean) * ((double)c[i] - mean);
  s2 = wave_sum_d(s2);
  if ((tid & 63) == 0) red[tid >> 6] = s2;
  __syncthreads();
  double var = 0.0;
  for (int k = 0; k < BG_NT / 64; ++k) var += red[k];
  const double m32 = (double)(float)mean, sd32 = (double)(float)sqrt(var / n);
  const float g0 = 1.0f - (float)(nx / 2.0);    // kx - nx/2 and ky - nx/2 at kx = ky = 1 (exact in float32)
  BgObj f{c, cx, n, g0, g0, red, 0};
  double x[6] = {m32, 1.0, 1.0, m32, (double)(float)(m32 * sd32), (double)(float)(m32 * sd32)};
  double dir[6][6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) dir[i][j] = i == j ? 1.0 : 0.0;
  const double xtol = 1e-4, ftol = 1e-4;
  const int maxit = 6000, maxfev = 6000;
  double fval = f(x), x1[6];
  for (int i = 0; i < 6; ++i) x1[i] = x[i];
  int status = 0, iter = 0;
  auto line_min = [&](const double* d, double* xnew, double* dnew) -> double {
    // returns the new value; xnew = x + a d, dnew = a d (a zero direction is kept as it is)
    bool any = false;
    for (int i = 0; i < 6; ++i) any |= d[i] != 0.0;
    if (!any) {
      for (int i = 0; i < 6; ++i) { xnew[i] = x[i]; dnew[i] = d[i]; }
      return fval;
    }
    BgLine ln{&f, {x[0], x[1], x[2], x[3], x[4], x[5]}, {d[0], d[1], d[2], d[3], d[4], d[5]}};
    double a, fa;
    bg_brent(ln, xtol * 100.0, &a, &fa);
    for (int i = 0; i < 6; ++i) { dnew[i] = a * d[i]; xnew[i] = x[i] + dnew[i]; }
    return fa;
  };
  while (true) {
    const double fx = fval;
    int bigind = 0;
    double delta = 0.0;
    for (int i = 0; i < 6; ++i) {
      const double fx2 = fval;
      double xn[6], dn[6];
      fval = line_min(dir[i], xn, dn);
      for (int k = 0; k < 6; ++k) { x[k] = xn[k]; dir[i][k] = dn[k]; }
      if (fx2 - fval > delta) { delta = fx2 - fval; bigind = i; }
    }
    ++iter;
    const double bnd = ftol * (fabs(fx) + fabs(fval)) + 1e-20;
    if (2.0 * (fx - fval) <= bnd) break;
    if (f.nfev >= maxfev) { status = 1; break; }
    if (iter >= maxit) { status = 2; break; }
    if (isnan(fx) && isnan(fval)) { status = 3; break; }
    double d1[6], x2[6];
    for (int k = 0; k < 6; ++k) { d1[k] = x[k] - x1[k]; x2[k] = 2.0 * x[k] - x1[k]; x1[k] = x[k]; }
    const double fx2 = f(x2);
    if (fx > fx2) {
      double t = 2.0 * (fx + fx2 - 2.0 * fval);
      double tmp = fx - fval - delta;
      t *= tmp * tmp;
      tmp = fx - fx2;
      t -= delta * tmp * tmp;
      if (t < 0.0) {
        double xn[6], dn[6];
        fval = line_min(d1, xn, dn);
        bool any = false;
        for (int k = 0; k < 6; ++k) { x[k] = xn[k]; any |= dn[k] != 0.0; }
        if (any) {
          for (int k = 0; k < 6; ++k) { dir[bigind][k] = dir[5][k]; dir[5][k] = dn[k]; }
        }
      }
    }
  }
  if (f.nfev >= maxfev && status == 0) status = 1;
  if (tid == 0) {
    for (int i = 0; i < 6; ++i) out->p[i] = x[i];
    out->f = fval;
    out->nfev = f.nfev;
    out->status = (isnan(fval) || status != 0) ? (status ? status : 3) : 0;
  }
}

// filtSum -= the fitted surface (in double, rounded back to float32), when the fit converged
__global__ __launch_bounds__(256) void k_bg_subtract(float* __restrict__ img, int H, int W,
                                                     const BgFit* __restrict__ fit) {
  const size_t e = (size_t)xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (e >= (size_t)H * W || fit->status != 0) return;
  const int y = (int)(e / W), x = (int)(e % W);
  const double* p = fit->p;
  const double xb = ((double)x - W / 2.0) / BG_BIN, yb = ((double)y - H / 2.0) / BG_BIN;
  const double xc = xb - p[1], yc = yb - p[2];
  const double back = p[0] + p[3] * exp(-xc * xc / (p[4] * p[4]) - yc * yc / (p[5] * p[5]));
  img[e] = (float)((double)img[e] - back);
}

// ==================================================================== host
namespace {
thread_local std::string s_err;
int serr(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  s_err = buf;
  return code;
}
#define ZS_HIP(call)                                                                                     \
  do {                                                                                                   \
    hipError_t e_ = (call);                                                                              \
    if (e_ != hipSuccess) return serr(ZR_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

std::vector<float2> fwd_table(int n) {
  std::vector<float2> t(n);
  for (int i = 0; i < n; ++i) {
    double a = -2.0 * M_PI * (double)i / (double)n;
    t[i] = make_float2((float)cos(a), (float)sin(a));
  }
  return t;
}
template <typename T>
hipError_t upload(T** p, const std::vector<T>& v) {
  hipError_t e = hipMalloc((void**)p, v.size() * sizeof(T));
  if (e != hipSuccess) return e;
  return hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

// The equations of shiftsSolver (zorro/zorro.py:1290-1320) for a pair set:
// tri[I][J-1] marks the pair (I, J); row I contributes the rows of
// rot90(tri[I:lc, I:lc], 2), lc = last marked column of row I + 1, and the
// measured shifts are the marked entries in row-major order.  Every equation
// row is a contiguous run of unknowns x_a..x_b (x_j = p_{j+1} - p_j), i.e. the
// graph edge (a, b+1) in position space; meas[e] is the trimat entry that
// holds its measurement.  For tri and diag patterns the edge of each equation
// is its own pair; other patterns can give an equation whose run differs from
// the pair whose shift it is matched with, exactly as in the reference.
// For every equation edge: the first edge with the same (i, j) and the next one
// (-1 at the end of the group), so that the solver sums parallel edges in order.
std::vector<int2> duplicate_groups(const std::vector<int2>& graph) {
  std::vector<int2> d(graph.size(), make_int2(-1, -1));
  std::map<long long, int> last;
  for (size_t e = 0; e < graph.size(); ++e) {
    const long long key = (long long)graph[e].x * 65536 + graph[e].y;
    auto it = last.find(key);
    if (it == last.end()) {
      d[e].x = (int)e;
      last.emplace(key, (int)e);
    } else {
      d[e].x = d[it->second].x;
      d[it->second].y = (int)e;
      it->second = (int)e;
    }
  }
  return d;
}

int build_equations(int N, const std::vector<int4>& pairs, std::vector<int2>& graph, std::vector<int2>& meas) {
  const int n = N - 1;
  std::vector<unsigned char> tri((size_t)n * n, 0);
  for (const int4& q : pairs) {
    if (q.x < 0 || q.y <= q.x || q.y > n) return -1;
    tri[(size_t)q.x * n + (q.y - 1)] = 1;
  }
  graph.clear();
  meas.clear();
  for (int I = 0; I < n; ++I)
    for (int c = 0; c < n; ++c)
      if (tri[(size_t)I * n + c]) meas.push_back(make_int2(I, c + 1));
  for (int I = 0; I < n; ++I) {
    int lc = -1;
    for (int c = 0; c < n; ++c)
      if (tri[(size_t)I * n + c]) lc = c + 1;
    if (lc < 0) continue;
    const int k = lc - I;
    if (k <= 0) return -1;
    for (int a = 0; a < k; ++a) {
      // rot[a][b] = tri[I + k-1-a][I + k-1-b]
      const int row = I + k - 1 - a;
      int c0 = -1, c1 = -1;
      bool gap = false;
      for (int b = 0; b < k; ++b) {
        const bool on = tri[(size_t)row * n + (I + k - 1 - b)] != 0;
        if (on) {
          if (c0 < 0) c0 = b;
          else if (c1 != b - 1) gap = true;
          c1 = b;
        }
      }
      if (c0 < 0 || gap) return -2;
      graph.push_back(make_int2(I + c0, I + c1 + 1));
    }
  }
  return graph.size() == meas.size() ? 0 : -3;
}
}  // namespace

struct zr_stage_state {
  zr_params p{};
  int N = 0, H = 0, W = 0, Hp = 0, Wp = 0, M = 0, Mcap = 0;
  bool eq_bad = false;         // the pairs' rot90 equations are not single frame spans
  size_t bytes = 0;            // device memory held
  int2 *edges = nullptr, *meas = nullptr, *dups = nullptr;
  float *hs = nullptr, *cdf = nullptr, *pdf = nullptr;
  double* obj = nullptr;
  SolveInfo* info = nullptr;
  double *weights = nullptr, *resid = nullptr, *pos = nullptr;
  double *trans_raw = nullptr, *trans = nullptr, *rel = nullptr, *err_stats = nullptr, *err_frames = nullptr;
  FrameShift* fs = nullptr;
  float *img_sum = nullptr, *even = nullptr, *odd = nullptr, *filt = nullptr;
  double* rowsum = nullptr;
  bool have_solution = false, have_sums = false, have_frc = false, have_dose = false;
  // FRC
  int box = 512, tmax_y = 0, tmax_x = 0;
  float2 *twB = nullptr, *twB2 = nullptr, *Tfrc = nullptr;
  double *frc2d = nullptr, *frc = nullptr, *frc_part = nullptr;
  float* hann = nullptr;
  // dose
  float2 *twHp = nullptr, *twWp = nullptr, *twWp2 = nullptr, *Tdose = nullptr, *G = nullptr;
  double *accum = nullptr, *thresq = nullptr;
  float* fjc = nullptr;
  float* gtmp = nullptr;       // 'gausslp': the axis-0 pass of the Gaussian
  float* bg_rows = nullptr;    // 'background': the sigma-4 axis-0 pass at the sample rows
  float* bg_nn = nullptr;      //   the samples every 128 pixels
  BgFit* bg_fit = nullptr;     //   the fitted surface
  // shiftMethod 'fourier': the aligned frames, identity shifts for the stages
  // that read them, and a chunk of column-inverted spectra
  bool fourier = false;
  float* aligned = nullptr;
  FrameShift* fs_id = nullptr;
  float2* Gf = nullptr;
  bool spectral = false;       // dose filter from the forward spectra (unpadded frames)
  float2 *R0 = nullptr, *RH = nullptr, *C0 = nullptr, *CW = nullptr, *Py = nullptr, *Px = nullptr;
};

const char* zr_stage_error() { return s_err.c_str(); }

// device allocations of a stage state, counted for the movie's memory footprint
static hipError_t smalloc(zr_stage_state* s, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e == hipSuccess) s->bytes += bytes;
  return e;
}
template <typename T>
static hipError_t supload(zr_stage_state* s, T** p, const std::vector<T>& v) {
  hipError_t e = upload(p, v);
  if (e == hipSuccess) s->bytes += v.size() * sizeof(T);
  return e;
}
size_t zr_stage_device_bytes(const zr_stage_state* s) { return s ? s->bytes : 0; }

int zr_stage_copy_sums_device(zr_stage_state* s, float* image_sum, float* filt_sum, float* even_sum, float* odd_sum,
                              hipStream_t st) {
  if (!s || !s->have_sums) return serr(ZR_ERR_STATE, "no sums");
  const size_t b = sizeof(float) * s->H * s->W;
  if (image_sum) ZS_HIP(hipMemcpyAsync(image_sum, s->img_sum, b, hipMemcpyDeviceToDevice, st));
  if (even_sum) ZS_HIP(hipMemcpyAsync(even_sum, s->even, b, hipMemcpyDeviceToDevice, st));
  if (odd_sum) ZS_HIP(hipMemcpyAsync(odd_sum, s->odd, b, hipMemcpyDeviceToDevice, st));
  if (filt_sum) {
    if (!s->have_dose) return serr(ZR_ERR_STATE, "no dose-filtered sum");
    ZS_HIP(hipMemcpyAsync(filt_sum, s->filt, b, hipMemcpyDeviceToDevice, st));
  }
  return ZR_OK;
}

void zr_stage_free(zr_stage_state* s) {
  if (!s) return;
  void* ptrs[] = {s->edges, s->meas, s->dups, s->hs, s->cdf, s->pdf, s->gtmp, s->bg_rows, s->bg_nn, s->bg_fit, s->aligned, s->fs_id, s->Gf, s->obj, s->info, s->weights, s->resid, s->pos, s->trans_raw, s->trans,
                  s->rel, s->err_stats, s->err_frames, s->fs, s->img_sum, s->even, s->odd, s->filt, s->rowsum,
                  s->twB, s->twB2, s->Tfrc, s->frc2d, s->frc, s->frc_part, s->hann, s->twHp, s->twWp, s->twWp2, s->Tdose, s->G,
                  s->accum, s->thresq, s->fjc, s->R0, s->RH, s->C0, s->CW, s->Py, s->Px};
  for (void* q : ptrs)
    if (q) hipFree(q);
  delete s;
}

int zr_stage_create(zr_stage_state** out, const zr_params& p, int N, int H, int W, int Hp, int Wp,
                    const std::vector<int4>& pairs, int spectral) {
  *out = nullptr;
  if (N > MAXN) return serr(ZR_ERR_UNSUPPORTED, "the solver handles up to %d frames (got %d)", MAXN, N);
  zr_stage_state* s = new zr_stage_state();
  s->p = p;
  s->N = N; s->H = H; s->W = W; s->Hp = Hp; s->Wp = Wp;
  s->M = s->Mcap = (int)pairs.size();
  if (s->M > 2016) { zr_stage_free(s); return serr(ZR_ERR_UNSUPPORTED, "too many equations"); }
  std::vector<int2> ed, ms;
  if (p.tri_mode == ZR_TRI_FIRST || p.tri_mode == ZR_TRI_REFINE) {
    // no solver: translations come straight from the trimats
    ed.assign(1, make_int2(0, 0));
    ms = ed;
    s->M = 0;
  } else if (build_equations(N, pairs, ed, ms) != 0) {
    // the pairs can still be correlated; zr_stage_solve refuses
    ed.assign(std::max(1, s->Mcap), make_int2(0, 1));   // placeholders: the buffers keep Mcap entries
    ms = ed;
    s->eq_bad = true;
  }
#define ZC(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { zr_stage_free(s); return serr(ZR_ERR_NOMEM, "%s: %s", #x, hipGetErrorString(e_)); } } while (0)
  ZC(supload(s, &s->edges, ed));
  ZC(supload(s, &s->dups, duplicate_groups(ed)));
  ZC(supload(s, &s->meas, ms));
  const int maxbins = std::max(1, s->Mcap / 7 + 1);
  if (maxbins > LOGI_MAXB) {    // N <= MAXN gives at most 2016 equations, 289 bins
    zr_stage_free(s);
    return serr(ZR_ERR_UNSUPPORTED, "%d equations: the logistic fit handles up to %d histogram bins", s->Mcap,
                LOGI_MAXB);
  }
  ZC(smalloc(s, (void**)&s->hs, sizeof(float) * maxbins));
  ZC(smalloc(s, (void**)&s->cdf, sizeof(float) * maxbins));
  ZC(smalloc(s, (void**)&s->pdf, sizeof(float) * maxbins));
  ZC(smalloc(s, (void**)&s->obj, sizeof(double) * LOGI_G0 * LOGI_G1 * LOGI_G2));
  ZC(smalloc(s, (void**)&s->info, sizeof(SolveInfo)));
  ZC(smalloc(s, (void**)&s->weights, sizeof(double) * std::max(1, s->Mcap)));
  ZC(smalloc(s, (void**)&s->resid, sizeof(double) * std::max(1, s->Mcap) * 2));
  ZC(smalloc(s, (void**)&s->pos, sizeof(double) * N * 2));
  ZC(smalloc(s, (void**)&s->trans_raw, sizeof(double) * N * 2));
  ZC(smalloc(s, (void**)&s->trans, sizeof(double) * N * 2));
  ZC(smalloc(s, (void**)&s->rel, sizeof(double) * N * 2));
  ZC(smalloc(s, (void**)&s->err_stats, sizeof(double) * 2));
  ZC(smalloc(s, (void**)&s->err_frames, sizeof(double) * N * 2));
  ZC(smalloc(s, (void**)&s->fs, sizeof(FrameShift) * N));
  ZC(smalloc(s, (void**)&s->img_sum, sizeof(float) * H * W));
  ZC(smalloc(s, (void**)&s->even, sizeof(float) * H * W));
  ZC(smalloc(s, (void**)&s->odd, sizeof(float) * H * W));
  ZC(smalloc(s, (void**)&s->filt, sizeof(float) * H * W));
  // partial sums of the shift+sum blocks (at most one per 64 columns x 4 rows)
  const int nrow_blocks = ((W + 63) / 64) * ((H + 3) / 4);
  static_assert(SS_R * (4 / SS_WX) >= 4 && SS_WX >= 1, "shift+sum tiles cover at least 64 x 4 pixels");
  ZC(smalloc(s, (void**)&s->rowsum, sizeof(double) * std::max(nrow_blocks, Hp / 4 + 1)));
  if (p.do_frc) {
    s->box = p.frc_box;
    if (s->box != 512 && s->box != 256 && s->box != 1024) { zr_stage_free(s); return serr(ZR_ERR_UNSUPPORTED, "FRC box %d: built sizes 256, 512, 1024", s->box); }
    double ov = p.frc_overlap > 0.8 ? 0.8 : p.frc_overlap;
    s->tmax_y = std::max(0, (int)std::floor((double)(H - s->box) / s->box / (1.0 - ov)));
    s->tmax_x = std::max(0, (int)std::floor((double)(W - s->box) / s->box / (1.0 - ov)));
    size_t nt = (size_t)std::max(1, s->tmax_y * s->tmax_x);
    ZC(supload(s, &s->twB, fwd_table(s->box)));
    ZC(supload(s, &s->twB2, fwd_table(s->box / 2)));
    ZC(smalloc(s, (void**)&s->hann, sizeof(float) * s->box * s->box));
    // on the null stream, waited for alone (other movies' streams keep running)
    hipLaunchKernelGGL(k_frc_hann, dim3((s->box * s->box + 255) / 256), dim3(256), 0, 0, s->box, s->hann);
    ZC(hipGetLastError());
    ZC(hipStreamSynchronize(nullptr));
    ZC(smalloc(s, (void**)&s->Tfrc, sizeof(float2) * nt * 2 * (s->box / 2 + 1) * s->box));
    ZC(smalloc(s, (void**)&s->frc2d, sizeof(double) * s->box * (s->box / 2 + 1)));
    ZC(smalloc(s, (void**)&s->frc_part, sizeof(double) * FRC_GROUPS * s->box * (s->box / 2 + 1)));
    ZC(smalloc(s, (void**)&s->frc, sizeof(double) * s->box));
  }
  if (p.do_dose) {
    if (!(p.pixelsize_nm > 0)) { zr_stage_free(s); return serr(ZR_ERR_ARG, "dose filter needs pixelsize"); }
    if (!(p.voltage_kv > 0)) { zr_stage_free(s); return serr(ZR_ERR_ARG, "Accelerating voltage not set in advance for dose filtering"); }
    ZC(supload(s, &s->twHp, fwd_table(Hp)));
    ZC(supload(s, &s->twWp, fwd_table(Wp)));
    ZC(supload(s, &s->twWp2, fwd_table(Wp / 2)));
    s->spectral = spectral != 0;
    if (s->spectral) {
      const size_t wh = (size_t)N * (Wp / 2 + 1), hh = (size_t)N * Hp;
      ZC(smalloc(s, (void**)&s->R0, sizeof(float2) * wh));
      ZC(smalloc(s, (void**)&s->RH, sizeof(float2) * wh));
      ZC(smalloc(s, (void**)&s->Px, sizeof(float2) * wh));
      ZC(smalloc(s, (void**)&s->C0, sizeof(float2) * hh));
      ZC(smalloc(s, (void**)&s->CW, sizeof(float2) * hh));
      ZC(smalloc(s, (void**)&s->Py, sizeof(float2) * hh));
    } else {
      // a chunk of frames in the row pass's column-pair layout (an odd Wp/2+1 padded by one column)
      const size_t kp = (size_t)(Wp / 2 + 2) / 2;
      ZC(smalloc(s, (void**)&s->Tdose, sizeof(float2) * (size_t)std::min(N, DOSE_CHUNK) * kp * 2 * Hp));
    }
    ZC(smalloc(s, (void**)&s->G, sizeof(float2) * (size_t)(Wp / 2 + 1) * Hp));
    ZC(smalloc(s, (void**)&s->accum, sizeof(double) * (N + 1)));
    ZC(smalloc(s, (void**)&s->thresq, sizeof(double) * N));
    ZC(smalloc(s, (void**)&s->fjc, sizeof(float) * 2 * N));
  }
  for (int i = 0; i < p.n_filters; ++i) {
    if (p.filters[i] == ZR_FILTER_GAUSSLP && !s->gtmp) ZC(smalloc(s, (void**)&s->gtmp, sizeof(float) * H * W));
    if (p.filters[i] == ZR_FILTER_BACKGROUND && !s->bg_nn) {
      const int ny = (H + BG_BIN - 1) / BG_BIN, nx = (W + BG_BIN - 1) / BG_BIN;
      ZC(smalloc(s, (void**)&s->bg_rows, sizeof(float) * (size_t)ny * W));
      ZC(smalloc(s, (void**)&s->bg_nn, sizeof(float) * (size_t)ny * nx));
      ZC(smalloc(s, (void**)&s->bg_fit, sizeof(BgFit)));
    }
  }
  s->fourier = p.shift_method == ZR_SHIFT_FOURIER && p.sub_pix_reg > 1;
  if (s->fourier) {
    ZC(smalloc(s, (void**)&s->aligned, sizeof(float) * (size_t)N * H * W));
    ZC(smalloc(s, (void**)&s->fs_id, sizeof(FrameShift) * N));
    ZC(hipMemset(s->fs_id, 0, sizeof(FrameShift) * N));          // no roll, no kernel: the frames as they are
    ZC(smalloc(s, (void**)&s->Gf, sizeof(float2) * (size_t)std::min(N, DOSE_CHUNK) * (Wp / 2 + 1) * H));
    if (!s->twHp) {
      ZC(supload(s, &s->twHp, fwd_table(Hp)));
      ZC(supload(s, &s->twWp, fwd_table(Wp)));
      ZC(supload(s, &s->twWp2, fwd_table(Wp / 2)));
    }
  }
#undef ZC
  *out = s;
  return ZR_OK;
}

int zr_stage_set_pairs(zr_stage_state* s, const std::vector<int4>& pairs) {
  if (!s) return serr(ZR_ERR_STATE, "no stage state");
  std::vector<int2> ed, ms;
  if (build_equations(s->N, pairs, ed, ms) != 0)
    return serr(ZR_ERR_UNSUPPORTED, "pair pattern gives equations the solver cannot represent");
  if ((int)ed.size() > s->Mcap || ed.empty()) return serr(ZR_ERR_ARG, "equation count %zu outside 1..%d", ed.size(), s->Mcap);
  ZS_HIP(hipMemcpy(s->edges, ed.data(), ed.size() * sizeof(int2), hipMemcpyHostToDevice));
  {
    const std::vector<int2> dg = duplicate_groups(ed);
    ZS_HIP(hipMemcpy(s->dups, dg.data(), dg.size() * sizeof(int2), hipMemcpyHostToDevice));
  }
  ZS_HIP(hipMemcpy(s->meas, ms.data(), ms.size() * sizeof(int2), hipMemcpyHostToDevice));
  s->M = (int)ed.size();
  s->eq_bad = false;
  return ZR_OK;
}

int zr_stage_equation_count(zr_stage_state* s) { return s ? s->M : 0; }

static void fill_post_args(zr_stage_state* s, PostArgs& q) {
  q.pos = s->pos; q.weights = s->weights; q.resid = s->resid; q.meas = s->meas;
  q.N = s->N; q.M = s->M; q.centroid = s->p.origin_centroid; q.lanczos = s->p.shift_method == ZR_SHIFT_LANCZOS;
  q.subpix = s->p.sub_pix_reg; q.H = s->H; q.W = s->W; q.box = s->box; q.overlap = s->p.frc_overlap;
  q.trans_raw = s->trans_raw; q.trans = s->trans; q.rel = s->rel; q.err_stats = s->err_stats;
  q.err_frames = s->err_frames; q.fs = s->fs; q.info = s->info;
}

int zr_stage_equations_ok(const zr_stage_state* s) { return s && !s->eq_bad; }

int zr_stage_solve(zr_stage_state* s, const float* shifts, const float* corr, const float* peaksig, hipStream_t st) {
  if (!s) return serr(ZR_ERR_STATE, "no stage state");
  if (s->eq_bad)
    return serr(ZR_ERR_UNSUPPORTED, "pair pattern gives equations the solver cannot represent (a row of the pair "
                "mask with a gap makes an equation over non-adjacent frame steps)");
  const int N = s->N, M = s->M;
  // a new registration: its sums and filter steps start afresh
  s->have_dose = s->have_sums = s->have_frc = false;
  ZS_HIP(zr_launch_fill_zero(s->info, sizeof(SolveInfo), st));   // a kernel, so the solve can be graph-captured
  if (s->p.tri_mode == ZR_TRI_FIRST || s->p.tri_mode == ZR_TRI_REFINE) {
    hipLaunchKernelGGL(k_direct_positions, dim3(1), dim3(64), 0, st, shifts, N,
                       s->p.tri_mode == ZR_TRI_REFINE ? 1 : 0, s->pos, s->info);
    ZS_HIP(hipGetLastError());
    PostArgs q;
    fill_post_args(s, q);
    hipLaunchKernelGGL(k_solution_post, dim3(1), dim3(POST_NT), 0, st, q);
    ZS_HIP(hipGetLastError());
    s->have_solution = true;
    return ZR_OK;
  }
  const int wm = s->p.weight_mode;
  if (wm == ZR_W_AUTOLOGISTIC) {
    const int maxbins = std::max(1, s->Mcap / 7 + 1);
    hipLaunchKernelGGL(k_ps_hist, dim3(1), dim3(256), 0, st, peaksig, N * N, s->hs, s->cdf, s->pdf, s->info, maxbins);
    ZS_HIP(hipGetLastError());
    const int total = LOGI_G0 * LOGI_G1 * LOGI_G2;
    hipLaunchKernelGGL(k_logi_grid, dim3((total + 255) / 256), dim3(256), 0, st, s->hs, s->cdf, s->info, s->obj);
    ZS_HIP(hipGetLastError());
    static std::once_flag lds_once;
    zr_allow_full_lds((const void*)k_logi_refine, lds_once);
    hipLaunchKernelGGL(k_logi_refine, dim3(1), dim3(LOGI_NT), logi_refine_lds_bytes(), st, s->hs, s->cdf, s->obj,
                       s->info);
    ZS_HIP(hipGetLastError());
  }
  SolveArgs a;
  a.shifts = shifts; a.corr = corr; a.peaksig = peaksig; a.edges = s->edges; a.meas = s->meas; a.dups = s->dups;
  a.N = N; a.M = M; a.weight_mode = wm;
  a.logi_fixed[0] = s->p.peaksig_thres; a.logi_fixed[1] = s->p.logistic_k; a.logi_fixed[2] = s->p.logistic_nu;
  a.bound = 0;
  a.weights = s->weights; a.resid = s->resid; a.pos = s->pos; a.info = s->info;
  if (M > L1_NT * L1_EPT) return serr(ZR_ERR_UNSUPPORTED, "the solver handles up to %d equations (got %d)", L1_NT * L1_EPT, M);
  const size_t dyn = l1_dyn_bytes(M);
  static std::once_flag lds_once;
  zr_allow_full_lds((const void*)k_l1_solve, lds_once);
  hipLaunchKernelGGL(k_l1_solve, dim3(2), dim3(L1_NT), dyn, st, a);
  ZS_HIP(hipGetLastError());
  PostArgs q;
  fill_post_args(s, q);
  hipLaunchKernelGGL(k_solution_post, dim3(1), dim3(POST_NT), 0, st, q);
  ZS_HIP(hipGetLastError());
  s->have_solution = true;
  return ZR_OK;
}

// Shift + sums over the frames; returns the number of blocks (the partial sums in rowsum).
static int launch_shift_sum(const float* frames, const FrameShift* fs, int N, int H, int W, float* img_sum,
                            float* even, float* odd, float* aligned, double* rowsum, hipStream_t st) {
  const int gx = (W + 64 * SS_WX - 1) / (64 * SS_WX), gy = (H + (4 / SS_WX) * SS_R - 1) / ((4 / SS_WX) * SS_R);
  hipLaunchKernelGGL((k_shift_sum<SS_R, SS_WX>), dim3(gx * gy), dim3(256), 0, st, frames, fs, N, H, W, img_sum, even,
                     odd, aligned, rowsum, gx);
  return gx * gy;
}

int zr_stage_shifts(zr_stage_state* s, const float* frames, float* aligned_out, hipStream_t st) {
  if (!s || !s->have_solution) return serr(ZR_ERR_STATE, "shifts before solve");
  const FrameShift* fs = s->fs;
  if (s->fourier) {            // the frames Fourier-shifted by zr_stage_fourier_align, summed as they are
    frames = s->aligned;
    fs = s->fs_id;
  }
  launch_shift_sum(frames, fs, s->N, s->H, s->W, s->img_sum, s->even, s->odd, aligned_out, s->rowsum, st);
  ZS_HIP(hipGetLastError());
  s->have_sums = true;
  return ZR_OK;
}

// tiledFRC(..., trans=...) with caller translations (zorro/zorro.py:2264-2274):
// the crop limits and tile counts that k_solution_post derives from the
// movie's own translations, recomputed on the host from `trans` [n][2].
int zr_stage_set_frc_crop(zr_stage_state* s, const double* trans, int n, hipStream_t st) {
  if (!s || !trans || n < 1) return serr(ZR_ERR_ARG, "FRC translations");
  double mny = INFINITY, mxy = -INFINITY, mnx = INFINITY, mxx = -INFINITY;
  bool any = false;
  for (int k = 0; k < n; ++k) {
    const double ty = trans[2 * k], tx = trans[2 * k + 1];
    mny = std::min(mny, ty); mxy = std::max(mxy, ty); mnx = std::min(mnx, tx); mxx = std::max(mxx, tx);
    if (ty != 0.0 || tx != 0.0) any = true;
  }
  int v[6];   // tiles_y, tiles_x, crop[4] (contiguous in SolveInfo)
  int* c = v + 2;
  if (!any) { c[0] = c[1] = c[2] = c[3] = 5; }
  else {
    c[2] = (int)(-std::min(std::floor(mny), 5.0));
    c[3] = (int)(-std::min(std::floor(mnx), 5.0));
    c[0] = (int)std::max(std::ceil(mxy), 5.0);
    c[1] = (int)std::max(std::ceil(mxx), 5.0);
  }
  const double ov = s->p.frc_overlap > 0.8 ? 0.8 : s->p.frc_overlap;
  v[1] = std::max(0, (int)std::floor((double)(s->W - c[1] - c[3] - s->box) / s->box / (1.0 - ov)));
  v[0] = std::max(0, (int)std::floor((double)(s->H - c[0] - c[2] - s->box) / s->box / (1.0 - ov)));
  static_assert(offsetof(SolveInfo, crop) == offsetof(SolveInfo, tiles_y) + 2 * sizeof(int), "SolveInfo layout");
  ZS_HIP(hipMemcpyAsync((char*)s->info + offsetof(SolveInfo, tiles_y), v, sizeof(v), hipMemcpyHostToDevice, st));
  ZS_HIP(hipStreamSynchronize(st));
  return ZR_OK;
}

int zr_stage_frc(zr_stage_state* s, hipStream_t st) {
  if (!s || !s->have_sums) return serr(ZR_ERR_STATE, "FRC before the half sums");
  if (!s->p.do_frc) return serr(ZR_ERR_STATE, "FRC disabled in the parameters");
  const int nt = s->tmax_y * s->tmax_x;
  if (nt < 1) return serr(ZR_ERR_ARG, "frame smaller than one FRC tile");
  const double step = 1.0 - (s->p.frc_overlap > 0.8 ? 0.8 : s->p.frc_overlap);
  const int B = s->box;
  switch (B) {
#define ZR_FRC(BX)                                                                                            \
  case BX:                                                                                                    \
    hipLaunchKernelGGL(k_frc_rows<BX>, dim3(nt * 2 * (BX / frc_rows_per_block(BX))), dim3(256), 0, st, s->even, s->odd, s->info, s->W, \
                       s->tmax_x, step, s->twB2, s->twB, s->hann, 0, s->Tfrc);                                \
    ZS_HIP(hipGetLastError());                                                                                \
    hipLaunchKernelGGL(k_frc_cols<BX>, dim3((BX / 2 + 1) * FRC_GROUPS), dim3(frc_cols_nt(BX)), 0, st, s->Tfrc, s->info, \
                       s->tmax_x, 0, nt, 0, s->twB, s->frc_part);                                            \
    ZS_HIP(hipGetLastError());                                                                                \
    hipLaunchKernelGGL(k_frc_reduce, dim3(256), dim3(256), 0, st, s->frc_part, (size_t)BX * (BX / 2 + 1),      \
                       s->frc2d);                                                                             \
    ZS_HIP(hipGetLastError());                                                                                \
    break;
    ZR_FRC(256) ZR_FRC(512) ZR_FRC(1024)
#undef ZR_FRC
    default: return serr(ZR_ERR_UNSUPPORTED, "FRC box");
  }
  hipLaunchKernelGGL(k_frc_rings, dim3(B / 2), dim3(256), 0, st, s->frc2d, B, s->frc);
  ZS_HIP(hipGetLastError());
  s->have_frc = true;
  return ZR_OK;
}

static void fill_dose_args(zr_stage_state* s, DoseArgs& a, const double* rowsum, int nrows, int normalize) {
  a.N = s->N; a.H = s->H; a.W = s->W; a.Hp = s->Hp; a.Wp = s->Wp;
  a.A = s->p.dose_a; a.B = s->p.dose_b; a.C = s->p.dose_c; a.order = s->p.dose_order; a.missing = s->p.dose_missing;
  a.vs = std::sqrt(s->p.voltage_kv / 300.0);
  a.ipx = 1.0 / (s->Wp * (s->p.pixelsize_nm * 10.0));
  a.ipy = 1.0 / (s->Hp * (s->p.pixelsize_nm * 10.0));
  a.rowsum = rowsum; a.nrows = nrows; a.accum = s->accum; a.thresq = s->thresq; a.fjc = s->fjc; a.info = s->info;
  a.order_sq = -1;
  a.normalize = normalize;
  a.cut_skip = a.order > 0.0 ? (float)std::exp(std::log(200.0) / a.order) : INFINITY;
  for (int k = 0; k < 8; ++k)
    if (a.order == (double)(1 << k)) a.order_sq = k;
}

static int dose_spectral(zr_stage_state* s, const float* frames, const float2* S, int Wk, int normalize,
                         hipStream_t st) {
  const int H = s->H, W = s->W;
  // aligned sums (imageSum, even, odd) and their total for the dose per frame
  const int nblk = launch_shift_sum(frames, s->fs, s->N, H, W, s->img_sum, s->even, s->odd, nullptr, s->rowsum, st);
  ZS_HIP(hipGetLastError());
  s->have_sums = true;
  DoseArgs a;
  fill_dose_args(s, a, s->rowsum, nblk, normalize);
  hipLaunchKernelGGL(k_dose_params, dim3(1), dim3(256), (size_t)(a.Hp / 2) * sizeof(double), st, a);
  ZS_HIP(hipGetLastError());
  if (H != W) return serr(ZR_ERR_UNSUPPORTED, "spectral dose path needs square frames");
  switch (H) {
#define ZR_DB(X) case X: hipLaunchKernelGGL(k_dose_border<X>, dim3(s->N * 4), dim3(256), 0, st, frames, s->fs, H, W, s->twHp, s->twHp, s->twWp, s->R0, s->RH, s->C0, s->CW, s->Py, s->Px); break;
    ZR_DB(256) ZR_DB(512) ZR_DB(1024) ZR_DB(2048) ZR_DB(3072) ZR_DB(3840) ZR_DB(4096)
#undef ZR_DB
    default: return serr(ZR_ERR_UNSUPPORTED, "dose size %d", H);
  }
  ZS_HIP(hipGetLastError());
  if (s->Hp != H || s->Wp != W) return serr(ZR_ERR_UNSUPPORTED, "spectral dose path needs unpadded frames");
  {
    const int whalf = s->Wp / 2 + 1;
    const int grid_acc = (s->Hp / DA_NT) * ((whalf + DA_U - 1) / DA_U);
    switch (s->Hp) {
#define ZR_DS(X) case X: \
      hipLaunchKernelGGL(k_dose_accum<X>, dim3(grid_acc), dim3(DA_NT), 0, st, S, Wk, a, s->R0, s->RH, s->C0, s->CW, s->Py, s->Px, s->twHp, s->twWp, s->G); \
      ZS_HIP(hipGetLastError()); \
      hipLaunchKernelGGL(k_dose_cols_ifft<X>, dim3(whalf), dim3(256), 0, st, s->twHp, s->G); break;
      ZR_DS(256) ZR_DS(512) ZR_DS(1024) ZR_DS(2048) ZR_DS(3072) ZR_DS(3840) ZR_DS(4096)
#undef ZR_DS
      default: return serr(ZR_ERR_UNSUPPORTED, "dose height %d", s->Hp);
    }
  }
  ZS_HIP(hipGetLastError());
  switch (s->Wp) {
#define ZR_DO(X) case X: hipLaunchKernelGGL(k_dose_out<X>, dim3(s->H), dim3(256), 0, st, s->G, s->Hp, s->H, s->W, s->twWp2, s->twWp, s->filt); break;
    ZR_DO(256) ZR_DO(512) ZR_DO(1024) ZR_DO(2048) ZR_DO(3072) ZR_DO(3840) ZR_DO(4096)
#undef ZR_DO
    default: return serr(ZR_ERR_UNSUPPORTED, "dose width %d", s->Wp);
  }
  ZS_HIP(hipGetLastError());
  s->have_dose = true;
  return ZR_OK;
}

int zr_stage_dose(zr_stage_state* s, const float* frames, const float2* S, int Wk, int normalize, hipStream_t st) {
  if (!s || !s->have_solution) return serr(ZR_ERR_STATE, "dose filter before solve");
  if (!s->p.do_dose) return serr(ZR_ERR_STATE, "dose filter disabled in the parameters");
  if (s->spectral) return dose_spectral(s, frames, S, Wk, normalize, st);
  const int H = s->H, W = s->W, Hp = s->Hp, Wp = s->Wp;
  const FrameShift* fs = s->fs;
  if (s->fourier) {            // the Fourier-shifted frames, taken as they are
    frames = s->aligned;
    fs = s->fs_id;
  }
  // aligned sums (imageSum, even, odd) and their total for the dose per frame
  const int nblk = launch_shift_sum(frames, fs, s->N, H, W, s->img_sum, s->even, s->odd, nullptr, s->rowsum, st);
  ZS_HIP(hipGetLastError());
  s->have_sums = true;
  DoseArgs a;
  fill_dose_args(s, a, s->rowsum, nblk, normalize);
  hipLaunchKernelGGL(k_dose_params, dim3(1), dim3(256), (size_t)(a.Hp / 2) * sizeof(double), st, a);
  ZS_HIP(hipGetLastError());
  // the filtered spectra of the aligned, symmetric-padded frames, DOSE_CHUNK frames at a time
  for (int j0 = 0; j0 < s->N; j0 += DOSE_CHUNK) {
    const int j1 = std::min(s->N, j0 + DOSE_CHUNK);
    switch (Wp) {
#define ZR_DR(X)                                                                                                  \
  case X: {                                                                                                       \
    constexpr int RB = dose_rows_rb(X), NT = RB * dose_rows_tpr(X);                                               \
    const size_t sh = (size_t)RB * zr_fft_lds(X / 2) * sizeof(float2);                                           \
    static std::once_flag lds_once;                                                                               \
    zr_allow_full_lds((const void*)k_dose_rows<X>, lds_once);                                                     \
    hipLaunchKernelGGL(k_dose_rows<X>, dim3((j1 - j0) * (Hp / RB)), dim3(NT), sh, st, frames, fs, j0, H, W, Hp, \
                       s->twWp2, s->twWp, a.thresq, a.ipx, a.cut_skip, s->Tdose);                                  \
  } break;
      ZR_DR(256) ZR_DR(512) ZR_DR(1024) ZR_DR(2048) ZR_DR(3072) ZR_DR(3840) ZR_DR(4096) ZR_DR(5760)
#undef ZR_DR
      default: return serr(ZR_ERR_UNSUPPORTED, "dose width %d", Wp);
    }
    ZS_HIP(hipGetLastError());
    switch (Hp) {
#define ZR_DC(X) case X: hipLaunchKernelGGL(k_dose_cols<X>, dim3(Wp / 2 + 1), dim3(dose_cols_nt(X)), 0, st, s->Tdose, a, j0, j1, s->twHp, s->G); break;
      ZR_DC(256) ZR_DC(512) ZR_DC(1024) ZR_DC(2048) ZR_DC(3072) ZR_DC(3840) ZR_DC(4096) ZR_DC(5760)
#undef ZR_DC
      default: return serr(ZR_ERR_UNSUPPORTED, "dose height %d", Hp);
    }
    ZS_HIP(hipGetLastError());
  }
  switch (Wp) {
#define ZR_DO(X) case X: hipLaunchKernelGGL(k_dose_out<X>, dim3(s->H), dim3(256), 0, st, s->G, s->Hp, s->H, s->W, s->twWp2, s->twWp, s->filt); break;
    ZR_DO(256) ZR_DO(512) ZR_DO(1024) ZR_DO(2048) ZR_DO(3072) ZR_DO(3840) ZR_DO(4096) ZR_DO(5760)
#undef ZR_DO
    default: return serr(ZR_ERR_UNSUPPORTED, "dose width %d", Wp);
  }
  ZS_HIP(hipGetLastError());
  s->have_dose = true;
  return ZR_OK;
}

int zr_stage_fourier_align(zr_stage_state* s, const float2* S, int Wk, int Hs, int mask_slot, const double* mu,
                           hipStream_t st) {
  if (!s || !s->have_solution) return serr(ZR_ERR_STATE, "Fourier shifts before solve");
  if (!s->fourier) return serr(ZR_ERR_STATE, "shiftMethod is not 'fourier'");
  const int N = s->N, H = s->H, W = s->W, Hp = s->Hp, Wp = s->Wp;
  if (Wk != Wp / 2 + 1 || Hs != Hp) return serr(ZR_ERR_STATE, "Fourier shifts need full-height half spectra");
  const float scale = (float)(1.0 / ((double)Hp * (double)Wp));
  for (int j0 = 0; j0 < N; j0 += DOSE_CHUNK) {
    const int nb = std::min(N, j0 + DOSE_CHUNK) - j0;
    switch (Hp) {
#define ZR_FC(X) case X: hipLaunchKernelGGL(k_fshift_cols<X>, dim3(nb * Wk), dim3(dose_cols_nt(X)), 0, st, S, Wk, Hs, mask_slot, mu, (double)H * W, s->trans, Wp, H, j0, s->twHp, s->Gf); break;
      ZR_FC(256) ZR_FC(512) ZR_FC(1024) ZR_FC(2048) ZR_FC(3072) ZR_FC(3840) ZR_FC(4096) ZR_FC(5760)
#undef ZR_FC
      default: return serr(ZR_ERR_UNSUPPORTED, "Fourier shift height %d", Hp);
    }
    ZS_HIP(hipGetLastError());
    switch (Wp) {
#define ZR_FR(X) case X: hipLaunchKernelGGL(k_fshift_rows<X>, dim3(nb * H), dim3(256), 0, st, s->Gf, Wk, H, W, j0, scale, s->twWp2, s->twWp, s->aligned); break;
      ZR_FR(256) ZR_FR(512) ZR_FR(1024) ZR_FR(2048) ZR_FR(3072) ZR_FR(3840) ZR_FR(4096) ZR_FR(5760)
#undef ZR_FR
      default: return serr(ZR_ERR_UNSUPPORTED, "Fourier shift width %d", Wp);
    }
    ZS_HIP(hipGetLastError());
  }
  return ZR_OK;
}

int zr_stage_gausslp(zr_stage_state* s, hipStream_t st) {
  if (!s || !s->have_sums) return serr(ZR_ERR_STATE, "gausslp before the sums");
  if (!s->gtmp) return serr(ZR_ERR_STATE, "gausslp is not among the parameters' filter steps");
  GaussTaps t;
  {
    double sum = 0.0, w[2 * GLP_R + 1];
    for (int j = -GLP_R; j <= GLP_R; ++j) sum += (w[j + GLP_R] = std::exp(-0.5 / 9.0 * (double)(j * j)));
    for (int j = 0; j <= GLP_R; ++j) t.w[j] = w[j + GLP_R] / sum;
  }
  // no filtSum yet: the filter starts from a copy of imageSum (zorro/zorro.py:1708-1710)
  const float* src = s->have_dose ? s->filt : s->img_sum;
  const size_t n = (size_t)s->H * s->W;
  hipLaunchKernelGGL(k_gauss_axis0, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, s->H, s->W, t, s->gtmp);
  ZS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_gauss_axis1, dim3((unsigned)(s->H * ((s->W + 255) / 256))), dim3(256), 0, st, s->gtmp, s->H,
                     s->W, t, s->filt);
  ZS_HIP(hipGetLastError());
  s->have_dose = true;
  return ZR_OK;
}

int zr_stage_background(zr_stage_state* s, hipStream_t st) {
  if (!s || !s->have_sums) return serr(ZR_ERR_STATE, "background before the sums");
  if (!s->bg_nn) return serr(ZR_ERR_STATE, "background is not among the parameters' filter steps");
  const int H = s->H, W = s->W;
  const int ny = (H + BG_BIN - 1) / BG_BIN, nx = (W + BG_BIN - 1) / BG_BIN;
  BgTaps t;
  {
    double sum = 0.0, w[2 * BG_R + 1];
    for (int j = -BG_R; j <= BG_R; ++j) sum += (w[j + BG_R] = std::exp(-0.5 / 16.0 * (double)(j * j)));
    for (int j = 0; j <= BG_R; ++j) t.w[j] = w[j + BG_R] / sum;
  }
  // no filtSum yet: the estimate starts from a copy of imageSum (zorro/zorro.py:1703-1704)
  if (!s->have_dose) ZS_HIP(hipMemcpyAsync(s->filt, s->img_sum, sizeof(float) * H * W, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_bg_rows, dim3((unsigned)(((size_t)ny * W + 255) / 256)), dim3(256), 0, st, s->filt, H, W, ny, t,
                     s->bg_rows);
  ZS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_bg_sample, dim3((ny * nx + 255) / 256), dim3(256), 0, st, s->bg_rows, W, ny, nx, t, s->bg_nn);
  ZS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_bg_fit, dim3(1), dim3(BG_NT), 0, st, s->bg_nn, ny, nx, s->bg_fit);
  ZS_HIP(hipGetLastError());
  const size_t n = (size_t)H * W;
  hipLaunchKernelGGL(k_bg_subtract, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s->filt, H, W, s->bg_fit);
  ZS_HIP(hipGetLastError());
  s->have_dose = true;
  return ZR_OK;
}

int zr_stage_get_background(zr_stage_state* s, double* p6, double* fval, int* status) {
  if (!s || !s->bg_fit) return serr(ZR_ERR_STATE, "background is not among the parameters' filter steps");
  BgFit f;
  ZS_HIP(hipMemcpy(&f, s->bg_fit, sizeof(f), hipMemcpyDeviceToHost));
  if (p6) for (int i = 0; i < 6; ++i) p6[i] = f.p[i];
  if (fval) *fval = f.f;
  if (status) *status = f.status;
  return ZR_OK;
}

int zr_stage_get_hist(zr_stage_state* s, float* pdf, float* cdf, float* hsigma, int cap, int* nbins) {
  if (!s || !s->have_solution) return serr(ZR_ERR_STATE, "no solution");
  SolveInfo inf;
  ZS_HIP(hipMemcpy(&inf, s->info, sizeof(inf), hipMemcpyDeviceToHost));
  const bool hist = s->p.weight_mode == ZR_W_AUTOLOGISTIC &&
                    !(s->p.tri_mode == ZR_TRI_FIRST || s->p.tri_mode == ZR_TRI_REFINE);
  const int nb = hist ? inf.nbins : 0;
  if (nbins) *nbins = nb;
  if (nb > cap && (pdf || cdf || hsigma)) return serr(ZR_ERR_ARG, "histogram has %d bins, buffers hold %d", nb, cap);
  if (nb > 0) {
    if (pdf) ZS_HIP(hipMemcpy(pdf, s->pdf, sizeof(float) * nb, hipMemcpyDeviceToHost));
    if (cdf) ZS_HIP(hipMemcpy(cdf, s->cdf, sizeof(float) * nb, hipMemcpyDeviceToHost));
    if (hsigma) ZS_HIP(hipMemcpy(hsigma, s->hs, sizeof(float) * nb, hipMemcpyDeviceToHost));
  }
  return ZR_OK;
}

// ---- standalone tiledFRC of two images (zr_tiled_frc, zr_tiled_frc_noalloc)
static size_t frc_al(size_t b) { return (b + 255) & ~(size_t)255; }

size_t zr_frc_work_bytes(int box, int chunk) {
  const size_t wh = box / 2 + 1;
  return frc_al(sizeof(SolveInfo)) + frc_al(sizeof(float2) * box) + frc_al(sizeof(float2) * (box / 2)) +
         frc_al(sizeof(float) * box * box) + frc_al(sizeof(float2) * (size_t)chunk * 2 * wh * box) +
         frc_al(sizeof(double) * FRC_GROUPS * box * wh) + frc_al(sizeof(double) * box * wh);
}

int zr_frc_rings_count(int box) {
  // rotmean's length: ceil(sqrt((box/2)^2 + (box/2)^2)) + 1 (zorro/zorro_util.py:174)
  const double h = box / 2;
  return (int)std::ceil(std::sqrt(h * h + h * h)) + 1;
}

int zr_frc_tiles_from_trans(const double* trans, int n, int h, int w, int box, double overlap, int crop[4],
                            int* tiles_y, int* tiles_x) {
  // tiledFRC's crop limits (minCrop 5) and tile counts (zorro/zorro.py:2255-2281)
  bool any = false;
  double mny = INFINITY, mxy = -INFINITY, mnx = INFINITY, mxx = -INFINITY;
  for (int k = 0; trans && k < n; ++k) {
    const double ty = trans[2 * k], tx = trans[2 * k + 1];
    mny = std::min(mny, ty); mxy = std::max(mxy, ty); mnx = std::min(mnx, tx); mxx = std::max(mxx, tx);
    if (ty != 0.0 || tx != 0.0) any = true;
  }
  if (!any) { crop[0] = crop[1] = crop[2] = crop[3] = 5; }
  else {
    crop[2] = (int)(-std::min(std::floor(mny), 5.0));
    crop[3] = (int)(-std::min(std::floor(mnx), 5.0));
    crop[0] = (int)std::max(std::ceil(mxy), 5.0);
    crop[1] = (int)std::max(std::ceil(mxx), 5.0);
  }
  const double ov = overlap > 0.8 ? 0.8 : overlap;
  *tiles_x = (int)std::floor((double)(w - crop[1] - crop[3] - box) / box / (1.0 - ov));
  *tiles_y = (int)std::floor((double)(h - crop[0] - crop[2] - box) / box / (1.0 - ov));
  return (*tiles_x > 0 && *tiles_y > 0) ? ZR_OK : ZR_ERR_ARG;
}

int zr_frc_tiles_run(const float* img1, const float* img2, int h, int w, const int crop[4], int tiles_y, int tiles_x,
                     int box, double overlap, void* work, int chunk, double* frc, hipStream_t st) {
  (void)h;
  const size_t wh = box / 2 + 1;
  char* p = static_cast<char*>(work);
  auto take = [&](size_t b) { char* q = p; p += frc_al(b); return q; };
  SolveInfo* info = reinterpret_cast<SolveInfo*>(take(sizeof(SolveInfo)));
  float2* twB = reinterpret_cast<float2*>(take(sizeof(float2) * box));
  float2* twB2 = reinterpret_cast<float2*>(take(sizeof(float2) * (box / 2)));
  float* hann = reinterpret_cast<float*>(take(sizeof(float) * box * box));
  float2* T = reinterpret_cast<float2*>(take(sizeof(float2) * (size_t)chunk * 2 * wh * box));
  double* part = reinterpret_cast<double*>(take(sizeof(double) * FRC_GROUPS * box * wh));
  double* frc2d = reinterpret_cast<double*>(take(sizeof(double) * box * wh));
  const double step = 1.0 - (overlap > 0.8 ? 0.8 : overlap);
  const int nt = tiles_y * tiles_x;
  hipLaunchKernelGGL(k_frc_setup, dim3((box + 255) / 256), dim3(256), 0, st, box, twB, twB2, info, tiles_y, tiles_x,
                     crop[0], crop[1], crop[2], crop[3]);
  ZS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_frc_hann, dim3((box * box + 255) / 256), dim3(256), 0, st, box, hann);
  ZS_HIP(hipGetLastError());
  for (int t0 = 0; t0 < nt; t0 += chunk) {
    const int t1 = std::min(nt, t0 + chunk);
    switch (box) {
#define ZR_FRCT(BX)                                                                                             \
  case BX:                                                                                                      \
    hipLaunchKernelGGL(k_frc_rows<BX>, dim3((t1 - t0) * 2 * (BX / frc_rows_per_block(BX))), dim3(256), 0, st, img1,  \
                       img2, info, w, tiles_x, step, twB2, twB, hann, t0, T);                                  \
    ZS_HIP(hipGetLastError());                                                                                  \
    hipLaunchKernelGGL(k_frc_cols<BX>, dim3((BX / 2 + 1) * FRC_GROUPS), dim3(frc_cols_nt(BX)), 0, st, T, info,   \
                       tiles_x, t0, t1, t0 > 0 ? 1 : 0, twB, part);                                            \
    ZS_HIP(hipGetLastError());                                                                                  \
    break;
      ZR_FRCT(256) ZR_FRCT(512) ZR_FRCT(1024)
#undef ZR_FRCT
      default: return serr(ZR_ERR_UNSUPPORTED, "FRC box %d: built sizes 256, 512, 1024", box);
    }
  }
  hipLaunchKernelGGL(k_frc_reduce, dim3(256), dim3(256), 0, st, part, (size_t)box * wh, frc2d);
  ZS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_frc_rings, dim3(zr_frc_rings_count(box)), dim3(256), 0, st, frc2d, box, frc);
  ZS_HIP(hipGetLastError());
  return ZR_OK;
}

int zr_stage_get_solution(zr_stage_state* s, double* translations, double* relative_est, double* weights,
                          double* logistic, double* error, double* error_stats, double* error_xy_frames) {
  if (!s || !s->have_solution) return serr(ZR_ERR_STATE, "no solution");
  SolveInfo inf;
  ZS_HIP(hipMemcpy(&inf, s->info, sizeof(inf), hipMemcpyDeviceToHost));
  if (inf.status_x < 0 || inf.status_y < 0) return serr(ZR_ERR_ARG, "equation graph is disconnected");
  if (inf.status_x != 1 || inf.status_y != 1)
    return serr(ZR_ERR_STATE, "shift solver stopped before proving optimality (status y %d, x %d)", inf.status_y,
                inf.status_x);
  const bool solver = !(s->p.tri_mode == ZR_TRI_FIRST || s->p.tri_mode == ZR_TRI_REFINE);
  if (solver && s->p.weight_mode == ZR_W_AUTOLOGISTIC && inf.nbins < 1)
    return serr(ZR_ERR_ARG, "autologistic weights need >= 7 positive peak significances (got %d)", inf.nps);
  const int N = s->N, M = s->M;
  if (translations) ZS_HIP(hipMemcpy(translations, s->trans, sizeof(double) * 2 * N, hipMemcpyDeviceToHost));
  if (relative_est) ZS_HIP(hipMemcpy(relative_est, s->rel, sizeof(double) * 2 * (N - 1), hipMemcpyDeviceToHost));
  if (weights) ZS_HIP(hipMemcpy(weights, s->weights, sizeof(double) * M, hipMemcpyDeviceToHost));
  if (logistic) {
    if (solver && s->p.weight_mode == ZR_W_AUTOLOGISTIC) for (int i = 0; i < 3; ++i) logistic[i] = inf.logi[i];
    else { logistic[0] = s->p.peaksig_thres; logistic[1] = s->p.logistic_k; logistic[2] = s->p.logistic_nu; }
  }
  if (error) ZS_HIP(hipMemcpy(error, s->resid, sizeof(double) * 2 * M, hipMemcpyDeviceToHost));
  if (error_stats) ZS_HIP(hipMemcpy(error_stats, s->err_stats, sizeof(double) * 2, hipMemcpyDeviceToHost));
  if (error_xy_frames) ZS_HIP(hipMemcpy(error_xy_frames, s->err_frames, sizeof(double) * 2 * N, hipMemcpyDeviceToHost));
  return ZR_OK;
}

int zr_stage_solver_info(zr_stage_state* s, int* out6) {
  if (!s || !s->have_solution) return serr(ZR_ERR_STATE, "no solution");
  SolveInfo inf;
  ZS_HIP(hipMemcpy(&inf, s->info, sizeof(inf), hipMemcpyDeviceToHost));
  const int v[6] = {inf.nps, inf.nbins, inf.status_y, inf.status_x, inf.pivots_y, inf.pivots_x};
  for (int i = 0; i < 6; ++i) out6[i] = v[i];
  return ZR_OK;
}

int zr_stage_get_sums(zr_stage_state* s, float* image_sum, float* filt_sum, float* even_sum, float* odd_sum) {
  if (!s || !s->have_sums) return serr(ZR_ERR_STATE, "no sums");
  size_t b = sizeof(float) * s->H * s->W;
  if (image_sum) ZS_HIP(hipMemcpy(image_sum, s->img_sum, b, hipMemcpyDeviceToHost));
  if (even_sum) ZS_HIP(hipMemcpy(even_sum, s->even, b, hipMemcpyDeviceToHost));
  if (odd_sum) ZS_HIP(hipMemcpy(odd_sum, s->odd, b, hipMemcpyDeviceToHost));
  if (filt_sum) {
    if (!s->have_dose) return serr(ZR_ERR_STATE, "no dose-filtered sum");
    ZS_HIP(hipMemcpy(filt_sum, s->filt, b, hipMemcpyDeviceToHost));
  }
  return ZR_OK;
}

int zr_stage_get_frc(zr_stage_state* s, double* frc, int* n) {
  if (!s || !s->have_frc) return serr(ZR_ERR_STATE, "no FRC");
  if (n) *n = s->box / 2;
  if (frc) ZS_HIP(hipMemcpy(frc, s->frc, sizeof(double) * (s->box / 2), hipMemcpyDeviceToHost));
  return ZR_OK;
}

int zr_stage_device_results(zr_stage_state* s, const float** image_sum, const float** filt_sum,
                            const double** translations) {
  if (!s) return serr(ZR_ERR_STATE, "no stage state");
  if (image_sum) *image_sum = s->img_sum;
  if (filt_sum) *filt_sum = s->filt;
  if (translations) *translations = s->trans;
  return ZR_OK;
}
