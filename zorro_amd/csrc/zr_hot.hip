// Outlier-pixel filter ('hot' in filterMode): hotpixFilter of the reference
// (zorro/zorro.py:2497-2662) on the frame stack in HBM, before the spectra.
//
//   k_hot_sum     unaligned sum over frames (float32, frame order) + FP64 partials
//   k_hot_setup   mean, Poisson std, histogram range (unit bins, half-integer edges)
//   k_hot_hist    histogram of the sum (LDS-privatised counts, integer atomics)
//   k_hot_fit     float32 CDF, bounded fit of a normal CDF (mu, sigma), sigma
//                 curve from erfinv and the two least-squares cut-offs (one block)
//   k_hot_mask    per-pixel logistic weight and the dead/hot counts
//   k_hot_nbr     3x3 PSF convolutions (reflect): neighbour mask, singleton term
//   k_hot_nfilt   clustered outliers: mean of the non-outlier 3x3 neighbours
//   k_hot_apply   frames = (w * frame + singleton term) + cluster term, in place
//
// Every pass is HBM-bound integer/elementwise work; the stack is read twice
// (sum, apply) and written once.  Arithmetic follows the reference's types:
// float32 per pixel, the logistic weights as numexpr evaluates them (float32
// exponent, float64 combination), convolutions accumulated in float64 and
// rounded to float32 as scipy.ndimage.convolve does.
#include "zr_common.h"
#include "zr_hot.h"

#include <cmath>

namespace {

constexpr int HOT_NT = 256;
constexpr int HOT_LDS_BINS = 8192;     // bins counted in LDS; more go straight to global atomics

ZR_DEV double norm_cdf(double x) { return 0.5 * erfc(-x * 0.70710678118654752440); }

// ------------------------------------------------------------------ passes
__global__ __launch_bounds__(HOT_NT) void k_hot_sum(const float* __restrict__ frames, int N, size_t hw,
                                                    float* __restrict__ usum, double* __restrict__ part) {
  __shared__ double red[HOT_NT / 64];
  const size_t p = (size_t)blockIdx.x * HOT_NT + threadIdx.x;
  double d = 0.0;
  if (p < hw) {
    float acc = frames[p];
    for (int J = 1; J < N; ++J) acc = __fadd_rn(acc, frames[(size_t)J * hw + p]);
    usum[p] = acc;
    d = acc;
  }
  d = wave_sum_d(d);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < HOT_NT / 64; ++i) t += red[i];
    part[blockIdx.x] = t;
  }
}

// One block: fixed-order reduction of the partial sums, the histogram range of
// np.arange(floor(mean - s*std) - 0.5, ceil(mean + s*std) + 0.5, 1), and the
// cleared counters.
__global__ __launch_bounds__(HOT_NT) void k_hot_setup(const double* __restrict__ part, int nblk, size_t hw,
                                                      double max_sigma, int maxbins, HotState* __restrict__ hs,
                                                      unsigned* __restrict__ hist) {
  __shared__ double red[HOT_NT];
  double t = 0.0;
  for (int i = threadIdx.x; i < nblk; i += HOT_NT) t += part[i];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int o = HOT_NT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double mean = red[0] / (double)hw;
  const double pstd = sqrt(mean);
  const double x1 = floor(mean - max_sigma * pstd), x2 = ceil(mean + max_sigma * pstd);
  const long nedges = (long)(x2 - x1) + 1;
  const int nb = (int)(nedges - 1);
  for (int i = threadIdx.x; i < maxbins; i += HOT_NT) hist[i] = 0u;
  if (threadIdx.x == 0) {
    hs->sum_mean = mean;
    hs->pstd = pstd;
    hs->lo = x1 - 0.5;
    hs->status = (nb >= 2 && nb <= maxbins && mean > 0.0) ? 0 : 1;
    hs->nbins = hs->status ? 0 : nb;
    hs->n_dead = hs->n_hot = hs->n_nbr = 0ull;
  }
}

// numpy.histogram on uniform unit bins [lo + k, lo + k + 1), the last bin
// closed; values outside [lo, lo + nbins] are not counted.
__global__ __launch_bounds__(HOT_NT) void k_hot_hist(const float* __restrict__ usum, size_t hw,
                                                     const HotState* __restrict__ hs, unsigned* __restrict__ hist) {
  __shared__ unsigned sh[HOT_LDS_BINS];
  const int nb = hs->nbins;
  if (nb <= 0) return;                                     // block-uniform
  const double lo = hs->lo;
  const bool local = nb <= HOT_LDS_BINS;
  if (local)
    for (int i = threadIdx.x; i < nb; i += HOT_NT) sh[i] = 0u;
  __syncthreads();
  for (size_t p = (size_t)blockIdx.x * HOT_NT + threadIdx.x; p < hw; p += (size_t)gridDim.x * HOT_NT) {
    const double v = (double)usum[p] - lo;                 // exact: float32 minus a half-integer
    if (!(v >= 0.0) || v > (double)nb) continue;
    int k = (int)floor(v);
    if (k == nb) k = nb - 1;
    if (local) atomicAdd(&sh[k], 1u); else atomicAdd(&hist[k], 1u);
  }
  if (local) {
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += HOT_NT)
      if (sh[i]) atomicAdd(&hist[i], sh[i]);
  }
}

// sum_k |cdf_k - Phi((c_k - mu)/sigma)| by one wave, bins in fixed lane order.
ZR_DEV double norm_objective(const float* cdf, const float* cnt, int nb, double mu, double sg, int lane) {
  double s = 0.0;
  for (int k = lane; k < nb; k += 64) s += fabs((double)cdf[k] - norm_cdf(((double)cnt[k] - mu) / sg));
  return wave_sum_d(s);
}

// Least-squares line y = a x + b through n points (numpy.linalg.lstsq; the
// minimum-norm solution for one point) and its root -b/a as float32.
ZR_DEV bool line_root(const float* xs, const float* ys, const int* idx, int n, float* root) {
  if (n <= 0) return false;
  if (n == 1) {
    *root = (float)(-1.0 / (double)xs[idx[0]]);
    return true;
  }
  double mx = 0.0, my = 0.0;
  for (int i = 0; i < n; ++i) { mx += xs[idx[i]]; my += ys[idx[i]]; }
  mx /= n; my /= n;
  double sxx = 0.0, sxy = 0.0;
  for (int i = 0; i < n; ++i) {
    const double dx = (double)xs[idx[i]] - mx;
    sxx += dx * dx;
    sxy += dx * ((double)ys[idx[i]] - my);
  }
  if (!(sxx > 0.0)) return false;
  const double a = sxy / sxx, b = my - a * mx;
  *root = (float)(-b / a);
  return true;
}

constexpr int HOT_FIT_NT = 1024, HOT_FIT_W = HOT_FIT_NT / 64;
constexpr int HOT_GRID = 17;            // coarse grid per axis of the bounded fit
constexpr int HOT_ROUNDS = 40;          // step halvings of the pattern search

// One block: CDF, the bounded normal-CDF fit (17x17 grid over the L-BFGS-B box
// of the reference, then a 3x3 pattern search with halving steps; or the
// caller's fixed (mu, sigma)), the sigma curve and the two cut-offs.
__global__ __launch_bounds__(HOT_FIT_NT) void k_hot_fit(const unsigned* __restrict__ hist, HotArgs a,
                                                        HotState* __restrict__ hs, float* __restrict__ work) {
  __shared__ double cval[HOT_GRID * HOT_GRID];
  __shared__ double cur[3];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int nb = hs->nbins;
  if (nb <= 0) return;
  float* cdf = work;              // [nb]
  float* cnt = work + nb;         // [nb]
  float* xs = work + 2 * nb;      // [nb] normalSigma
  float* ys = work + 3 * nb;      // [nb] normalSigma - sigmaFromCDF
  int* keep = reinterpret_cast<int*>(work + 4 * nb);   // [nb] indices of finite ys
  const double lo = hs->lo;
  if (tid == 0) {
    float run = 0.f;
    for (int k = 0; k < nb; ++k) { run = __fadd_rn(run, (float)hist[k]); cdf[k] = run; }
    for (int k = 0; k < nb; ++k) cdf[k] = __fdiv_rn(cdf[k], run);
  }
  for (int k = tid; k < nb; k += HOT_FIT_NT) cnt[k] = __fadd_rn((float)(lo + k), 0.5f);
  __syncthreads();
  const double m0 = hs->sum_mean, s0 = hs->pstd;
  double mu, sg;
  if (a.mu_fix > 0.0 && a.sigma_fix > 0.0) {
    mu = a.mu_fix;
    sg = a.sigma_fix;
  } else {
    const double lob[2] = {m0 - 0.5 * s0, 0.7 * s0}, hib[2] = {m0 + 0.5 * s0, 1.3 * s0};
    for (int c = wid; c < HOT_GRID * HOT_GRID; c += HOT_FIT_W) {
      const double cm = lob[0] + (hib[0] - lob[0]) * (c / HOT_GRID) / (HOT_GRID - 1);
      const double cs = lob[1] + (hib[1] - lob[1]) * (c % HOT_GRID) / (HOT_GRID - 1);
      const double v = norm_objective(cdf, cnt, nb, cm, cs, lane);
      if (lane == 0) cval[c] = v;
    }
    __syncthreads();
    if (tid == 0) {
      int bi = 0;
      for (int c = 1; c < HOT_GRID * HOT_GRID; ++c)
        if (cval[c] < cval[bi]) bi = c;
      cur[0] = lob[0] + (hib[0] - lob[0]) * (bi / HOT_GRID) / (HOT_GRID - 1);
      cur[1] = lob[1] + (hib[1] - lob[1]) * (bi % HOT_GRID) / (HOT_GRID - 1);
      cur[2] = cval[bi];
    }
    __syncthreads();
    double step[2] = {(hib[0] - lob[0]) / (HOT_GRID - 1), (hib[1] - lob[1]) / (HOT_GRID - 1)};
    for (int it = 0, halv = 0; it < 400 && halv < HOT_ROUNDS; ++it) {
      const double c0 = cur[0], c1 = cur[1], f0 = cur[2];
      if (wid < 9) {
        const double cm = fmin(fmax(c0 + (wid / 3 - 1) * step[0], lob[0]), hib[0]);
        const double cs = fmin(fmax(c1 + (wid % 3 - 1) * step[1], lob[1]), hib[1]);
        const double v = norm_objective(cdf, cnt, nb, cm, cs, lane);
        if (lane == 0) cval[wid] = v;
      }
      __syncthreads();
      int bi = 4;
      for (int c = 0; c < 9; ++c)
        if (cval[c] < cval[bi]) bi = c;
      const bool moved = cval[bi] < f0;
      if (moved && tid == 0) {
        cur[0] = fmin(fmax(c0 + (bi / 3 - 1) * step[0], lob[0]), hib[0]);
        cur[1] = fmin(fmax(c1 + (bi % 3 - 1) * step[1], lob[1]), hib[1]);
        cur[2] = cval[bi];
      }
      if (!moved) { step[0] *= 0.5; step[1] *= 0.5; ++halv; }
      __syncthreads();
    }
    mu = cur[0];
    sg = cur[1];
  }
  // objective at the chosen point, the sigma curve and the kept indices
  if (wid == 0) {
    const double v = norm_objective(cdf, cnt, nb, mu, sg, lane);
    if (lane == 0) hs->objective = v;
  }
  const float muf = (float)mu, sgf = (float)sg;
  const float sq2 = (float)1.4142135623730951;
  for (int k = tid; k < nb; k += HOT_FIT_NT) {
    const float arg = __fsub_rn(__fmul_rn(2.f, cdf[k]), 1.f);
    const float sc = __fmul_rn(sq2, (float)erfinv((double)arg));
    const float ns = __fdiv_rn(__fsub_rn(cnt[k], muf), sgf);
    xs[k] = ns;
    ys[k] = __fsub_rn(ns, sc);
  }
  __syncthreads();
  if (tid == 0) {
    int n = 0;
    for (int k = 0; k < nb; ++k)
      if (!isinf(ys[k])) keep[n++] = k;
    // lower: the points before the first with error > -0.5; upper: from the
    // last with error < 0.5 on (zorro/zorro.py:2572-2592); failures fall back
    // to the defaults
    int li = -1, ui = -1;
    for (int i = 0; i < n; ++i)
      if (ys[keep[i]] > -0.5f) { li = i; break; }
    for (int i = n - 1; i >= 0; --i)
      if (ys[keep[i]] < 0.5f) { ui = i; break; }
    float cl = (float)a.cut_lo_default, cu = (float)a.cut_hi_default, r;
    if (!a.fixed_cutoffs) {
      if (li >= 0 && line_root(xs, ys, keep, li, &r)) cl = r;
      if (ui >= 0 && line_root(xs, ys, keep + ui, n - ui, &r)) cu = r;
    }
    hs->mu = mu;
    hs->sigma = sg;
    hs->cut_lo = cl;
    hs->cut_hi = cu;
    hs->best_mean = mu / a.N;
  }
}

__global__ __launch_bounds__(HOT_NT) void k_hot_mask(const float* __restrict__ usum, size_t hw, HotArgs a,
                                                     HotState* __restrict__ hs, float* __restrict__ lmask) {
  __shared__ unsigned cnt[2];
  if (threadIdx.x < 2) cnt[threadIdx.x] = 0u;
  __syncthreads();
  const float muf = (float)hs->mu, sgf = (float)hs->sigma;
  const float cl = (float)hs->cut_lo, cu = (float)hs->cut_hi;
  const float lk = a.logistic_k, rel = a.relax;
  const float tl = __fmul_rn(cl, rel), th = __fmul_rn(cu, rel);
  unsigned nd = 0, nh = 0;
  for (size_t p = (size_t)blockIdx.x * HOT_NT + threadIdx.x; p < hw; p += (size_t)gridDim.x * HOT_NT) {
    const float us = __fdiv_rn(__fsub_rn(usum[p], muf), sgf);
    const float e1 = expf(__fmul_rn(lk, __fsub_rn(us, tl)));
    const float e2 = expf(__fmul_rn(lk, __fsub_rn(us, th)));
    const double l1 = 1.0 - 1.0 / (1.0 + (double)e1);
    lmask[p] = (float)(l1 / (1.0 + (double)e2));
    nd += us < cl;
    nh += us > cu;
  }
  atomicAdd(&cnt[0], nd);
  atomicAdd(&cnt[1], nh);
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&hs->n_dead, (unsigned long long)cnt[0]);
    atomicAdd(&hs->n_hot, (unsigned long long)cnt[1]);
  }
}

// scipy.ndimage 'reflect' (half-sample symmetric) index
ZR_DEV int refl(int i, int n) { return i < 0 ? -i - 1 : (i >= n ? 2 * n - i - 1 : i); }

// 3x3 convolution with the symmetric zero-centre PSF kernel (k1: edge weight,
// k2: corner weight), float64 accumulation over the kernel in row-major order.
template <typename F>
ZR_DEV float conv3(F val, int y, int x, int H, int W, double k1, double k2) {
  double acc = 0.0;
#pragma unroll
  for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx) {
      if (dy == 0 && dx == 0) continue;
      const double w = (dy != 0 && dx != 0) ? k2 : k1;
      acc += w * (double)val(refl(y + dy, H), refl(x + dx, W));
    }
  return (float)acc;
}

__global__ __launch_bounds__(HOT_NT) void k_hot_nbr(const float* __restrict__ lmask, const float* __restrict__ usum,
                                                    int H, int W, HotArgs a, float* __restrict__ nmask,
                                                    float* __restrict__ sp) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * (HOT_NT / 64) + (threadIdx.x >> 6);
  if (x >= W || y >= H) return;
  const float nfr = (float)a.N;
  auto cl = [&](int yy, int xx) { return __fsub_rn(1.f, lmask[(size_t)yy * W + xx]); };
  auto um = [&](int yy, int xx) { return __fdiv_rn(usum[(size_t)yy * W + xx], nfr); };
  const size_t p = (size_t)y * W + x;
  const float c = cl(y, x);
  const float nm = __fmul_rn(c, conv3(cl, y, x, H, W, a.psf_k1, a.psf_k2));
  nmask[p] = nm;
  const float single = (double)nm <= a.thres ? c : 0.f;
  sp[p] = __fmul_rn(single, conv3(um, y, x, H, W, a.psf_k1, a.psf_k2));
}

// python slice [c-1 : c+2] of an axis of length n -> [s, e)
ZR_DEV void slice3(int c, int n, int& s, int& e) {
  s = c - 1;
  if (s < 0) { s += n; if (s < 0) s = 0; }
  e = c + 2 < n ? c + 2 : n;
  if (e < s) e = s;
}

__global__ __launch_bounds__(HOT_NT) void k_hot_nfilt(const float* __restrict__ nmask, const float* __restrict__ lmask,
                                                      const float* __restrict__ usum, int H, int W, HotArgs a,
                                                      HotState* __restrict__ hs, float* __restrict__ nf) {
  __shared__ unsigned cnt;
  if (threadIdx.x == 0) cnt = 0u;
  __syncthreads();
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * (HOT_NT / 64) + (threadIdx.x >> 6);
  if (x < W && y < H) {
    const size_t p = (size_t)y * W + x;
    float v = 0.f;
    if ((double)nmask[p] > a.thres) {
      atomicAdd(&cnt, 1u);
      int ys, ye, xs, xe;
      slice3(y, H, ys, ye);
      slice3(x, W, xs, xe);
      const int size = (ye - ys) * (xe - xs);
      int nbr = 0;
      for (int yy = ys; yy < ye; ++yy)
        for (int xx = xs; xx < xe; ++xx) nbr += (double)nmask[(size_t)yy * W + xx] > a.thres;
      const double ratio = size > 0 ? (double)nbr / (double)size : NAN;
      if (!(ratio <= 0.66) || ratio <= 0.001) {
        v = (float)hs->best_mean;
      } else {
        const float nfr = (float)a.N;
        float s = 0.f;
        int k = 0;
        for (int yy = ys; yy < ye; ++yy)
          for (int xx = xs; xx < xe; ++xx)
            if (!((double)nmask[(size_t)yy * W + xx] > a.thres)) {
              s = __fadd_rn(s, __fdiv_rn(usum[(size_t)yy * W + xx], nfr));
              ++k;
            }
        v = __fmul_rn(__fsub_rn(1.f, lmask[p]), __fdiv_rn(s, (float)k));
      }
    }
    nf[p] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(&hs->n_nbr, (unsigned long long)cnt);
}

__global__ __launch_bounds__(HOT_NT) void k_hot_apply(float* __restrict__ frames, int N, size_t hw,
                                                      const float* __restrict__ lmask, const float* __restrict__ sp,
                                                      const float* __restrict__ nf) {
  for (size_t p = (size_t)blockIdx.x * HOT_NT + threadIdx.x; p < hw; p += (size_t)gridDim.x * HOT_NT) {
    const float w = lmask[p], s = sp[p], n = nf[p];
    for (int J = 0; J < N; ++J) {
      float* f = frames + (size_t)J * hw + p;
      *f = __fadd_rn(__fadd_rn(__fmul_rn(w, *f), s), n);
    }
  }
}

}  // namespace

size_t zr_hot_bytes(int H, int W, int maxbins) {
  const size_t hw = (size_t)H * W;
  const size_t nblk = (hw + HOT_NT - 1) / HOT_NT;
  // nine regions, each rounded up to 256 bytes by the caller's carving
  return 5 * hw * sizeof(float) + nblk * sizeof(double) + (size_t)maxbins * sizeof(unsigned) +
         (size_t)5 * maxbins * sizeof(float) + sizeof(HotState) + 9 * 256;
}

hipError_t zr_hot_filter(float* frames, int N, int H, int W, const HotArgs& a, const HotBuffers& b, hipStream_t st) {
  const size_t hw = (size_t)H * W;
  const int nblk = (int)((hw + HOT_NT - 1) / HOT_NT);
  hipLaunchKernelGGL(k_hot_sum, dim3(nblk), dim3(HOT_NT), 0, st, frames, N, hw, b.usum, b.part);
  hipLaunchKernelGGL(k_hot_setup, dim3(1), dim3(HOT_NT), 0, st, b.part, nblk, hw, a.max_sigma, a.maxbins, b.state,
                     b.hist);
  const int gsz = 1024;
  hipLaunchKernelGGL(k_hot_hist, dim3(gsz), dim3(HOT_NT), 0, st, b.usum, hw, b.state, b.hist);
  hipLaunchKernelGGL(k_hot_fit, dim3(1), dim3(HOT_FIT_NT), 0, st, b.hist, a, b.state, b.fitwork);
  hipLaunchKernelGGL(k_hot_mask, dim3(gsz), dim3(HOT_NT), 0, st, b.usum, hw, a, b.state, b.lmask);
  const dim3 tiles((W + 63) / 64, (H + HOT_NT / 64 - 1) / (HOT_NT / 64));
  hipLaunchKernelGGL(k_hot_nbr, tiles, dim3(HOT_NT), 0, st, b.lmask, b.usum, H, W, a, b.nmask, b.sp);
  hipLaunchKernelGGL(k_hot_nfilt, tiles, dim3(HOT_NT), 0, st, b.nmask, b.lmask, b.usum, H, W, a, b.state, b.nf);
  hipLaunchKernelGGL(k_hot_apply, dim3(2048), dim3(HOT_NT), 0, st, frames, N, hw, b.lmask, b.sp, b.nf);
  return hipGetLastError();
}
