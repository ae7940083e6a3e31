// Pair cross-correlation path for gfx950:
//   K1  frame -> cropped half-spectrum (row r2c pass + column c2c pass)
//   K2  pair  -> 2m x 2m lag window of the raw inverse DFT of conj(F_I) F_J
//       (pass 1: pruned column transform along ky, pass 2: row transform along kx)
//   KP  window -> origin suppression, B filter, disc argmax, peak significance,
//       Fourier-interpolated sub-pixel shift
//
// Layouts (all row-major, complex64 as float2):
//   frames  f32 [N][H][W]
//   T       c64 [B][Wk][Hp]     row-pass output, kx-major (Wk kept columns)
//   S       c64 [N][Wk][Hs]     kx-major half-spectrum.  Either cropped (Hs = Hc,
//                               rows in FFT order of the cropped grid, plus Sside
//                               [N][Wk] holding the original row +Hc/2), or full
//                               (Hs = Hp, all rows; the crop is a row map with
//                               offset Hp-Hc for ky' >= Hc/2).
//   Z       c64 [P][Wh][2m]     pass-1 output
//   windows f32 [P][2m][2m]     index i <-> lag i-m
#include "zr_common.h"

// --------------------------------------------------------------------- K1
// Row pass: real FFT of length WP (as a complex FFT of WP/2), kept columns
// kx in [0, Wh).  One block = RB consecutive rows of one frame; each row is
// transformed by its own group of TPR threads (the groups share the block's
// barriers, so the RB transforms proceed together).  Every thread issues all of
// its row loads before the first use.
//
// The transformed image is selected by `mode` (frames zero-padded to Hp x WP,
// mask m = 1 on the H x W frame region):
//   K1_FRAME   f                      (frame sums go to `part`)
//   K1_CENTRED m (f - mu_f)           mu_f = stats[2f] / (H W)
//   K1_SQUARED (m (f - mu_f))^2
//   K1_MASK    m
// (m: ones on the H x W frame region, times the user's binary mask umask
// [H][W] when one is given).  Subtracting the frame mean inside the mask leaves the masked normalised
// cross-correlation unchanged (m^2 = m makes every mask term shift with the
// mean in a way that cancels exactly in numerator and denominator) and keeps
// the float32 spectra free of the large DC term.
enum { K1_FRAME = 0, K1_CENTRED = 1, K1_SQUARED = 2, K1_MASK = 3 };
template <int WP, int RB, int TPR>
__global__ __launch_bounds__(RB * TPR) void k1_rows(const float* __restrict__ frames, int H, int W, int Hp, int Wh,
                                                    int frame0, const float2* __restrict__ twL,
                                                    const float2* __restrict__ tw2L, float2* __restrict__ T,
                                                    double* __restrict__ part, int mode,
                                                    const double* __restrict__ stats,
                                                    const float* __restrict__ umask) {
  constexpr int L = WP / 2, NT = RB * TPR, PER = L / TPR;
  static_assert(L % TPR == 0, "row length must split evenly over the row group");
  extern __shared__ float2 k1_lds[];
  float2 (*buf)[zr_fft_lds(L)] = reinterpret_cast<float2 (*)[zr_fft_lds(L)]>(k1_lds);
  __shared__ double red[2][NT / 64];
  const int tid = threadIdx.x;
  const int g = tid / TPR, gt = tid - g * TPR;     // row group, thread in group
  const int nrb = Hp / RB;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int fb = blk / nrb;                 // frame within batch
  const int y0 = (blk % nrb) * RB;
  const int f = frame0 + fb;
  const int y = y0 + g;
  const float* row = frames + ((size_t)f * H + y) * W;
  // twiddles of the row transform (length L) and of the real-FFT split
  // (length WP) in LDS, after the row buffers
  const LdsTw tws = zr_ldstw_fill(k1_lds + RB * zr_fft_lds(L), twL, L, tid, NT);
  const LdsTw tws2 = zr_ldstw_fill(k1_lds + RB * zr_fft_lds(L) + zr_ldstw_size(L), tw2L, WP, tid, NT);
  float2 v[PER];
  // in-mask test: the frame region, times the user mask where one is given
  auto inm = [&](int x) -> bool { return y < H && x < W && (!umask || umask[(size_t)y * W + x] != 0.f); };
  if (mode == K1_MASK) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int x = 2 * (gt + i * TPR);
      v[i].x = inm(x) ? 1.f : 0.f;
      v[i].y = inm(x + 1) ? 1.f : 0.f;
    }
  } else if (y < H && W == WP) {
    const float2* row2 = reinterpret_cast<const float2*>(row);
#pragma unroll
    for (int i = 0; i < PER; ++i) v[i] = row2[gt + i * TPR];
  } else {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int x = 2 * (gt + i * TPR);
      v[i].x = (y < H && x < W) ? row[x] : 0.f;
      v[i].y = (y < H && x + 1 < W) ? row[x + 1] : 0.f;
    }
  }
  if (mode == K1_CENTRED || mode == K1_SQUARED) {
    // stats == NULL: no centring (m f and (m f)^2 as they are)
    const float mu = stats ? (float)(stats[2 * f] / ((double)H * (double)W)) : 0.f;
    const bool sq = (mode == K1_SQUARED);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int x = 2 * (gt + i * TPR);
      const float a = inm(x) ? v[i].x - mu : 0.f;
      const float b = inm(x + 1) ? v[i].y - mu : 0.f;
      v[i].x = sq ? a * a : a;
      v[i].y = sq ? b * b : b;
    }
  }
  double s1 = 0.0, s2 = 0.0;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const double a = v[i].x, b = v[i].y;
    s1 += a + b;
    s2 += a * a + b * b;
  }
  if constexpr (wg_fft_regs_first<L, TPR>()) {   // v[i] = element gt + i TPR: the first pass's inputs
    wg_fft_first_regs<L, -1, TPR>(v, buf[g], gt);
    zr_lds_barrier();
    wg_fft_rest<L, -1, TPR>(buf[g], tws, gt);
  } else {
#pragma unroll
    for (int i = 0; i < PER; ++i) buf[g][gt + i * TPR] = v[i];
    __syncthreads();
    wg_fft<L, -1, TPR>(buf[g], tws, gt);
  }
  // post-process and write transposed in column pairs: T[fb][k/2][y][k%2], so a
  // thread writes RB rows x 2 columns = 16 RB contiguous bytes (the odd pad
  // column of an odd Wh is written as zero)
  const int KP = (Wh + 1) / 2;
  for (int kp = tid; kp < KP; kp += NT) {
    float2 out[2][RB];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int k = 2 * kp + h;
      const bool live = k < Wh;
      const int kk = live ? k : 0;
      const float2 w = tws2.get(kk);   // exp(-2*pi*i*k/WP)
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const float2 zk = buf[r][kk % L];
        const float2 zn = c_conj(buf[r][(L - kk) % L]);
        const float2 e = c_scale(c_add(zk, zn), 0.5f);
        const float2 o = c_scale(c_sub(zk, zn), 0.5f);           // = i * O
        // X = E + w * O,  O = -i * o
        const float2 Ok = make_float2(o.y, -o.x);
        out[h][r] = live ? c_add(e, c_mul(w, Ok)) : make_float2(0.f, 0.f);
      }
    }
    float4* dst = reinterpret_cast<float4*>(T + (((size_t)fb * KP + kp) * Hp + y0) * 2);
#pragma unroll
    for (int r = 0; r < RB; ++r) dst[r] = make_float4(out[0][r].x, out[0][r].y, out[1][r].x, out[1][r].y);
  }
  if (mode != K1_FRAME) return;   // block-uniform
  // frame sums (for the intensity normalisation)
  s1 = wave_sum_d(s1);
  s2 = wave_sum_d(s2);
  if ((tid & 63) == 0) { red[0][tid >> 6] = s1; red[1][tid >> 6] = s2; }
  __syncthreads();
  if (tid == 0) {
    double a = 0, b = 0;
    for (int i = 0; i < NT / 64; ++i) { a += red[0][i]; b += red[1][i]; }
    part[((size_t)f * nrb + (blk % nrb)) * 2 + 0] = a;
    part[((size_t)f * nrb + (blk % nrb)) * 2 + 1] = b;
  }
}

// Column pass: complex FFT of length HP along y for a pair of kept columns
// (kx = 2 kp, 2 kp + 1: one contiguous 16-byte-per-row stream of T), then the
// four-quadrant crop to Hc rows and the side row +Hc/2 (or all HP rows when
// full != 0).  Each column is transformed by its own group of NTC threads (the
// groups share the block's barriers).  Loads are staged in registers before
// the LDS fill.
template <int HP, int NTC>
__global__ __launch_bounds__(2 * NTC) void k1_cols(const float2* __restrict__ T, int Wh, int Hc, int frame0,
                                                   const float2* __restrict__ twH, float2* __restrict__ S,
                                                   float2* __restrict__ Sside, int full) {
  constexpr int NT = 2 * NTC;
  constexpr int PER = (HP + NT - 1) / NT;        // rows per thread in the load
  constexpr int PERC = (HP + NTC - 1) / NTC;     // rows per thread in the store
  extern __shared__ float2 k1c_lds[];
  float2 (*buf)[zr_fft_lds(HP)] = reinterpret_cast<float2 (*)[zr_fft_lds(HP)]>(k1c_lds);
  const int tid = threadIdx.x;
  const int KP = (Wh + 1) / 2;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int fb = blk / KP, kp = blk % KP;
  const float4* src = reinterpret_cast<const float4*>(T + ((size_t)fb * KP + kp) * HP * 2);
  float4 v[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i)
    if (HP % NT == 0 || tid + i * NT < HP) v[i] = src[tid + i * NT];
#pragma unroll
  for (int i = 0; i < PER; ++i)
    if (HP % NT == 0 || tid + i * NT < HP) {
      buf[0][tid + i * NT] = make_float2(v[i].x, v[i].y);
      buf[1][tid + i * NT] = make_float2(v[i].z, v[i].w);
    }
  __syncthreads();
  const int g = tid / NTC, gt = tid - g * NTC;
  wg_fft<HP, -1, NTC>(buf[g], twH, gt);
  const int k = 2 * kp + g;
  if (k >= Wh) return;                            // pad column of an odd Wh (after the last barrier)
  const int f = frame0 + fb;
  const float2* col = buf[g];
  if (full) {
    float2* dst = S + ((size_t)f * Wh + k) * HP;
#pragma unroll
    for (int i = 0; i < PERC; ++i)
      if (HP % NTC == 0 || gt + i * NTC < HP) dst[gt + i * NTC] = col[gt + i * NTC];
    return;
  }
  float2* dst = S + ((size_t)f * Wh + k) * Hc;
  const int h2 = Hc / 2;
  for (int n = gt; n < Hc; n += NTC) dst[n] = col[n < h2 ? n : n + HP - Hc];
  if (gt == 0) Sside[(size_t)f * Wh + k] = col[h2];
}

// Column pass, one column per block: column k read from the column-pair layout
// of T (every other float2; the sibling block of column k^1 reads the rest of
// the same lines, from L2), then as k1_cols.
template <int HP, int NT>
__global__ __launch_bounds__(NT) void k1_cols1(const float2* __restrict__ T, int Wh, int Hc, int frame0,
                                               const float2* __restrict__ twH, float2* __restrict__ S,
                                               float2* __restrict__ Sside, int full) {
  constexpr int PER = (HP + NT - 1) / NT;
  __shared__ float2 buf[zr_fft_lds(HP)];
  __shared__ float2 twl[zr_ldstw_size(HP)];
  const int tid = threadIdx.x;
  const int KP = (Wh + 1) / 2;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int fb = blk / Wh, k = blk % Wh;
  const float2* src = T + (((size_t)fb * KP + k / 2) * HP) * 2 + (k & 1);
  float2 v[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i)
    if (HP % NT == 0 || tid + i * NT < HP) v[i] = src[2 * (tid + i * NT)];
  const LdsTw tws = zr_ldstw_fill(twl, twH, HP, tid, NT);
  if constexpr (wg_fft_regs_first<HP, NT>()) {   // v[i] = element tid + i NT: the first pass's inputs
    wg_fft_first_regs<HP, -1, NT>(v, buf, tid);
    zr_lds_barrier();
    wg_fft_rest<HP, -1, NT>(buf, tws, tid);
  } else {
#pragma unroll
    for (int i = 0; i < PER; ++i)
      if (HP % NT == 0 || tid + i * NT < HP) buf[tid + i * NT] = v[i];
    __syncthreads();
    wg_fft<HP, -1, NT>(buf, tws, tid);
  }
  const int f = frame0 + fb;
  if (full) {
    float2* dst = S + ((size_t)f * Wh + k) * HP;
#pragma unroll
    for (int i = 0; i < PER; ++i)
      if (HP % NT == 0 || tid + i * NT < HP) dst[tid + i * NT] = buf[tid + i * NT];
    return;
  }
  float2* dst = S + ((size_t)f * Wh + k) * Hc;
  const int h2 = Hc / 2;
  for (int n = tid; n < Hc; n += NT) dst[n] = buf[n < h2 ? n : n + HP - Hc];
  if (tid == 0) Sside[(size_t)f * Wh + k] = buf[h2];
}

__global__ __launch_bounds__(256) void k1_frame_stats(const double* __restrict__ part, int nrb, int n,
                                                      double* __restrict__ stats) {
  // stats[f] = (sum, sumsq); one block per frame, fixed reduction order
  __shared__ double ra[4], rb[4];
  const int f = blockIdx.x, tid = threadIdx.x;
  double a = 0, b = 0;
  for (int i = tid; i < nrb; i += 256) { a += part[((size_t)f * nrb + i) * 2]; b += part[((size_t)f * nrb + i) * 2 + 1]; }
  a = wave_sum_d(a);
  b = wave_sum_d(b);
  if ((tid & 63) == 0) { ra[tid >> 6] = a; rb[tid >> 6] = b; }
  __syncthreads();
  if (tid == 0) {
    stats[2 * f] = (ra[0] + ra[1]) + (ra[2] + ra[3]);
    stats[2 * f + 1] = (rb[0] + rb[1]) + (rb[2] + rb[3]);
  }
}

// Frame sums ahead of the masked row pass (which needs each frame's mean):
// block = K1S_RB rows of one frame, FP64 partials in the layout k1_frame_stats
// reduces (part[(f * nrb + row block) * 2 + {0, 1}]).
constexpr int K1S_RB = 8;
__global__ __launch_bounds__(256) void k1_frame_sums(const float* __restrict__ frames, int H, int W,
                                                     double* __restrict__ part) {
  __shared__ double ra[4], rb[4];
  const int nrb = (H + K1S_RB - 1) / K1S_RB;
  const int f = blockIdx.x / nrb, r = blockIdx.x % nrb;
  const int tid = threadIdx.x;
  double a = 0.0, b = 0.0;
  const int y1 = min(H, (r + 1) * K1S_RB);
  for (int y = r * K1S_RB; y < y1; ++y) {
    const float* row = frames + ((size_t)f * H + y) * W;
    for (int x = tid; x < W; x += 256) {
      const double v = row[x];
      a += v;
      b += v * v;
    }
  }
  a = wave_sum_d(a);
  b = wave_sum_d(b);
  if ((tid & 63) == 0) { ra[tid >> 6] = a; rb[tid >> 6] = b; }
  __syncthreads();
  if (tid == 0) {
    part[((size_t)f * nrb + r) * 2 + 0] = (ra[0] + ra[1]) + (ra[2] + ra[3]);
    part[((size_t)f * nrb + r) * 2 + 1] = (rb[0] + rb[1]) + (rb[2] + rb[3]);
  }
}

// Masked normalised cross-correlation from its lag-window terms
// (zorro/zorro.py:1031-1074).  Every term is nc2 * Re R[.] on the 2m x 2m lag
// window, produced by the pair transform for spectrum pairs (a, b) -> conj(S_a) S_b:
//   X  (F_I, F_J)            the pair itself, window p
//   Ct (F_I, M)   Cb (M, F_J)    T (Q_I, M)   B (M, Q_J)   mp (M, M)
// laid out after the P pair windows as [Ct 0..N) [Cb 0..N) [T 0..N) [B 0..N) [mp].
//   C = (X - Ct Cb / mp) / clip(sqrt((B - Cb^2/mp) (T - Ct^2/mp)), 1, inf)
// evaluated in FP64; as with numpy's clip, a NaN denominator stays NaN.
__global__ __launch_bounds__(256) void k_mnxc_combine(float* __restrict__ windows, const int4* __restrict__ pairs,
                                                      int P, int N, int nn) {
  const int4 pr = pairs[blockIdx.x];
  const int I = pr.x, J = pr.y;
  float* X = windows + (size_t)pr.z * nn;
  const float* Ct = windows + (size_t)(P + I) * nn;
  const float* Cb = windows + (size_t)(P + N + J) * nn;
  const float* Tt = windows + (size_t)(P + 2 * N + I) * nn;
  const float* Bt = windows + (size_t)(P + 3 * N + J) * nn;
  const float* mp = windows + (size_t)(P + 4 * N) * nn;
  for (int e = threadIdx.x; e < nn; e += 256) {
    const double ct = Ct[e], cb = Cb[e], mpe = mp[e];
    double den = sqrt((Bt[e] - cb * (cb / mpe)) * (Tt[e] - ct * (ct / mpe)));
    den = (den < 1.0) ? 1.0 : den;
    X[e] = (float)(((double)X[e] - ct * cb / mpe) / den);
  }
}

// triMode 'refine' (zorro/zorro.py:934-956): the template of frame I is the sum
// of all spectra minus its own, so the sums of the spectra are kept as extra
// slots.  dst[e] = sum_k src[slot0 + k][e] over `count` slots, frame order.
__global__ __launch_bounds__(256) void k_spec_sum(float2* __restrict__ S, size_t slot_elems, int slot0, int count,
                                                  int out_slot) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= slot_elems) return;
  float2 acc = make_float2(0.f, 0.f);
  for (int k = 0; k < count; ++k) acc = c_add(acc, S[(size_t)(slot0 + k) * slot_elems + e]);
  S[(size_t)out_slot * slot_elems + e] = acc;
}

// Unmasked 'refine' normalisation: the closed form with the template's sums
// S_T = (S_tot - S_I)/N and SQ_T = (SQ_tot - SQ_I)/N (the template's squared
// spectrum is (sum Q - Q_I)/N, zorro/zorro.py:945-946), both windows of the
// pair (sum, I) and (I, I) take the same scale.
__global__ __launch_bounds__(64) void k_refine_den(const double* __restrict__ stats, int N, double npix, double nc2,
                                                   double* __restrict__ den) {
  __shared__ double tot[2];
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int f = 0; f < N; ++f) { a += stats[2 * f]; b += stats[2 * f + 1]; }
    tot[0] = a;
    tot[1] = b;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < N - 1; k += 64) {
    const double sI = stats[2 * k], qI = stats[2 * k + 1];
    const double sT = (tot[0] - sI) / N, qT = (tot[1] - qI) / N;
    const double a = fmax(npix * qI - sI * sI, 0.0), b = fmax(npix * qT - sT * sT, 0.0);
    den[2 * k] = den[2 * k + 1] = nc2 * sqrt(a * b);
  }
}

// Masked 'refine' transforms every frame centred by one common value c (the
// mean over all frames): unlike a pair, the refine template's squared term is a
// mean of squares, so the formula is not invariant to the frames' means.  Every
// frame gets c through the per-frame sums k1_rows reads: out[2f] = sum_k S_k / N.
__global__ __launch_bounds__(64) void k_common_mean(const double* __restrict__ stats, int N,
                                                    double* __restrict__ out) {
  __shared__ double tot;
  if (threadIdx.x == 0) {
    double a = 0.0;
    for (int f = 0; f < N; ++f) a += stats[2 * f];
    tot = a / N;
  }
  __syncthreads();
  for (int f = threadIdx.x; f < N; f += 64) { out[2 * f] = tot; out[2 * f + 1] = 0.0; }
}

// 'refine' windows from the pair-transform windows (in units of nc2 Re R[.]):
// X = (W[sum, I] - W[I, I]) / N; masked: Ct and T likewise from their sum terms,
// Cb, B and mp as for a pair, then the five-term combination.  Writes window
// out0 + I for I < N-1 (one block each).  Layout of the masked terms after the
// 2(N-1) X windows, base b: Ct [b, b+N), Cb [b+N, ..), T [b+2N, ..), B [b+3N, ..),
// mp b+4N, Ct_sum b+4N+1, T_sum b+4N+2.
// The masked spectra are of frames centred by c (cm: c * H * W as summed by
// k_common_mean, npix = H W); the reference's uncentred terms follow exactly
// (F = F' + cM, Q = Q' + 2cF' + c^2 M, k1 = (N-1)/N):
//   Ct_T = Ct'_T + c mp k1          T_T = T'_T + 2c Ct'_T + c^2 mp k1
//   Cb   = Cb'   + c mp             B   = B'   + 2c Cb'    + c^2 mp
//   X    = X'_T  + c Cb' k1 + c Ct'_T + c^2 mp k1
// and the combination runs in FP64 on them.
__global__ __launch_bounds__(256) void k_refine_combine(float* __restrict__ windows, int N, int nn, int masked,
                                                        int out0, const double* __restrict__ cm, double npix) {
  const int I = blockIdx.x;
  const float invN = 1.f / (float)N;
  const float* Xs = windows + (size_t)(2 * I) * nn;
  const float* Xi = windows + (size_t)(2 * I + 1) * nn;
  float* out = windows + (size_t)(out0 + I) * nn;
  const int b = 2 * (N - 1);
  const double c = masked ? cm[0] / npix : 0.0;
  const double k1 = (double)(N - 1) / (double)N, dN = (double)N;
  for (int e = threadIdx.x; e < nn; e += 256) {
    if (!masked) {
      out[e] = (Xs[e] - Xi[e]) * invN;
      continue;
    }
    const double xc = ((double)Xs[e] - (double)Xi[e]) / dN;
    const double ctc = ((double)windows[(size_t)(b + 4 * N + 1) * nn + e] - (double)windows[(size_t)(b + I) * nn + e]) / dN;
    const double ttc = ((double)windows[(size_t)(b + 4 * N + 2) * nn + e] -
                        (double)windows[(size_t)(b + 2 * N + I) * nn + e]) / dN;
    const double cbc = windows[(size_t)(b + N + I) * nn + e];
    const double bbc = windows[(size_t)(b + 3 * N + I) * nn + e];
    const double mp = windows[(size_t)(b + 4 * N) * nn + e];
    const double ct = ctc + c * mp * k1;
    const double tt = ttc + 2.0 * c * ctc + c * c * mp * k1;
    const double cb = cbc + c * mp;
    const double bb = bbc + 2.0 * c * cbc + c * c * mp;
    const double x = xc + c * cbc * k1 + c * ctc + c * c * mp * k1;
    double den = sqrt((bb - cb * (cb / mp)) * (tt - ct * (ct / mp)));
    den = (den < 1.0) ? 1.0 : den;
    out[e] = (float)((x - ct * cb / mp) / den);
  }
}

// --------------------------------------------------------------------- K2
// Pass 1.  One wave = one (pair, column kx).  ky = a + 64*b with lane a:
//   Y_a[r] = sum_b x[a+64b] e^{+2 pi i b r / L}                  (per lane, registers)
//   Z[dy]  = sum_a e^{+2 pi i a dy / Hc} Y_a[dy mod L]             (LDS exchange)
// with x = c(kx) * conj(F_I) F_J (DC removed on the unmasked path) and the
// Nyquist row combining the original rows -Hc/2 and +Hc/2.  The twiddles
// e^{+2 pi i a dy/Hc} = T1[dy][a/8] T0[dy][a%8] live in LDS, shared by the block.
// Loads are issued in groups so that at most K2_LOADG products are in flight,
// which keeps the kernel at two or more waves per SIMD.
constexpr int K2_LOADG = 8;
template <int L, int WPB, bool PS>
__global__ __launch_bounds__(64 * WPB, 2) void k2_cols(const float2* __restrict__ S, const float2* __restrict__ Sside,
                                                        const int4* __restrict__ work, int nwork_pairs, int Wh,
                                                        int Hc, int Wc, int m, int cols_per_wave, int dc_zero,
                                                        const float2* __restrict__ twHc, float2* __restrict__ Z,
                                                        int Wk, int Hs, int hoff, int side_in_col,
                                                        const int2* __restrict__ ctr,
                                                        const float2* __restrict__ twWc, float* __restrict__ s0) {
  constexpr int LP = 65;  // padded LDS row (conflict-free 64-bit reads per half wave)
  extern __shared__ float2 k2_lds[];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int twoM = 2 * m;
  constexpr int LH = L / 2;                         // Y rows exchanged per half
  float2* T0 = k2_lds + (size_t)WPB * LH * LP;     // [2m][8]
  float2* T1 = T0 + (size_t)twoM * 8;              // [2m][8]
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int kgroup = blk / nwork_pairs;
  const int4 wk = work[blk % nwork_pairs];          // (I, J, pair index, unused)
  const int I = wk.x, J = wk.y, p = wk.z;
  // window centre (preShift): lags t - m + cy, and the sums modulated by e^{2 pi i kx cx/Wc}
  const int2 cc = PS ? ctr[p] : make_int2(0, 0);
  for (int e = threadIdx.x; e < twoM * 8; e += 64 * WPB) {
    const int t = e >> 3, a0 = e & 7;
    const int dy = t - m + cc.x;
    const long dym = ((dy % Hc) + Hc) % Hc;
    const float2 u = twHc[(a0 * dym) % Hc], v = twHc[(8 * a0 * dym) % Hc];
    T0[e] = make_float2(u.x, -u.y);
    T1[e] = make_float2(v.x, -v.y);
  }
  __syncthreads();
  // the second sweep (outputs 64..2m-1) is spread over `sf` lanes per output
  const int c2 = twoM > 64 ? twoM - 64 : 0;
  const int sf = c2 == 0 ? 1 : (c2 <= 16 ? 4 : (c2 <= 32 ? 2 : 1));
  const int span = 64 / sf;                          // lanes per part
  const int o2 = lane % span, part = lane / span;
  const int t1 = lane, t2 = 64 + o2;
  const int r1 = (((t1 - m + cc.x) % L) + L) % L, r2 = (((t2 - m + cc.x) % L) + L) % L;
  const int a1lo = part * (8 / sf), a1hi = a1lo + 8 / sf;
  const int kbase = kgroup * WPB * cols_per_wave;
  float2* my = k2_lds + (size_t)w * LH * LP;
  for (int c = 0; c < cols_per_wave; ++c) {
    const int kx = kbase + c * WPB + w;
    if (kx >= Wh) break;
    const float2* sI = S + ((size_t)I * Wk + kx) * Hs;
    const float2* sJ = S + ((size_t)J * Wk + kx) * Hs;
    const float cw = (kx == 0 || 2 * kx == Wc) ? 1.f : 2.f;
    // Nyquist row (ky' = Hc/2 <-> lane 0, b = L/2): rows -Hc/2 (main) and +Hc/2
    // (side), combined per column; wave-uniform loads, applied by a lane select
    float2 nv;
    {
      const float2 nm = c_cjmul(sI[Hc / 2 + hoff], sJ[Hc / 2 + hoff]);
      const float2 ns = side_in_col >= 0 ? c_cjmul(sI[side_in_col], sJ[side_in_col])
                                         : c_cjmul(Sside[(size_t)I * Wk + kx], Sside[(size_t)J * Wk + kx]);
      nv = (kx == 0) ? nm : ((2 * kx == Wc) ? ns : c_add(nm, ns));
    }
    float2 v[L];
#pragma unroll
    for (int g0 = 0; g0 < L; g0 += K2_LOADG) {
      float2 xi[K2_LOADG], xj[K2_LOADG];
#pragma unroll
      for (int j = 0; j < K2_LOADG; ++j) {
        const int b = g0 + j;
        const int off = (b >= L / 2) ? hoff : 0;   // cropped row ky' -> stored row
        xi[j] = sI[lane + 64 * b + off];
        xj[j] = sJ[lane + 64 * b + off];
      }
#pragma unroll
      for (int j = 0; j < K2_LOADG; ++j) v[g0 + j] = c_scale(c_cjmul(xi[j], xj[j]), cw);
      __builtin_amdgcn_sched_barrier(0);
    }
    {
      const bool l0 = (lane == 0);
      const bool dc = l0 && dc_zero && kx == 0;
      v[0] = dc ? make_float2(0.f, 0.f) : v[0];
      v[L / 2] = l0 ? nv : v[L / 2];
    }
    Dft<L, 1>::run(v);
    float2 xmod = make_float2(1.f, 0.f);
    if constexpr (PS) {
      const float sr = wave_sum_f(v[0].x);    // lag-0 column sum (c(kx) is inside the products)
      if (lane == 0) s0[(size_t)p * Wh + kx] = sr;
      const float2 t = twWc[(int)((((long)kx * cc.y) % Wc + Wc) % Wc)];
      xmod = make_float2(t.x, -t.y);
    }
    float2* zdst = Z + ((size_t)p * Wh + kx) * twoM;
    // two halves of r through LDS; each output reads the half holding its r
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
      for (int r = 0; r < LH; ++r) my[r * LP + lane] = v[h * LH + r];
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (t1 < twoM && (r1 / LH) == h) {   // first sweep: outputs 0..63, one per lane
        const float2* yr = my + (r1 - h * LH) * LP;
        const float2* t0 = T0 + t1 * 8;
        const float2* tt1 = T1 + t1 * 8;
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll
        for (int a1 = 0; a1 < 8; ++a1) {
          float2 pa = make_float2(0.f, 0.f);
#pragma unroll
          for (int a0 = 0; a0 < 8; ++a0) {
            const float2 yv = yr[8 * a1 + a0], tw = t0[a0];
            pa.x += tw.x * yv.x - tw.y * yv.y;
            pa.y += tw.x * yv.y + tw.y * yv.x;
          }
          const float2 tw = tt1[a1];
          acc.x += tw.x * pa.x - tw.y * pa.y;
          acc.y += tw.x * pa.y + tw.y * pa.x;
        }
        zdst[t1] = PS ? c_mul(acc, xmod) : acc;
      }
      if (c2 > 0) {      // second sweep: outputs 64..2m-1, `sf` partial sums each
        float2 acc = make_float2(0.f, 0.f);
        const bool mine = (o2 < c2) && ((r2 / LH) == h);
        if (mine) {
          const float2* yr = my + (r2 - h * LH) * LP;
          const float2* t0 = T0 + t2 * 8;
          const float2* tt1 = T1 + t2 * 8;
          for (int a1 = a1lo; a1 < a1hi; ++a1) {
            float2 pa = make_float2(0.f, 0.f);
#pragma unroll
            for (int a0 = 0; a0 < 8; ++a0) {
              const float2 yv = yr[8 * a1 + a0], tw = t0[a0];
              pa.x += tw.x * yv.x - tw.y * yv.y;
              pa.y += tw.x * yv.y + tw.y * yv.x;
            }
            const float2 tw = tt1[a1];
            acc.x += tw.x * pa.x - tw.y * pa.y;
            acc.y += tw.x * pa.y + tw.y * pa.x;
          }
        }
        for (int o = span; o < 64; o <<= 1) {
          acc.x += __shfl_xor(acc.x, o, 64);
          acc.y += __shfl_xor(acc.y, o, 64);
        }
        if (part == 0 && mine) zdst[t2] = PS ? c_mul(acc, xmod) : acc;
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
  }
}

// Pass 1 for m <= L (every residue r = dy mod L has at most two outputs:
// dy = r and dy = r - L).  With Hc = 64 L and dy = r + L j,
//   Z[dy] = sum_a e^{+2 pi i a r / Hc} e^{+2 pi i a j / 64} Y_a[r],
// so lane a multiplies its Y_a[r] by the twiddle e^{+2 pi i a r/Hc}; output
// dy = r is the plain sum of these products over the 64 lanes and output
// dy = r - L their sum weighted by e^{-2 pi i a/64}.  The sums go through LDS,
// one sixteen-residue group at a time: each residue's row of 64 products is
// read by four lanes (16 terms each, lanes a = 16 p + i), each forming both
// sums (the weight is (-i)^p e^{-2 pi i i/64}, with compile-time constants),
// and the four partial sums are added by two lane swaps.  The twiddle table
// ([L+1][64]) is copied to LDS once per block.  The column weight c(kx) is applied to the sums (a power of two: the
// same result as scaling every product).  Loads: the column's rows in two
// batches (K2R2_B1 rows of both frames, then the rest), each issued at once.
// A window centre (cy, cx) per pair (ctr != NULL, preShift) moves the lags to
// [c - m, c + m): residue r then has lags dy_a = r + L j1(r) (the largest
// <= cy + m - 1) and dy_a - L, so row r of the block's table becomes
// e^{+2 pi i a dy_a/Hc} (from twHc), and the sums are modulated by
// e^{+2 pi i kx cx/Wc}.  s0 != NULL also stores c(kx) Re sum_ky x per column
// (the lag-0 correlation's column sums, for the origin value).
constexpr int K2R2_ROWS = 16, K2R2_B1 = 32, K2R2_WPB = 4, K2R2_LP = 65, K2R2_OCC = 2;
// cos and sin of 2 pi i / 64, i < 16 (float32 of the FP64 values): the lag
// dy - L reader weights e^{-2 pi i i / 64}
constexpr float K2R2_WC[16] = {1.000000000e+00f, 9.951847196e-01f, 9.807852507e-01f, 9.569403529e-01f,
                               9.238795042e-01f, 8.819212914e-01f, 8.314695954e-01f, 7.730104327e-01f,
                               7.071067691e-01f, 6.343932748e-01f, 5.555702448e-01f, 4.713967443e-01f,
                               3.826834261e-01f, 2.902846634e-01f, 1.950903237e-01f, 9.801714122e-02f};
constexpr float K2R2_WS[16] = {0.000000000e+00f, 9.801714122e-02f, 1.950903237e-01f, 2.902846634e-01f,
                               3.826834261e-01f, 4.713967443e-01f, 5.555702448e-01f, 6.343932748e-01f,
                               7.071067691e-01f, 7.730104327e-01f, 8.314695954e-01f, 8.819212914e-01f,
                               9.238795042e-01f, 9.569403529e-01f, 9.807852507e-01f, 9.951847196e-01f};
// Residue r of output k2 of sixteen-point group k1 in the length-L = 16 R1 lane
// DFT of k2_cols_r2.  R1 odd (coprime to 16): the prime-factor (Good-Thomas)
// split, input n = (16 n1 + R1 n2) mod L, output r = (16 a k1 + R1 b k2) mod L
// with a = 16^-1 mod R1, b = R1^-1 mod 16, which needs no twiddles between the
// R1-point and sixteen-point DFTs; R1 even: Cooley-Tukey, n = 16 n1 + n2,
// r = k1 + R1 k2, with the twiddles.
constexpr int k2r2_inv(int x, int m) {
  for (int i = 1; i < m; ++i)
    if ((x * i) % m == 1) return i;
  return 1;
}
template <int L> ZR_DEV constexpr int k2r2_residue(int k1, int k2) {
  constexpr int R1 = L / 16;
  if constexpr (R1 % 2 == 1) {
    constexpr int A = 16 * k2r2_inv(16 % R1, R1), B = R1 * k2r2_inv(R1, 16);
    return (A * k1 + B * k2) % L;
  } else {
    return k1 + R1 * k2;
  }
}
constexpr size_t k2r2_lds_bytes(int L) {
  return ((size_t)K2R2_WPB * K2R2_ROWS * K2R2_LP + (size_t)(L + 1) * 64) * sizeof(float2);
}
template <int L, int WPB, bool PS>
__global__ __launch_bounds__(64 * WPB, K2R2_OCC) void k2_cols_r2(const float2* __restrict__ S, const float2* __restrict__ Sside,
                                                           const int4* __restrict__ work, int nwork_pairs, int Wh,
                                                           int Hc, int Wc, int m, int cols_per_wave, int pairs_per_wave,
                                                           int dc_zero,
                                                           const float2* __restrict__ ltw, float2* __restrict__ Z,
                                                           int Wk, int Hs, int hoff, int side_in_col,
                                                           const int2* __restrict__ ctr,
                                                           const float2* __restrict__ twHc,
                                                           const float2* __restrict__ twWc, float* __restrict__ s0) {
  constexpr int LP = K2R2_LP;             // padded LDS row: rows 2 dwords apart in the banks
  constexpr int ROWS = K2R2_ROWS;         // one sixteen-residue group
  static_assert(ROWS * 4 == 64, "four readers per row");
  extern __shared__ float2 k2_lds[];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave index (uniform)
  float2* my = k2_lds + (size_t)w * ROWS * LP;
  float2* tab = k2_lds + (size_t)WPB * ROWS * LP;                    // [L+1][64]
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  // Column-group mode (pairs_per_wave == 0): block = (group of WPB *
  // cols_per_wave columns, one pair), each wave walks its columns.  Pair mode
  // (pairs_per_wave > 0, not with PS): block = (one column, WPB * pairs_per_wave
  // consecutive entries of the pair list), each wave walks pairs w, w + WPB, ...
  // so the blocks resident on an XCD at one time read one column of every frame
  // (40 frames x 24 KiB at Hc = 3072) and the spectra are fetched from HBM about
  // once.
  const bool pmode = !PS && pairs_per_wave > 0;
  const int npc = pmode ? (nwork_pairs + WPB * pairs_per_wave - 1) / (WPB * pairs_per_wave) : 1;
  const int kgroup = pmode ? blk / npc : blk / nwork_pairs;
  const int4 wk = work[pmode ? 0 : blk % nwork_pairs];
  const int twoM = 2 * m;
  const int2 cc = PS ? ctr[wk.z] : make_int2(0, 0);
  auto j1of = [&](int r) -> int {           // largest j with r + L j <= cy + m - 1
    const int t = cc.x + m - 1 - r;
    return t >= 0 ? t / L : -((-t + L - 1) / L);
  };
  if constexpr (!PS) {
    for (int e = threadIdx.x; e < (L + 1) * 64; e += 64 * WPB) tab[e] = ltw[e];
  } else {
    for (int e = threadIdx.x; e < (L + 1) * 64; e += 64 * WPB) {
      const int r = e / 64, a = e % 64;
      if (r == L) { tab[e] = ltw[e]; continue; }
      const long dy = r + (long)L * j1of(r);
      const long ph = (((long)a * dy) % Hc + Hc) % Hc;
      const float2 t = twHc[ph];
      tab[e] = make_float2(t.x, -t.y);
    }
  }
  __syncthreads();
  // reader role: row slot (residue r0 + slot/2, kind slot%2), part of the lanes
  const int slot = lane & (ROWS - 1), part = lane / ROWS;
  const int kbase = kgroup * WPB * cols_per_wave;
  const int pbase = (blk % npc) * WPB * pairs_per_wave;
  const int nsteps = pmode ? pairs_per_wave : cols_per_wave;
  for (int c = 0; c < nsteps; ++c) {
    int kx, I, J, p;
    if (pmode) {
      const int e = pbase + c * WPB + w;
      if (e >= nwork_pairs) break;
      const int4 we = work[e];
      kx = kgroup; I = we.x; J = we.y; p = we.z;
    } else {
      kx = kbase + c * WPB + w;
      if (kx >= Wh) break;
      I = wk.x; J = wk.y; p = wk.z;
    }
    const float2* sI = S + ((size_t)I * Wk + kx) * Hs;
    const float2* sJ = S + ((size_t)J * Wk + kx) * Hs;
    const float cw = (kx == 0 || 2 * kx == Wc) ? 1.f : 2.f;
    float2 nv;
    {
      const float2 nm = c_cjmul(sI[Hc / 2 + hoff], sJ[Hc / 2 + hoff]);
      const float2 ns = side_in_col >= 0 ? c_cjmul(sI[side_in_col], sJ[side_in_col])
                                         : c_cjmul(Sside[(size_t)I * Wk + kx], Sside[(size_t)J * Wk + kx]);
      nv = (kx == 0) ? nm : ((2 * kx == Wc) ? ns : c_add(nm, ns));
      nv = c_scale(nv, 1.f / cw);          // exact: the sums are scaled by cw below
    }
    // column loads through buffer resources: per-row offsets stay scalar
    const __amdgpu_buffer_rsrc_t rI = __builtin_amdgcn_make_buffer_rsrc((void*)sI, (short)0, Hs * 8, 0x00020000);
    const __amdgpu_buffer_rsrc_t rJ = __builtin_amdgcn_make_buffer_rsrc((void*)sJ, (short)0, Hs * 8, 0x00020000);
    float2 v[L];
    auto ld = [&](const __amdgpu_buffer_rsrc_t& rs, int b) -> float2 {
      const int soff = (64 * b + ((b >= L / 2) ? hoff : 0)) * 8;
      return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, lane * 8, soff, 0));
    };
    constexpr int B1 = L < K2R2_B1 ? L : K2R2_B1;
    {
      float2 t[B1];
#pragma unroll
      for (int b = 0; b < B1; ++b) { v[b] = ld(rI, b); t[b] = ld(rJ, b); }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int b = 0; b < B1; ++b) v[b] = c_cjmul(v[b], t[b]);
      __builtin_amdgcn_sched_barrier(0);   // the second batch's loads after these products
    }
    if constexpr (L > B1) {
      float2 t[L - B1];
#pragma unroll
      for (int b = B1; b < L; ++b) { v[b] = ld(rI, b); t[b - B1] = ld(rJ, b); }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int b = B1; b < L; ++b) v[b] = c_cjmul(v[b], t[b - B1]);
    }
    {
      const bool l0 = (lane == 0);
      const bool dc = l0 && dc_zero && kx == 0;
      v[0] = dc ? make_float2(0.f, 0.f) : v[0];
      v[L / 2] = l0 ? nv : v[L / 2];
    }
    // The length-L DFT as L/16-point DFTs then L/16 sixteen-point DFTs (output
    // k2 of group k1 is residue k2r2_residue<L>(k1, k2)), with each
    // sixteen-point group's residues twiddled and lane-summed as soon as it is
    // done, so that at most one group's outputs are live beside the remaining
    // inputs.
    constexpr int R1 = L / 16;
    constexpr bool PFA = (R1 % 2) == 1;
    float2 t[L];
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) {
      float2 u[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) u[n1] = v[PFA ? (16 * n1 + R1 * n2) % L : 16 * n1 + n2];
      Dft<R1, 1>::run(u);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) t[n2 * R1 + k1] = u[k1];
    }
    if constexpr (!PFA) {
#pragma unroll
      for (int n2 = 1; n2 < 16; ++n2)
#pragma unroll
        for (int k1 = 1; k1 < R1; ++k1) t[n2 * R1 + k1] = DftCT<R1, 16, 1>::mul_tw(t[n2 * R1 + k1], n2 * k1);
    }
    float2 xmod = make_float2(cw, 0.f);
    float2* zdst = Z + ((size_t)p * Wh + kx) * twoM;
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      __builtin_amdgcn_sched_barrier(0);
      // the group's sixteen lane twiddles, read before its DFT so that the
      // LDS latency hides behind the DFT
      float2 tw16[16];
#pragma unroll
      for (int n = 0; n < 16; ++n) tw16[n] = tab[k2r2_residue<L>(k1, n) * 64 + lane];
      __builtin_amdgcn_sched_barrier(0);
      float2 y[16];                            // residues k2r2_residue<L>(k1, k2), k2 = 0 .. 15
#pragma unroll
      for (int n2 = 0; n2 < 16; ++n2) y[n2] = t[n2 * R1 + k1];
      Dft<16, 1>::run(y);
      if constexpr (PS) {
        if (k1 == 0) {
          const float sr = wave_sum_f(y[0].x);    // lag-0 column sum: sum over lanes of Y_a[0]
          if (lane == 0) s0[(size_t)p * Wh + kx] = sr * cw;
          const float2 tt = twWc[(int)((((long)kx * cc.y) % Wc + Wc) % Wc)];
          xmod = make_float2(tt.x * cw, -tt.y * cw);
        }
      }
      // the group's sixteen residues, u = Y_a[r] times the lane twiddle, as the
      // rows of the wave's LDS tile (rows of lags outside the window are
      // written too and never stored)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < 16; ++i) my[i * LP + lane] = c_mul(y[i], tw16[i]);
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      {
        // reader (row slot, part p = lanes a = 16 p + i): its sixteen terms as
        // they are (lag dy) and times e^{-2 pi i a / 64} = (-i)^p e^{-2 pi i i / 64}
        // (lag dy - L)
        const int r = k2r2_residue<L>(k1, slot);
        const int dy = PS ? r + L * j1of(r) : r;     // the larger of the residue's two lags
        const float2* row = my + slot * LP + 16 * part;
        float2 a0 = row[0], a1 = a0;            // term 0 has weight 1
#pragma unroll
        for (int i = 1; i < 16; ++i) {
          const float2 x = row[i];
          a0 = c_add(a0, x);
          a1.x = fmaf(x.x, K2R2_WC[i], fmaf(x.y, K2R2_WS[i], a1.x));
          a1.y = fmaf(x.y, K2R2_WC[i], fmaf(-x.x, K2R2_WS[i], a1.y));
        }
        {
          const bool sw = part & 1, nx = part >= 2, ny = part == 1 || part == 2;
          const float bx = sw ? a1.y : a1.x, by = sw ? a1.x : a1.y;
          a1 = make_float2(nx ? -bx : bx, ny ? -by : by);
        }
        a0.x = xor32_sum(xor16_sum(a0.x));
        a0.y = xor32_sum(xor16_sum(a0.y));
        a1.x = xor32_sum(xor16_sum(a1.x));
        a1.y = xor32_sum(xor16_sum(a1.y));
        if (part == 0) {
          if (dy >= cc.x - m && dy <= cc.x + m - 1) zdst[dy - cc.x + m] = PS ? c_mul(a0, xmod) : c_scale(a0, cw);
          const int dy1 = dy - L;
          if (dy1 >= cc.x - m && dy1 <= cc.x + m - 1) zdst[dy1 - cc.x + m] = PS ? c_mul(a1, xmod) : c_scale(a1, cw);
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
  }
}

// Lane twiddles of k2_cols_r2, row-major [L + 1][64], from the angle in FP64:
// rows r < L hold e^{+2 pi i a r / Hc}, row L holds e^{-2 pi i a / 64}.
__global__ __launch_bounds__(256) void k2_lane_twiddles(int L, int Hc, float2* __restrict__ ltw) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= (L + 1) * 64) return;
  const int r = e / 64, a = e % 64;
  const long ph = r < L ? (long)a * r % Hc : (long)(Hc - a * L) % Hc;
  const double ang = 2.0 * 3.14159265358979323846 * (double)ph / (double)Hc;
  ltw[e] = make_float2((float)cos(ang), (float)sin(ang));
}

// Pass 2.  One block per pair: W(dy,dx) = sum_kx Re(Z[kx][dy] e^{+2 pi i kx dx/Wc}),
// then C = nc2 * W / max(1, Denom).  Thread owns 4 dx x 5 dy outputs.
constexpr int K2R_DX = 4, K2R_DY = 5, K2R_KC = 32;
__global__ __launch_bounds__(1024) void k2_rows(const float2* __restrict__ Z, const int4* __restrict__ pairs, int Wh,
                                                int Wc, int m, const float2* __restrict__ twWc,
                                                const double* __restrict__ den, float nc2,
                                                float* __restrict__ windows, const int4* __restrict__ sub) {
  extern __shared__ float2 zs[];  // [K2R_KC][2m]
  const int twoM = 2 * m;
  const int ndxg = (twoM + K2R_DX - 1) / K2R_DX;
  const int p = sub ? sub[blockIdx.x].z : blockIdx.x;     // sub: a subset of the pairs (preShift diagonals)
  const int tid = threadIdx.x;
  const int gx = tid % ndxg, gy = tid / ndxg;
  const int dx0 = gx * K2R_DX, dy0 = gy * K2R_DY;
  float acc[K2R_DY][K2R_DX];
#pragma unroll
  for (int i = 0; i < K2R_DY; ++i)
#pragma unroll
    for (int j = 0; j < K2R_DX; ++j) acc[i][j] = 0.f;
  int idx[K2R_DX], dd[K2R_DX];
#pragma unroll
  for (int j = 0; j < K2R_DX; ++j) {
    int d = dx0 + j - m;
    dd[j] = ((d % Wc) + Wc) % Wc;
    idx[j] = 0;
  }
  const float2* zp = Z + (size_t)p * Wh * twoM;
  for (int k0 = 0; k0 < Wh; k0 += K2R_KC) {
    const int kc = min(K2R_KC, Wh - k0);
    __syncthreads();
    for (int e = tid; e < kc * twoM; e += blockDim.x) zs[e] = zp[(size_t)k0 * twoM + e];
    __syncthreads();
    if (dy0 < twoM) {
      for (int k = 0; k < kc; ++k) {
        float2 tw[K2R_DX];
#pragma unroll
        for (int j = 0; j < K2R_DX; ++j) {
          float2 t = twWc[idx[j]];
          tw[j] = make_float2(t.x, -t.y);
          idx[j] += dd[j];
          if (idx[j] >= Wc) idx[j] -= Wc;
        }
        const float2* zr = zs + k * twoM;
#pragma unroll
        for (int i = 0; i < K2R_DY; ++i) {
          int dy = dy0 + i;
          float2 z = (dy < twoM) ? zr[dy] : make_float2(0.f, 0.f);
#pragma unroll
          for (int j = 0; j < K2R_DX; ++j) acc[i][j] += z.x * tw[j].x - z.y * tw[j].y;
        }
      }
    }
  }
  if (dy0 < twoM && dx0 < twoM) {
    const float s = (float)((double)nc2 / fmax(1.0, den[p]));
    float* wdst = windows + (size_t)p * twoM * twoM;
#pragma unroll
    for (int i = 0; i < K2R_DY; ++i)
#pragma unroll
      for (int j = 0; j < K2R_DX; ++j) {
        int dy = dy0 + i, dx = dx0 + j;
        if (dy < twoM && dx < twoM) wdst[dy * twoM + dx] = acc[i][j] * s;
      }
  }
}

// Pass 2 on the matrix cores (2m a multiple of 16).  The row DFT over kx is a
// real GEMM per pair with K = 2 Wh (real and imaginary parts along kx):
//   W[dy][dx] = sum_kx Re(Z[kx][dy] e^{+2 pi i kx (dx - m) / Wc}) = sum_kx (Zr Bc + Zi Bs)
// with Btab[kx][dx] = (Bc, Bs) = (cos, -sin)(2 pi kx (dx - m) / Wc), then
// C = nc2 W / max(1, den).  v_mfma_f32_16x16x4f32: lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16] of a 16x4 / 4x16 step and holds
// D[4 (l / 16) + v][l % 16].
//
// Work item = (pair, 16-row strip dy in [16 s, 16 s + 16)) with all 2m/16
// column tiles, one wave each; a block is K2M_WPB = 4 items, so that its waves
// fill the CU's four SIMDs evenly (one block per pair, with 2m/16 = 5 waves,
// left one SIMD with two of them and every barrier waited for it).  Any items
// can share a block: the twiddle table is the same for every pair.  K order:
// for the four columns kx = kx0 + 4u + (0..3) one step takes the real parts and
// the next the imaginary parts, so both of a lane's A operands are one float2
// of Z (column kx0 + 4u + l / 16, row 16 s + l % 16).  The strip is the wave's
// alone, so Z goes from global memory straight into registers one chunk ahead
// of its use; the twiddle chunk (L2-resident) is staged through a
// double-buffered LDS tile with one barrier per chunk.  The table's zero rows
// pad Wh to an even number of chunks: the chunks are taken two at a time with
// no branch in the loop body (a branch lets the compiler sink the next chunk's
// Z loads past the current chunk's MFMAs), and the clamped Z reads past the
// last column meet zero twiddles.
typedef float zr_v4f __attribute__((ext_vector_type(4)));
constexpr int K2M_WPB = 4;           // work items (waves) per block
constexpr int K2M_KC = 16;           // kx per chunk (LDS: 2 x K2M_KC x 2m float2)
template <int NT>
__global__ __launch_bounds__(64 * K2M_WPB) void k2_rows_mfma(const float2* __restrict__ Z, int Wh, int nitems,
                                                             const float2* __restrict__ Btab,
                                                             const double* __restrict__ den, float nc2,
                                                             float* __restrict__ windows,
                                                             const int4* __restrict__ sub) {
  constexpr int TWOM = 16 * NT, NTH = 64 * K2M_WPB;
  constexpr int KC = K2M_KC, UG = KC / 4;            // four-column groups per chunk
  constexpr int PERB = KC * TWOM / NTH;             // twiddle float2 per thread per chunk
  static_assert(PERB * NTH == KC * TWOM, "the twiddle chunk splits evenly over the block");
  __shared__ float2 Bs[2][KC * TWOM];               // [buffer][kx][dx]
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // past the last item a wave only helps stage the twiddles
  const int item = min((int)blockIdx.x * K2M_WPB + w, nitems - 1);
  const int pi = item / NT, strip = item - pi * NT;
  const int p = sub ? sub[pi].z : pi;               // sub: a subset of the pairs (preShift diagonals)
  const float2* zp = Z + (size_t)p * Wh * TWOM + 16 * strip + (lane & 15);
  const int kq = lane >> 4;
  const int nch = (Wh + 2 * KC - 1) / (2 * KC) * 2;
  zr_v4f acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = zr_v4f{0.f, 0.f, 0.f, 0.f};
  auto loadA = [&](float2 (&a)[UG], int c) {
#pragma unroll
    for (int u = 0; u < UG; ++u) a[u] = zp[(size_t)min(c * KC + 4 * u + kq, Wh - 1) * TWOM];
  };
  float2 rb[PERB];
  auto loadB = [&](int c) {
#pragma unroll
    for (int i = 0; i < PERB; ++i) rb[i] = Btab[(size_t)c * KC * TWOM + tid + i * NTH];
  };
  auto storeB = [&](int buf) {
#pragma unroll
    for (int i = 0; i < PERB; ++i) Bs[buf][tid + i * NTH] = rb[i];
  };
  // chunk c: its MFMAs (A from a, B from Bs[c & 1]); then chunk c + 1's
  // twiddles go to the other buffer and chunk c + 2's are requested
  auto chunk = [&](const float2 (&a)[UG], int c) {
    const float2* bs = Bs[c & 1] + kq * TWOM + (lane & 15);
#pragma unroll
    for (int u = 0; u < UG; ++u) {
      __builtin_amdgcn_sched_barrier(0);   // one group's twiddle reads live at a time
      float2 b[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) b[t] = bs[4 * u * TWOM + 16 * t];
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].x, b[t].x, acc[t], 0, 0, 0);
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].y, b[t].y, acc[t], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    storeB((c + 1) & 1);                   // after the last chunk: padding rows, never read
    zr_lds_barrier();                      // LDS only: the Z loads stay in flight
    loadB(min(c + 2, nch - 1));
  };
  float2 a0[UG], a1[UG];
  loadB(0);
  loadA(a0, 0);
  storeB(0);
  loadB(1);
  zr_lds_barrier();
  for (int c = 0; c < nch; c += 2) {
    loadA(a1, c + 1);
    chunk(a0, c);
    loadA(a0, c + 2);
    chunk(a1, c + 1);
  }
  if ((int)blockIdx.x * K2M_WPB + w >= nitems) return;
  const float sc = (float)((double)nc2 / fmax(1.0, den[p]));
  float* wdst = windows + (size_t)p * TWOM * TWOM;
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) wdst[(16 * strip + 4 * kq + v) * TWOM + 16 * t + (lane & 15)] = acc[t][v] * sc;
}

// Twiddles of the matrix-core row pass: Btab[kx][dx] = (cos, -sin)(2 pi kx (dx - m) / Wc),
// phase index kx (dx - m) mod Wc taken exactly in integers, FP64 sin/cos; rows
// Wh .. Whp-1 (the padding to whole pairs of the pass's chunks) are zero.
__global__ __launch_bounds__(256) void k2_rows_twiddles(int Wh, int Whp, int Wc, int m, float2* __restrict__ Btab) {
  const int twoM = 2 * m;
  for (int e = blockIdx.x * 256 + threadIdx.x; e < Whp * twoM; e += gridDim.x * 256) {
    const int kx = e / twoM, dx = e - kx * twoM;
    if (kx >= Wh) { Btab[e] = make_float2(0.f, 0.f); continue; }
    const long d = (((long)(dx - m) % Wc) + Wc) % Wc;
    const long ph = ((long)kx * d) % Wc;
    const double a = 2.0 * 3.14159265358979323846 * (double)ph / (double)Wc;
    Btab[e] = make_float2((float)cos(a), (float)(-sin(a)));
  }
}

// Per-pair normalisation denominator on the unmasked path (FP64):
// D = nc2 * sqrt((Np*SQ_J - S_J^2)(Np*SQ_I - S_I^2))
__global__ void k2_denominators(const int4* __restrict__ pairs, int npairs, const double* __restrict__ stats,
                                double npix, double nc2, double* __restrict__ den) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  int4 pr = pairs[i];
  double sI = stats[2 * pr.x], qI = stats[2 * pr.x + 1];
  double sJ = stats[2 * pr.y], qJ = stats[2 * pr.y + 1];
  double a = fmax(npix * qJ - sJ * sJ, 0.0), b = fmax(npix * qI - sI * sI, 0.0);
  den[pr.z] = nc2 * sqrt(a * b);
}

// --------------------------------------------------------------------- KP
struct PeakParams {
  int m;             // maxShift
  int suppress;      // suppressOrigin
  int bconv;         // 1: Gaussian on the window (Bmode 'conv')
  int bopti;         // 1: Gaussian width chosen per pair (Bmode 'opti')
  int ry, rx;        // Gaussian radii
  int subpix;        // subPixReg (<=1: integer)
  float scale_y, scale_x;   // Hp/Hc, Wp/Wc (1 when no crop)
  int n;             // frames (trimat side)
  int tmp_floats;    // LDS floats ahead of the filtered window
  int disc;          // radius of the maxShift disc (= m unless preShift)
  int preshift;      // 1: windows centred on per-pair lags ctr[p] (preShift)
};

ZR_DEV int reflect_idx(int i, int n) {
  int p = 2 * n;
  i %= p;
  if (i < 0) i += p;
  return i < n ? i : p - 1 - i;
}

// scipy.ndimage.gaussian_filter of the n2 x n2 window (mode 'reflect'): axis 0
// then axis 1 through reflect-padded copies in P.  The window value at index
// cidx is replaced by sup (origin suppression).  Ends synchronised.
ZR_DEV void gauss_window(const float* __restrict__ win, int cidx, float sup, int n2, const float* wy, int ry,
                         const float* wx, int rx, float* P, float* Cf, int tid) {
  const int nn = n2 * n2;
  const int prows = n2 + 2 * ry;
  for (int e = tid; e < prows * n2; e += 256) {
    const int i = e / n2 - ry, j = e - (e / n2) * n2;
    const int src = reflect_idx(i, n2) * n2 + j;
    P[e] = src == cidx ? sup : win[src];
  }
  __syncthreads();
  for (int e = tid; e < nn; e += 256) {
    const int i = e / n2, j = e - i * n2;
    const float* col = P + i * n2 + j;
    float acc = 0.f;
    for (int k = 0; k <= 2 * ry; ++k) acc += wy[k] * col[k * n2];
    Cf[e] = acc;
  }
  __syncthreads();
  const int pc = n2 + 2 * rx;
  for (int e = tid; e < n2 * pc; e += 256) {
    const int i = e / pc, j = e - i * pc - rx;
    P[e] = Cf[i * n2 + reflect_idx(j, n2)];
  }
  __syncthreads();
  for (int e = tid; e < nn; e += 256) {
    const int i = e / n2, j = e - i * n2;
    const float* row = P + i * pc + j;
    float acc = 0.f;
    for (int k = 0; k <= 2 * rx; ++k) acc += wx[k] * row[k];
    Cf[e] = acc;
  }
  __syncthreads();
}

// Bmode 'opti' (zorro/zorro.py:1105-1124): the Gaussian width sigma in [0, 7]
// minimising std(Cf) / (max(Cf) - mean(Cf)) over the whole window, found by
// bounded Brent minimisation with maxiter 15 and xatol 0.01.  The steps are
// those of Brent's fminbound (Brent 1973, ch. 5; the form scipy's
// minimize_scalar(method='bounded') implements): golden-section steps,
// parabolic steps when acceptable, tolerance sqrt(2.2e-16)|x| + xatol/3.
// All threads run the scalar logic redundantly; each objective evaluation is
// a block-wide filter + FP64 reductions.  Returns the chosen sigma with Cf
// holding the window filtered by it.
constexpr int OPTI_RMAX = 28;        // int(4 * 7 + 0.5)
ZR_DEV double opti_objective(const float* __restrict__ win, int cidx, float sup, int n2, double sigma, float* wts,
                             float* P, float* Cf, double* red, int tid) {
  const int nn = n2 * n2;
  __syncthreads();   // previous users of wts / Cf are done
  if (sigma > 1e-15) {
    const int r = (int)(4.0 * sigma + 0.5);
    if (tid <= 2 * r) {
      double sum = 0.0;
      for (int k = -r; k <= r; ++k) sum += exp(-0.5 / (sigma * sigma) * (double)(k * k));
      const int k = tid - r;
      wts[tid] = (float)(exp(-0.5 / (sigma * sigma) * (double)(k * k)) / sum);
    }
    __syncthreads();
    gauss_window(win, cidx, sup, n2, wts, r, wts, r, P, Cf, tid);
  } else {
    for (int e = tid; e < nn; e += 256) Cf[e] = e == cidx ? sup : win[e];
    __syncthreads();
  }
  double s1 = 0.0, mx = -INFINITY;
  for (int e = tid; e < nn; e += 256) {
    const double v = Cf[e];
    s1 += v;
    mx = fmax(mx, v);
  }
  s1 = wave_sum_d(s1);
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  if ((tid & 63) == 0) { red[tid >> 6] = s1; red[4 + (tid >> 6)] = mx; }
  __syncthreads();
  const double mean = (red[0] + red[1] + red[2] + red[3]) / nn;
  const double vmax = fmax(fmax(red[4], red[5]), fmax(red[6], red[7]));
  double s2 = 0.0;
  for (int e = tid; e < nn; e += 256) {
    const double d = (double)Cf[e] - mean;
    s2 += d * d;
  }
  s2 = wave_sum_d(s2);
  __syncthreads();
  if ((tid & 63) == 0) red[8 + (tid >> 6)] = s2;
  __syncthreads();
  const double sd = sqrt((red[8] + red[9] + red[10] + red[11]) / nn);
  return sd / (vmax - mean);
}

ZR_DEV double opti_sigma(const float* __restrict__ win, int cidx, float sup, int n2, float* wts, float* P, float* Cf,
                         double* red, int tid) {
  const double lo = 0.0, hi = 7.0, xatol = 0.01;
  const int maxfun = 15;
  const double sqrt_eps = sqrt(2.2e-16);
  const double golden_mean = 0.5 * (3.0 - sqrt(5.0));
  double a = lo, b = hi;
  double fulc = a + golden_mean * (b - a);
  double nfc = fulc, xf = fulc;
  double rat = 0.0, e = 0.0;
  double x = xf;
  double fx = opti_objective(win, cidx, sup, n2, x, wts, P, Cf, red, tid);
  int num = 1;
  double ffulc = fx, fnfc = fx;
  double xm = 0.5 * (a + b);
  double tol1 = sqrt_eps * fabs(xf) + xatol / 3.0;
  double tol2 = 2.0 * tol1;
  while (fabs(xf - xm) > (tol2 - 0.5 * (b - a))) {
    bool golden = true;
    if (fabs(e) > tol1) {
      golden = false;
      double r = (xf - nfc) * (fx - ffulc);
      double q = (xf - fulc) * (fx - fnfc);
      double p = (xf - fulc) * q - (xf - nfc) * r;
      q = 2.0 * (q - r);
      if (q > 0.0) p = -p;
      q = fabs(q);
      r = e;
      e = rat;
      if ((fabs(p) < fabs(0.5 * q * r)) && (p > q * (a - xf)) && (p < q * (b - xf))) {
        rat = (p + 0.0) / q;
        x = xf + rat;
        if (((x - a) < tol2) || ((b - x) < tol2)) {
          const double dxm = xm - xf;
          const double si = (dxm > 0.0) - (dxm < 0.0) + (dxm == 0.0);
          rat = tol1 * si;
        }
      } else {
        golden = true;
      }
    }
    if (golden) {
      e = (xf >= xm) ? a - xf : b - xf;
      rat = golden_mean * e;
    }
    const double si = (rat > 0.0) - (rat < 0.0) + (rat == 0.0);
    x = xf + si * fmax(fabs(rat), tol1);
    const double fu = opti_objective(win, cidx, sup, n2, x, wts, P, Cf, red, tid);
    ++num;
    if (fu <= fx) {
      if (x >= xf) a = xf; else b = xf;
      fulc = nfc; ffulc = fnfc;
      nfc = xf; fnfc = fx;
      xf = x; fx = fu;
    } else {
      if (x < xf) a = x; else b = x;
      if ((fu <= fnfc) || (nfc == xf)) {
        fulc = nfc; ffulc = fnfc;
        nfc = x; fnfc = fu;
      } else if ((fu <= ffulc) || (fulc == xf) || (fulc == nfc)) {
        fulc = x; ffulc = fu;
      }
    }
    xm = 0.5 * (a + b);
    tol1 = sqrt_eps * fabs(xf) + xatol / 3.0;
    tol2 = 2.0 * tol1;
    if (num >= maxfun) break;
  }
  // leave Cf filtered with the chosen width
  opti_objective(win, cidx, sup, n2, xf, wts, P, Cf, red, tid);
  return xf;
}

// One block (256 threads) per pair.
__global__ __launch_bounds__(256) void k_peaks(const float* __restrict__ windows, const int4* __restrict__ pairs,
                                              PeakParams pp, const float* __restrict__ gy,
                                              const float* __restrict__ gx, const float2* __restrict__ ktab,
                                              float* __restrict__ shifts, float* __restrict__ corr,
                                              float* __restrict__ peaksig, float* __restrict__ origin,
                                              float* __restrict__ bsigma, const int2* __restrict__ ctr,
                                              float* __restrict__ pshift, int* __restrict__ pflag) {
  // LDS: [tmp: max(2m*2m, reflect-padded window, 34*16*subpix)] [Cf: 2m*2m]
  // preShift: the window of pair p covers lags [c - m, c + m), c = ctr[p] (the
  // rolled disc's centre, zorro/zorro.py:1148-1163); lag 0 sits at (m - cy,
  // m - cx) when inside, and its suppression stencil is applied only when the
  // stencil is inside too (elsewhere it cannot reach the disc through the
  // Gaussian: the window margin is maxShift + 8 + radius).  The origin value
  // comes from the lag-0 column sums (k_origin_sums).
  extern __shared__ float smem[];
  const int m = pp.m, n2 = 2 * pp.m, nn = n2 * n2;
  float* tmp = smem;          // padded filter input, then sub-pixel scratch
  float* Cf = smem + pp.tmp_floats;   // filtered window
  __shared__ double rs[2][4];
  __shared__ float rv[4];
  __shared__ int ri[4];
  __shared__ float s_sup;
  const int tid = threadIdx.x;
  const int4 pr = pairs[blockIdx.x];
  const float* win = windows + (size_t)pr.z * nn;
  const int2 cc = pp.preshift ? ctr[pr.z] : make_int2(0, 0);
  const int li = m - cc.x, lj = m - cc.y;       // window index of lag 0
  const bool lag0 = li >= 1 && li <= n2 - 2 && lj >= 1 && lj <= n2 - 2;
  const int cidx = lag0 ? li * n2 + lj : -1;   // zero lag (-1: outside, no substitution)
  if (tid == 0) {
    float v = lag0 ? win[cidx] : 0.f;
    if (!pp.preshift) origin[pr.x * pp.n + pr.y] = v;
    if (pp.suppress && lag0) {
      // lag (ly,lx) -> index (li+ly)*2m + (lj+lx); stencil with duplicated terms
      v = 0.125f * (win[(li + 1) * n2 + lj] + win[li * n2 + lj + 1] + win[(li - 1) * n2 + lj] +
                    win[(li - 1) * n2 + lj] + win[(li + 1) * n2 + lj + 1] + win[(li - 1) * n2 + lj + 1] +
                    win[(li - 1) * n2 + lj + 1] + win[(li - 1) * n2 + lj - 1]);
    }
    s_sup = v;
  }
  __syncthreads();
  const float sup = s_sup;   // window value at zero lag after suppression
  if (pp.bopti) {
    __shared__ float wts[2 * OPTI_RMAX + 1];
    __shared__ double ored[12];
    const double sg = opti_sigma(win, cidx, sup, n2, wts, tmp, Cf, ored, tid);
    if (tid == 0 && bsigma) bsigma[pr.x * pp.n + pr.y] = (float)sg;
  } else if (pp.bconv) {
    gauss_window(win, cidx, sup, n2, gy, pp.ry, gx, pp.rx, tmp, Cf, tid);
  } else {
    for (int e = tid; e < nn; e += 256) Cf[e] = e == cidx ? sup : win[e];
  }
  __syncthreads();
  // disc statistics and argmax of Cf*disc (zeros outside take part; first index wins ties)
  double s1 = 0.0, s2 = 0.0, cnt = 0.0;
  float bv = -INFINITY;
  int bi = 0x7fffffff;
  const long m2 = (long)pp.disc * pp.disc;
  for (int e = tid; e < nn; e += 256) {
    int i = e / n2, j = e % n2;
    long yy = i - m, xx = j - m;
    bool in = (yy * yy + xx * xx) < m2;
    float v = in ? Cf[e] : 0.f;
    if (in) { s1 += v; cnt += 1.0; }
    if (v > bv || (v == bv && e < bi)) { bv = v; bi = e; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    float ov = __shfl_xor(bv, o, 64);
    int oi = __shfl_xor(bi, o, 64);
    if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    s1 += __shfl_xor(s1, o, 64);
    cnt += __shfl_xor(cnt, o, 64);
  }
  if ((tid & 63) == 0) { rv[tid >> 6] = bv; ri[tid >> 6] = bi; rs[0][tid >> 6] = s1; rs[1][tid >> 6] = cnt; }
  __syncthreads();
  double mean, total_cnt;
  {
    float v = rv[0]; int ix = ri[0];
    double a = rs[0][0], c = rs[1][0];
    for (int k = 1; k < 4; ++k) {
      if (rv[k] > v || (rv[k] == v && ri[k] < ix)) { v = rv[k]; ix = ri[k]; }
      a += rs[0][k]; c += rs[1][k];
    }
    bv = v; bi = ix; mean = a / c; total_cnt = c;
  }
  __syncthreads();
  for (int e = tid; e < nn; e += 256) {
    int i = e / n2, j = e % n2;
    long yy = i - m, xx = j - m;
    if ((yy * yy + xx * xx) < m2) { double d = (double)Cf[e] - mean; s2 += d * d; }
  }
  s2 = wave_sum_d(s2);
  if ((tid & 63) == 0) rs[0][tid >> 6] = s2;
  __syncthreads();
  const double var = (rs[0][0] + rs[0][1] + rs[0][2] + rs[0][3]) / total_cnt;
  const int cy = bi / n2, cx = bi % n2;
  const float psig = (float)(((double)bv - mean) / sqrt(var));
  const int I = pr.x, J = pr.y, n = pp.n;
  // preShift: with no positive value in the disc the reference's argmax leaves
  // the disc (first zero of the whole map), which no window holds: flagged
  if (pp.preshift && tid == 0 && !(bv > 0.f)) atomicOr(pflag, 1);
  float sh_y = (float)(cy - m + cc.x), sh_x = (float)(cx - m + cc.y), cval = bi == cidx ? sup : win[bi];
  // sub-pixel: python-slice semantics of C[c-8:c+8]
  bool valid = false;
  int sy = cy - 8, sx = cx - 8;
  if (pp.subpix > 1) {
    auto slice_ok = [&](int c0, int& s) {
      int st = c0 - 8, en = c0 + 8;
      if (st < 0) { st += n2; if (st < 0) st = 0; }
      en = en < n2 ? en : n2;
      s = st;
      return (en - st) == 16;
    };
    valid = slice_ok(cy, sy) & slice_ok(cx, sx);
  }
  if (valid) {
    const int P = pp.subpix, G = 16 * P;
    float2* Tm = reinterpret_cast<float2*>(tmp);     // [16][G]
    float2* Ks = reinterpret_cast<float2*>(tmp + (size_t)32 * G);   // K table copy [G]
    __syncthreads();
    for (int e = tid; e < G; e += 256) Ks[e] = ktab[e];
    __syncthreads();
    // T[p][t] = sum_q c[p][q] K(t - P q)   (K is G-periodic); with t = P a + b
    // the table index is P ((a - q) mod 16) + b: one (row p, residue b) per
    // thread forms the 16 outputs a from registers
    for (int e = tid; e < 16 * P; e += 256) {
      const int pr_ = e / P, b = e % P;
      float cr[16];
      float2 kb[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) cr[q] = Cf[(sy + pr_) * n2 + sx + q];
#pragma unroll
      for (int j = 0; j < 16; ++j) kb[j] = Ks[P * j + b];
#pragma unroll
      for (int a = 0; a < 16; ++a) {
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const float2 kv = kb[(a - q) & 15];
          acc.x += cr[q] * kv.x;
          acc.y += cr[q] * kv.y;
        }
        Tm[pr_ * G + P * a + b] = acc;
      }
    }
    __syncthreads();
    // Y[s][t] = sum_p K(s - P p) T[p][t];  argmax |Y|^2, first in C order.
    // With s = P a + b (a < 16, b < P) the table index is P ((a - q) mod 16) + b:
    // the 16 entries of residue b sit in registers while its 16 rows a are formed
    float best = -1.f;
    int bidx = 0x7fffffff;
    for (int t = tid; t < G; t += 256) {
      float2 tc[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) tc[q] = Tm[q * G + t];
      for (int b = 0; b < P; ++b) {
        float2 kb[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) kb[j] = Ks[P * j + b];
#pragma unroll
        for (int a = 0; a < 16; ++a) {
          float2 acc = make_float2(0.f, 0.f);
#pragma unroll
          for (int q = 0; q < 16; ++q) {
            const float2 kv = kb[(a - q) & 15];
            acc.x += kv.x * tc[q].x - kv.y * tc[q].y;
            acc.y += kv.x * tc[q].y + kv.y * tc[q].x;
          }
          const float a2 = acc.x * acc.x + acc.y * acc.y;
          const int id = (P * a + b) * G + t;
          if (a2 > best || (a2 == best && id < bidx)) { best = a2; bidx = id; }
        }
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      float ov = __shfl_xor(best, o, 64);
      int oi = __shfl_xor(bidx, o, 64);
      if (ov > best || (ov == best && oi < bidx)) { best = ov; bidx = oi; }
    }
    __syncthreads();
    if ((tid & 63) == 0) { rv[tid >> 6] = best; ri[tid >> 6] = bidx; }
    __syncthreads();
    if (tid == 0) {
      float v = rv[0]; int ix = ri[0];
      for (int k = 1; k < 4; ++k)
        if (rv[k] > v || (rv[k] == v && ri[k] < ix)) { v = rv[k]; ix = ri[k]; }
      int smy = ix / G, smx = ix % G;
      float rpy = (float)(cy - m + cc.x), rpx = (float)(cx - m + cc.y);
      float sgy = (rpy > 0.f) - (rpy < 0.f), sgx = (rpx > 0.f) - (rpx < 0.f);
      sh_y = rpy + ((float)smy - (float)(G / 2) + sgy) / (float)P;
      sh_x = rpx + ((float)smx - (float)(G / 2) + sgx) / (float)P;
      cval = sqrtf(v) / 256.f;
    }
  }
  if (tid == 0) {
    shifts[((size_t)I * n + J) * 2 + 0] = sh_y * pp.scale_y;
    shifts[((size_t)I * n + J) * 2 + 1] = sh_x * pp.scale_x;
    corr[I * n + J] = cval;
    peaksig[I * n + J] = psig;
    if (pshift) {                             // cropped pixels, for the next pair's disc
      pshift[2 * pr.z] = sh_y;
      pshift[2 * pr.z + 1] = sh_x;
    }
  }
}

// preShift (zorro/zorro.py:1148-1153): the disc of pair (I, J) is centred on
// round(shiftsTriMat[I, J-1]) in cropped pixels (banker's rounding, np.round);
// a pair whose previous one is the diagonal or was not correlated uses 0.
__global__ __launch_bounds__(64) void k_preshift_centres(const int4* __restrict__ dpairs, int cnt,
                                                         const int* __restrict__ pidx, const float* __restrict__ pshift,
                                                         int N, int2* __restrict__ ctr) {
  for (int i = blockIdx.x * 64 + threadIdx.x; i < cnt; i += gridDim.x * 64) {
    const int4 pr = dpairs[i];
    const int q = pr.y - 1 > pr.x ? pidx[pr.x * N + pr.y - 1] : -1;
    ctr[pr.z] = q >= 0 ? make_int2((int)rintf(pshift[2 * q]), (int)rintf(pshift[2 * q + 1])) : make_int2(0, 0);
  }
}

// originTriMat under preShift: C at lag 0 = nc2 sum_kx s0[p][kx] / max(1, D_p)
// (the column sums of the pair transform), one block per pair, fixed order.
__global__ __launch_bounds__(256) void k_origin_sums(const int4* __restrict__ pairs, const float* __restrict__ s0,
                                                     int Wh, const double* __restrict__ den, double nc2, int N,
                                                     float* __restrict__ origin) {
  __shared__ double red[4];
  const int4 pr = pairs[blockIdx.x];
  double a = 0.0;
  for (int k = threadIdx.x; k < Wh; k += 256) a += (double)s0[(size_t)pr.z * Wh + k];
  a = wave_sum_d(a);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = (red[0] + red[1]) + (red[2] + red[3]);
    origin[pr.x * N + pr.y] = (float)(nc2 * t / fmax(1.0, den[pr.z]));
  }
}

// triMode 'auto' (zorro/zorro.py:992-996): row I visits J = I+1, I+2, ... and
// stops after the first pair whose peak significance is <= peaksigThres or whose
// J - I reaches autoMax.  The pairs after the stop are never correlated in the
// reference, so their trimat entries return to zero.  One thread per row.
__global__ __launch_bounds__(64) void k_auto_prune(float* __restrict__ shifts, float* __restrict__ corr,
                                                   float* __restrict__ peaksig, float* __restrict__ origin, int N,
                                                   int auto_max, double thres) {
  for (int I = threadIdx.x; I < N; I += 64) {
    bool stopped = false;
    for (int J = I + 1; J < N; ++J) {
      const size_t e = (size_t)I * N + J;
      if (stopped) {
        shifts[2 * e] = 0.f;
        shifts[2 * e + 1] = 0.f;
        corr[e] = 0.f;
        peaksig[e] = 0.f;
        origin[e] = 0.f;
      } else if ((double)peaksig[e] <= thres || J - I >= auto_max) {
        stopped = true;
      }
    }
  }
}

// Zero fill (a kernel rather than hipMemsetAsync, so that captured graphs hold
// only kernel nodes); bytes must be a multiple of 4.
__global__ __launch_bounds__(256) void k_fill_zero(uint32_t* __restrict__ p, size_t words) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < words; i += (size_t)gridDim.x * 256) p[i] = 0u;
}

// ----------------------------------------------------------- launchers
// Host-side launch helpers (instantiated for the supported sizes).
#include "zr_launch.h"

#include <algorithm>

// rows per k1_rows block for a padded width (also the granularity of the
// frame-sum partials); at 5760 two rows (49 KB of LDS, three blocks per CU)
// beat four (98 KB, one block per CU)
constexpr int k1_rows_per_block(int wp) { return wp == 5760 ? 2 : (wp >= 3072 ? 4 : 8); }
int zr_k1_rows_per_block(int Wp) { return k1_rows_per_block(Wp); }
// threads per row: the half-length L = WP/2 must split evenly over them
constexpr int k1_threads_per_row(int wp) {
  return (wp >= 2048 && (wp / 2) % 128 == 0) ? 128 : ((wp / 2) % 192 == 0 ? 192 : 64);
}

template <int WP>
static hipError_t launch_k1_rows_t(const float* frames, int H, int W, int Hp, int Wh, int frame0, int nb,
                                   const float2* twL, const float2* tw2L, float2* T, double* part, int mode,
                                   const double* stats, const float* umask, hipStream_t st) {
  constexpr int RB = k1_rows_per_block(WP);
  constexpr int TPR = k1_threads_per_row(WP);
  if (Hp % RB != 0) return hipErrorInvalidValue;
  int grid = nb * (Hp / RB);
  size_t sh = ((size_t)RB * zr_fft_lds(WP / 2) + zr_ldstw_size(WP / 2) + zr_ldstw_size(WP)) * sizeof(float2);
  static std::once_flag lds_once;
  zr_allow_full_lds((const void*)k1_rows<WP, RB, TPR>, lds_once);
  hipLaunchKernelGGL((k1_rows<WP, RB, TPR>), dim3(grid), dim3(RB * TPR), sh, st, frames, H, W, Hp, Wh, frame0, twL,
                     tw2L, T, part, mode, stats, mode == 0 ? nullptr : umask);
  return hipGetLastError();
}

hipError_t zr_launch_k1_rows(int Wp, const float* frames, int H, int W, int Hp, int Wh, int frame0, int nb,
                             const float2* twL, const float2* tw2L, float2* T, double* part, int mode,
                             const double* stats, const float* umask, int* rb_out, hipStream_t st) {
  switch (Wp) {
#define ZR_CASE(X) case X: *rb_out = k1_rows_per_block(X); return launch_k1_rows_t<X>(frames, H, W, Hp, Wh, frame0, nb, twL, tw2L, T, part, mode, stats, umask, st);
    ZR_CASE(256) ZR_CASE(512) ZR_CASE(1024) ZR_CASE(2048) ZR_CASE(3072) ZR_CASE(3840) ZR_CASE(4096) ZR_CASE(5760)
#undef ZR_CASE
    default: return hipErrorInvalidValue;
  }
}

constexpr bool K1_COLS_PAIRED = false;  // column pairs per block (else one column per block, the faster)
// threads per column of k1_cols1 (timed on MI355X: 128 at heights up to 2048,
// 256 at 3072 to 4096, 512 at 5760).  At 4096, 128 threads are as fast alone,
// but 256 (about 70 VGPRs instead of 135) let the column blocks of one movie
// share CUs with another movie's pair transform (two waves of about 200
// VGPRs per SIMD), which is what two movies in flight need.
constexpr int k1_cols_threads(int hp) { return hp == 5760 ? 512 : (hp >= 3072 ? 256 : 128); }
hipError_t zr_launch_k1_cols(int Hp, const float2* T, int Wh, int Hc, int frame0, int nb, const float2* twH,
                             float2* S, float2* Sside, int full, hipStream_t st) {
  if (!K1_COLS_PAIRED) {
    switch (Hp) {
#define ZR_CASE(X) case X: hipLaunchKernelGGL((k1_cols1<X, k1_cols_threads(X)>), dim3(nb * Wh), dim3(k1_cols_threads(X)), 0, st, T, Wh, Hc, frame0, twH, S, Sside, full); return hipGetLastError();
      ZR_CASE(256) ZR_CASE(512) ZR_CASE(1024) ZR_CASE(2048) ZR_CASE(3072) ZR_CASE(3840) ZR_CASE(4096) ZR_CASE(5760)
#undef ZR_CASE
      default: return hipErrorInvalidValue;
    }
  }
  const int grid = nb * ((Wh + 1) / 2);
  switch (Hp) {
#define ZR_CASE(X) case X: { constexpr int NTC = X >= 2048 ? 512 : 256; static std::once_flag lds_once; \
      zr_allow_full_lds((const void*)k1_cols<X, NTC>, lds_once); \
      hipLaunchKernelGGL((k1_cols<X, NTC>), dim3(grid), dim3(2 * NTC), 2 * zr_fft_lds(X) * sizeof(float2), st, T, Wh, Hc, frame0, twH, S, Sside, full); return hipGetLastError(); }
    ZR_CASE(256) ZR_CASE(512) ZR_CASE(1024) ZR_CASE(2048) ZR_CASE(3072) ZR_CASE(3840) ZR_CASE(4096) ZR_CASE(5760)
#undef ZR_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t zr_launch_k1_sums(const float* frames, int n, int H, int W, double* part, int* nrb_out, hipStream_t st) {
  const int nrb = (H + K1S_RB - 1) / K1S_RB;
  *nrb_out = nrb;
  hipLaunchKernelGGL(k1_frame_sums, dim3(n * nrb), dim3(256), 0, st, frames, H, W, part);
  return hipGetLastError();
}

hipError_t zr_launch_mnxc_combine(float* windows, const int4* pairs, int P, int N, int m, hipStream_t st) {
  hipLaunchKernelGGL(k_mnxc_combine, dim3(P), dim3(256), 0, st, windows, pairs, P, N, 4 * m * m);
  return hipGetLastError();
}

hipError_t zr_launch_k1_stats(const double* part, int nrb, int n, double* stats, hipStream_t st) {
  hipLaunchKernelGGL(k1_frame_stats, dim3(n), dim3(256), 0, st, part, nrb, n, stats);
  return hipGetLastError();
}

static constexpr int K2C_WPB = 2;
hipError_t zr_launch_k2_cols(int Lper, const float2* S, const float2* Sside, const int4* work, int npairs, int Wh,
                             int Hc, int Wc, int m, int cols_per_wave, int dc_zero, const float2* twHc, float2* Z,
                             int Wk, int Hs, int hoff, int side_in_col, const int2* ctr, const float2* twWc,
                             float* s0, hipStream_t st) {
  int cols_per_block = K2C_WPB * cols_per_wave;
  int ngroups = (Wh + cols_per_block - 1) / cols_per_block;
  int grid = ngroups * npairs;
  switch (Lper) {
#define ZR_CASE(X) case X: { size_t sh = ((size_t)K2C_WPB * (X / 2) * 65 + (size_t)2 * 2 * m * 8) * sizeof(float2); \
      static std::once_flag lds_once, lds_once_ps; \
      if (ctr) { zr_allow_full_lds((const void*)k2_cols<X, K2C_WPB, true>, lds_once_ps); \
        hipLaunchKernelGGL((k2_cols<X, K2C_WPB, true>), dim3(grid), dim3(64 * K2C_WPB), sh, st, S, Sside, work, npairs, Wh, Hc, Wc, m, cols_per_wave, dc_zero, twHc, Z, Wk, Hs, hoff, side_in_col, ctr, twWc, s0); } \
      else { zr_allow_full_lds((const void*)k2_cols<X, K2C_WPB, false>, lds_once); \
        hipLaunchKernelGGL((k2_cols<X, K2C_WPB, false>), dim3(grid), dim3(64 * K2C_WPB), sh, st, S, Sside, work, npairs, Wh, Hc, Wc, m, cols_per_wave, dc_zero, twHc, Z, Wk, Hs, hoff, side_in_col, ctr, twWc, s0); } \
      return hipGetLastError(); }
    ZR_CASE(4) ZR_CASE(8) ZR_CASE(16) ZR_CASE(32) ZR_CASE(48) ZR_CASE(64)
#undef ZR_CASE
    default: return hipErrorInvalidValue;
  }
}

bool zr_k2_cols_r2_ok(int Lper, int m) { return (Lper == 32 || Lper == 48 || Lper == 64) && m >= 1 && m <= Lper; }

hipError_t zr_launch_k2_lane_twiddles(int Lper, int Hc, float2* ltw, hipStream_t st) {
  const int n = (Lper + 1) * 64;
  hipLaunchKernelGGL(k2_lane_twiddles, dim3((n + 255) / 256), dim3(256), 0, st, Lper, Hc, ltw);
  return hipGetLastError();
}

hipError_t zr_launch_k2_cols_r2(int Lper, const float2* S, const float2* Sside, const int4* work, int npairs, int Wh,
                                int Hc, int Wc, int m, int cols_per_wave, int pairs_per_wave, int dc_zero,
                                const float2* ltw, float2* Z, int Wk, int Hs, int hoff, int side_in_col,
                                const int2* ctr, const float2* twHc, const float2* twWc, float* s0, hipStream_t st) {
  if (!zr_k2_cols_r2_ok(Lper, m) || Hc != 64 * Lper || npairs < 1) return hipErrorInvalidValue;
  if (ctr) pairs_per_wave = 0;                    // the window centres are per pair: column-group mode
  if (pairs_per_wave <= 0 && cols_per_wave <= 0) return hipErrorInvalidValue;
  int grid;
  if (pairs_per_wave > 0) {
    const int per_block = K2R2_WPB * pairs_per_wave;
    grid = Wh * ((npairs + per_block - 1) / per_block);
  } else {
    const int cols_per_block = K2R2_WPB * cols_per_wave;
    grid = ((Wh + cols_per_block - 1) / cols_per_block) * npairs;
  }
  const size_t sh = k2r2_lds_bytes(Lper);
  switch (Lper) {
#define ZR_CASE(X) case X: { static std::once_flag lds_once, lds_once_ps; \
      if (ctr) { zr_allow_full_lds((const void*)k2_cols_r2<X, K2R2_WPB, true>, lds_once_ps); \
        hipLaunchKernelGGL((k2_cols_r2<X, K2R2_WPB, true>), dim3(grid), dim3(64 * K2R2_WPB), sh, st, S, Sside, work, npairs, Wh, Hc, Wc, m, cols_per_wave, pairs_per_wave, dc_zero, ltw, Z, Wk, Hs, hoff, side_in_col, ctr, twHc, twWc, s0); } \
      else { zr_allow_full_lds((const void*)k2_cols_r2<X, K2R2_WPB, false>, lds_once); \
        hipLaunchKernelGGL((k2_cols_r2<X, K2R2_WPB, false>), dim3(grid), dim3(64 * K2R2_WPB), sh, st, S, Sside, work, npairs, Wh, Hc, Wc, m, cols_per_wave, pairs_per_wave, dc_zero, ltw, Z, Wk, Hs, hoff, side_in_col, ctr, twHc, twWc, s0); } \
      return hipGetLastError(); }
    ZR_CASE(32) ZR_CASE(48) ZR_CASE(64)
#undef ZR_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t zr_launch_k2_rows(const float2* Z, const int4* pairs, int npairs, int Wh, int Wc, int m,
                             const float2* twWc, const float2* btab, const double* den, float nc2, float* windows,
                             const int4* sub, hipStream_t st) {
  int twoM = 2 * m;
  if (npairs <= 0) return hipSuccess;
  if (btab && twoM % 16 == 0 && twoM <= 128) {
    switch (twoM / 16) {
#define ZR_MF(T) case T: hipLaunchKernelGGL(k2_rows_mfma<T>, dim3((npairs * T + K2M_WPB - 1) / K2M_WPB), dim3(64 * K2M_WPB), 0, st, Z, Wh, npairs * T, btab, den, nc2, windows, sub); return hipGetLastError();
      ZR_MF(1) ZR_MF(2) ZR_MF(3) ZR_MF(4) ZR_MF(5) ZR_MF(6) ZR_MF(7) ZR_MF(8)
#undef ZR_MF
      default: break;
    }
  }
  int ndxg = (twoM + K2R_DX - 1) / K2R_DX, ndyg = (twoM + K2R_DY - 1) / K2R_DY;
  int nt = ndxg * ndyg;
  if (nt > 1024) return hipErrorInvalidValue;
  nt = ((nt + 63) / 64) * 64;
  size_t sh = (size_t)K2R_KC * twoM * sizeof(float2);
  hipLaunchKernelGGL(k2_rows, dim3(npairs), dim3(nt), sh, st, Z, pairs, Wh, Wc, m, twWc, den, nc2, windows, sub);
  return hipGetLastError();
}

hipError_t zr_launch_k2_den(const int4* pairs, int npairs, const double* stats, double npix, double nc2, double* den,
                            hipStream_t st) {
  hipLaunchKernelGGL(k2_denominators, dim3((npairs + 127) / 128), dim3(128), 0, st, pairs, npairs, stats, npix, nc2,
                     den);
  return hipGetLastError();
}

hipError_t zr_launch_peaks(const float* windows, const int4* pairs, int npairs, const PeakParamsHost* hp,
                           const float* gy, const float* gx, const float2* ktab, float* shifts, float* corr,
                           float* peaksig, float* origin, float* bsigma, const PeakPreShift* ps, hipStream_t st) {
  PeakParams pp;
  pp.m = hp->m; pp.suppress = hp->suppress; pp.bconv = hp->bconv; pp.bopti = hp->bopti; pp.ry = hp->ry;
  pp.rx = hp->rx;
  pp.subpix = hp->subpix; pp.scale_y = hp->scale_y; pp.scale_x = hp->scale_x; pp.n = hp->n;
  pp.disc = hp->disc > 0 ? hp->disc : hp->m;
  pp.preshift = ps ? 1 : 0;
  const size_t sh = zr_peaks_lds_bytes(hp, &pp.tmp_floats);
  if (sh > (size_t)ZR_LDS_BYTES) return hipErrorInvalidConfiguration;
  static std::once_flag lds_once;
  zr_allow_full_lds((const void*)k_peaks, lds_once);
  hipLaunchKernelGGL(k_peaks, dim3(npairs), dim3(256), sh, st, windows, pairs, pp, gy, gx, ktab, shifts, corr,
                     peaksig, origin, bsigma, ps ? ps->ctr : nullptr, ps ? ps->pshift : nullptr,
                     ps ? ps->flag : nullptr);
  return hipGetLastError();
}

hipError_t zr_launch_preshift_centres(const int4* dpairs, int cnt, const int* pidx, const float* pshift, int N,
                                      int2* ctr, hipStream_t st) {
  hipLaunchKernelGGL(k_preshift_centres, dim3((cnt + 63) / 64), dim3(64), 0, st, dpairs, cnt, pidx, pshift, N, ctr);
  return hipGetLastError();
}

hipError_t zr_launch_origin_sums(const int4* pairs, int npairs, const float* s0, int Wh, const double* den, double nc2,
                                 int N, float* origin, hipStream_t st) {
  hipLaunchKernelGGL(k_origin_sums, dim3(npairs), dim3(256), 0, st, pairs, s0, Wh, den, nc2, N, origin);
  return hipGetLastError();
}

hipError_t zr_launch_auto_prune(float* shifts, float* corr, float* peaksig, float* origin, int N, int auto_max,
                                double thres, hipStream_t st) {
  hipLaunchKernelGGL(k_auto_prune, dim3(1), dim3(64), 0, st, shifts, corr, peaksig, origin, N, auto_max, thres);
  return hipGetLastError();
}

hipError_t zr_launch_spec_sum(float2* S, size_t slot_elems, int slot0, int count, int out_slot, hipStream_t st) {
  hipLaunchKernelGGL(k_spec_sum, dim3((unsigned)((slot_elems + 255) / 256)), dim3(256), 0, st, S, slot_elems, slot0,
                     count, out_slot);
  return hipGetLastError();
}

hipError_t zr_launch_refine_den(const double* stats, int N, double npix, double nc2, double* den, hipStream_t st) {
  hipLaunchKernelGGL(k_refine_den, dim3(1), dim3(64), 0, st, stats, N, npix, nc2, den);
  return hipGetLastError();
}

hipError_t zr_launch_refine_combine(float* windows, int N, int m, int masked, int out0, const double* cm,
                                    double npix, hipStream_t st) {
  hipLaunchKernelGGL(k_refine_combine, dim3(N - 1), dim3(256), 0, st, windows, N, 4 * m * m, masked, out0, cm, npix);
  return hipGetLastError();
}

hipError_t zr_launch_common_mean(const double* stats, int N, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_common_mean, dim3(1), dim3(64), 0, st, stats, N, out);
  return hipGetLastError();
}

hipError_t zr_launch_fill_zero(void* p, size_t bytes, hipStream_t st) {
  const size_t words = bytes / 4;
  const size_t blocks = std::min<size_t>((words + 255) / 256, 4096);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, st, (uint32_t*)p, words);
  return hipGetLastError();
}

int zr_k2_btab_rows(int Wh) { return (Wh + 2 * K2M_KC - 1) / (2 * K2M_KC) * (2 * K2M_KC); }

hipError_t zr_launch_k2_twiddles(int Wh, int Wc, int m, float2* btab, hipStream_t st) {
  const int Whp = zr_k2_btab_rows(Wh);
  const int n = Whp * 2 * m;
  hipLaunchKernelGGL(k2_rows_twiddles, dim3((n + 255) / 256), dim3(256), 0, st, Wh, Whp, Wc, m, btab);
  return hipGetLastError();
}

// ---------------------------------------------------------------- batches
// A movie of a batch reads its block [off, off + n)^2 of the batch's
// triangular matrices (shifts [Nt][Nt][2], corr, peaksig, origin [Nt][Nt])
// into its own [n][n] ones.
__global__ __launch_bounds__(256) void k_gather_trimats(const float* __restrict__ shifts, const float* __restrict__ corr,
                                                        const float* __restrict__ peaksig,
                                                        const float* __restrict__ origin, int Nt, int off, int n,
                                                        float* __restrict__ d_shifts, float* __restrict__ d_corr,
                                                        float* __restrict__ d_peaksig, float* __restrict__ d_origin) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n * n) return;
  const int I = e / n, J = e % n;
  const size_t src = (size_t)(off + I) * Nt + (off + J);
  d_shifts[2 * e] = shifts[2 * src];
  d_shifts[2 * e + 1] = shifts[2 * src + 1];
  d_corr[e] = corr[src];
  d_peaksig[e] = peaksig[src];
  d_origin[e] = origin[src];
}

hipError_t zr_launch_gather_trimats(const float* shifts, const float* corr, const float* peaksig, const float* origin,
                                    int Nt, int off, int n, float* d_shifts, float* d_corr, float* d_peaksig,
                                    float* d_origin, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_trimats, dim3((n * n + 255) / 256), dim3(256), 0, st, shifts, corr, peaksig, origin, Nt,
                     off, n, d_shifts, d_corr, d_peaksig, d_origin);
  return hipGetLastError();
}
