// Input stage of super-resolution counting movies, for gfx950:
//   gain reference product (loadData, zorro/zorro.py:3455-3482) fused with the
//   integer-to-float conversion, and Fourier binning (binStack 'fourier',
//   zorro/zorro.py:1866-1901): forward DFT at the full size, four-quadrant crop
//   to the binned size, 1/(Hb Wb), raw inverse DFT at the binned size, real part.
//
// Detector sizes are not 2^a 3^b 5^c (7676 = 4 19 101, 7420 = 4 5 7 53), so every
// DFT here is a Bluestein chirp-z transform (Bluestein 1970): for length L,
//   X[k] = c[k] sum_n (x[n] c[n]) b[k - n],  c[n] = e^{-i pi n^2 / L},
//   b[m] = e^{+i pi m^2 / L},
// the convolution done circularly with a power-of-two LDS FFT of size M >= 2L-1
// against the precomputed spectrum B = FFT_M(b) / M.  The inverse DFT is
// conj(DFT(conj(.))).  Phases come from n^2 mod 2L in integers (exact).
//
// Per batch of frames (row-major, complex64 as float2):
//   Tr  [B][H][Wb]   row pass: kept columns of the row DFT (cropped order)
//   Tc  [B][Wb][H]   its transpose
//   Gc  [B][Wb][Hb]  column pass: kept rows, 1/(Hb Wb), inverse DFT along y
//   Gr  [B][Hb][Wb]  its transpose
// then the inverse DFT along x gives the binned frame f32 [Hb][Wb].
#include "zr_common.h"
#include "zr_launch.h"

// the values of ZR_DTYPE_* in the public header
enum { ZR_RAW_F32 = 0, ZR_RAW_U8 = 1, ZR_RAW_U16 = 2, ZR_RAW_U4 = 3, ZR_RAW_I16 = 4 };

// Sample i (row-major over frames) of a raw stack.  U4: two samples per byte,
// the even one in the low nibble (frame width even, so rows need no padding).
ZR_DEV float raw_value(const void* raw, int dtype, size_t i) {
  if (dtype == ZR_RAW_U8) return (float)reinterpret_cast<const uint8_t*>(raw)[i];
  if (dtype == ZR_RAW_U16) return (float)reinterpret_cast<const uint16_t*>(raw)[i];
  if (dtype == ZR_RAW_U4) {
    const unsigned b = reinterpret_cast<const uint8_t*>(raw)[i >> 1];
    return (float)((i & 1) ? (b >> 4) : (b & 15u));
  }
  if (dtype == ZR_RAW_I16) return (float)reinterpret_cast<const int16_t*>(raw)[i];
  return reinterpret_cast<const float*>(raw)[i];
}

// Samples i and i+1 (i even) of a raw stack with one load: two bytes, two
// 16-bit words, two floats, or the two nibbles of one 4-bit byte.
ZR_DEV float2 raw_pair(const void* raw, int dtype, size_t i) {
  if (dtype == ZR_RAW_U8) {
    const unsigned v = reinterpret_cast<const uint16_t*>(raw)[i >> 1];
    return make_float2((float)(v & 255u), (float)(v >> 8));
  }
  if (dtype == ZR_RAW_U16) {
    const unsigned v = reinterpret_cast<const uint32_t*>(raw)[i >> 1];
    return make_float2((float)(v & 65535u), (float)(v >> 16));
  }
  if (dtype == ZR_RAW_U4) {
    const unsigned b = reinterpret_cast<const uint8_t*>(raw)[i >> 1];
    return make_float2((float)(b & 15u), (float)(b >> 4));
  }
  if (dtype == ZR_RAW_I16) {
    const int v = reinterpret_cast<const int*>(raw)[i >> 1];
    return make_float2((float)(short)(v & 65535), (float)(short)(v >> 16));
  }
  return reinterpret_cast<const float2*>(raw)[i >> 1];
}

// gain product in float32 as loadData does it: gainRef * data, then / gain
ZR_DEV float gained(float v, const float* gain, size_t pix, float gscale, int use_scale) {
  if (gain) v = gain[pix] * v;
  if (use_scale) v = v / gscale;
  return v;
}

// The chirp-z convolution on buf (LDS, M entries, holding x[n] c[n], zero past
// L): on return buf[k] * c[k] is the DFT for k < L.  Synchronised on entry and exit.
// The FFT twiddles come from an LDS table after buf (bin_twiddles).
template <int M, int NT>
ZR_DEV void bluestein_conv(float2* buf, const float2* __restrict__ B, const LdsTw& tws, int tid) {
  wg_fft<M, -1, NT>(buf, tws, tid);
#pragma unroll 4
  for (int e = tid; e < M; e += NT) buf[e] = c_mul(buf[e], B[e]);
  __syncthreads();
  wg_fft<M, 1, NT>(buf, tws, tid);
}

// The LDS twiddle table of the length-M transforms, placed after the M-point
// buffer; filled by the block before its first barrier.
template <int M, int NT>
ZR_DEV LdsTw bin_twiddles(float2* lds, const float2* __restrict__ twM, int tid) {
  return zr_ldstw_fill(lds + zr_fft_lds(M), twM, M, tid, NT);
}
constexpr size_t bin_lds_bytes(int M) { return (size_t)(zr_fft_lds(M) + zr_ldstw_size(M)) * sizeof(float2); }
// threads per block: at M = 16384 a 1024-thread block would cap a thread at 128
// VGPRs and spill; 512 threads hold two butterflies of a pass each in registers
constexpr int bin_threads(int M) { return M >= 16384 ? 512 : 1024; }

// Row pass: one block per (frame in batch, raw row y).
template <int M>
__global__ __launch_bounds__(bin_threads(M)) void k_bin_rows(const void* __restrict__ raw, int dtype, int frame0, int H, int W,
                                                   const float* __restrict__ gain, float gscale, int use_scale,
                                                   BinChirp cw, const float2* __restrict__ twM, int Wb,
                                                   float2* __restrict__ Tr) {
  constexpr int NT = bin_threads(M);
  extern __shared__ float2 bin_lds[];
  float2* buf = bin_lds;
  const int tid = threadIdx.x;
  const int fb = blockIdx.x / H, y = blockIdx.x % H;
  const size_t base = ((size_t)(frame0 + fb) * H + y) * W;
  const LdsTw tws = bin_twiddles<M, NT>(bin_lds, twM, tid);
#pragma unroll 4
  for (int x = tid; x < M; x += NT) {
    float2 v = make_float2(0.f, 0.f);
    if (x < W) {
      const float a = gained(raw_value(raw, dtype, base + x), gain, (size_t)y * W + x, gscale, use_scale);
      v = c_scale(cw.c[x], a);
    }
    buf[x] = v;
  }
  __syncthreads();
  bluestein_conv<M, NT>(buf, cw.B, tws, tid);
  float2* dst = Tr + ((size_t)fb * H + y) * Wb;
  const int h2 = Wb / 2;
#pragma unroll 4
  for (int k = tid; k < Wb; k += NT) {
    const int kx = k < h2 ? k : W - Wb + k;
    dst[k] = c_mul(cw.c[kx], buf[kx]);
  }
}

// 32x32 LDS-tiled transpose of B matrices [R][C] -> [C][R].
__global__ __launch_bounds__(256) void k_bin_transpose(const float2* __restrict__ in, int R, int C,
                                                       float2* __restrict__ out) {
  __shared__ float2 tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
  const int tilesC = (C + 31) / 32, tilesR = (R + 31) / 32;
  const int b = blockIdx.x / (tilesC * tilesR);
  const int t = blockIdx.x % (tilesC * tilesR);
  const int r0 = (t / tilesC) * 32, c0 = (t % tilesC) * 32;
  const float2* src = in + (size_t)b * R * C;
  float2* dst = out + (size_t)b * R * C;
  for (int i = ty; i < 32; i += 8) {
    const int r = r0 + i, c = c0 + tx;
    if (r < R && c < C) tile[i][tx] = src[(size_t)r * C + c];
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    const int c = c0 + i, r = r0 + tx;
    if (r < R && c < C) dst[(size_t)c * R + r] = tile[tx][i];
  }
}

// Column pass: one block per (frame in batch, kept column).  DFT along y (length
// H), keep Hb rows, scale, inverse DFT of length Hb.
template <int M>
__global__ __launch_bounds__(bin_threads(M)) void k_bin_cols(const float2* __restrict__ Tc, int H, int Wb, int Hb, BinChirp ch,
                                                   BinChirp chb, const float2* __restrict__ twM, float scale,
                                                   float2* __restrict__ Gc) {
  constexpr int NT = bin_threads(M);
  constexpr int PK = (M / 2 + NT - 1) / NT;    // kept rows per thread (Hb <= M/2)
  extern __shared__ float2 bin_lds[];
  float2* buf = bin_lds;
  const int tid = threadIdx.x;
  const size_t col = blockIdx.x;               // fb * Wb + kx'
  const float2* src = Tc + col * H;
  const LdsTw tws = bin_twiddles<M, NT>(bin_lds, twM, tid);
#pragma unroll 4
  for (int y = tid; y < M; y += NT) buf[y] = y < H ? c_mul(src[y], ch.c[y]) : make_float2(0.f, 0.f);
  __syncthreads();
  bluestein_conv<M, NT>(buf, ch.B, tws, tid);
  // kept rows in cropped order, conjugated and chirped for the inverse
  float2 keep[PK];
  const int h2 = Hb / 2;
#pragma unroll
  for (int i = 0; i < PK; ++i) {
    const int k = tid + i * NT;
    keep[i] = make_float2(0.f, 0.f);
    if (k < Hb) {
      const int ky = k < h2 ? k : H - Hb + k;
      const float2 X = c_scale(c_mul(ch.c[ky], buf[ky]), scale);
      keep[i] = c_mul(c_conj(X), chb.c[k]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PK; ++i) {
    const int k = tid + i * NT;
    if (k < Hb) buf[k] = keep[i];
  }
#pragma unroll 4
  for (int e = Hb + tid; e < M; e += NT) buf[e] = make_float2(0.f, 0.f);
  __syncthreads();
  bluestein_conv<M, NT>(buf, chb.B, tws, tid);
  float2* dst = Gc + col * Hb;
#pragma unroll 4
  for (int n = tid; n < Hb; n += NT) dst[n] = c_conj(c_mul(chb.c[n], buf[n]));
}

// Final rows: one block per (frame in batch, binned row).  Inverse DFT along x
// (length Wb), real part.
template <int M>
__global__ __launch_bounds__(bin_threads(M)) void k_bin_out(const float2* __restrict__ Gr, int Hb, int Wb, BinChirp cwb,
                                                  const float2* __restrict__ twM, int frame0,
                                                  float* __restrict__ frames) {
  constexpr int NT = bin_threads(M);
  extern __shared__ float2 bin_lds[];
  float2* buf = bin_lds;
  const int tid = threadIdx.x;
  const int fb = blockIdx.x / Hb, ny = blockIdx.x % Hb;
  const float2* src = Gr + ((size_t)fb * Hb + ny) * Wb;
  const LdsTw tws = bin_twiddles<M, NT>(bin_lds, twM, tid);
#pragma unroll 4
  for (int k = tid; k < M; k += NT) buf[k] = k < Wb ? c_mul(c_conj(src[k]), cwb.c[k]) : make_float2(0.f, 0.f);
  __syncthreads();
  bluestein_conv<M, NT>(buf, cwb.B, tws, tid);
  float* dst = frames + ((size_t)(frame0 + fb) * Hb + ny) * Wb;
#pragma unroll 4
  for (int n = tid; n < Wb; n += NT) dst[n] = c_mul(cwb.c[n], buf[n]).x;   // Re(conj(z)) = Re(z)
}

// ---------------------------------------------------------------- 2x binning
// Binning by exactly two in each direction with even binned sizes (the
// super-resolution case: 7676 x 7420 -> 3838 x 3710) needs no transform at the
// full size.  Along one axis of length N = 2 Nh, the kept frequencies
// K = [-Nh/2, Nh/2) are a whole period of the length-Nh DFT, so the raw
// inverse of the cropped spectrum is
//   y[m] = Nh x[2m] + sum_j x[2j+1] s[m - j],  s[d] = sum_{k in K} e^{i pi k (2d-1)/Nh}
// (circular in Nh): the even samples as they are plus the odd samples through a
// half-sample delay.  Im s[d] = (-1)^d comes only from the unpaired k = -Nh/2;
// with the 1/(Hb Wb) scale and the real part taken at the end, the 2D result is
//   Z[y, mx]  = x[y, 2mx] + (1/Wb) (x[y, odd] (*) g_x)[mx]           (rows)
//   out[my, mx] = Z[2my, mx] + (1/Hb) (Z[odd, mx] (*) g_y)[my]
//                 - (-1)^(my + mx) q / (Hb Wb),   q = sum_{j,i} (-1)^(i+j) x[2j+1, 2i+1]
// with g = Re s, which is binStack('fourier') (zorro/zorro.py:1866-1901) exactly
// in real arithmetic.  Each circular convolution of length Nh runs as a linear
// one through a power-of-two FFT of size M >= 2 Nh - 1 against the
// precomputed spectrum G = FFT_M(g wrapped) / (M Nh), two real lines packed as
// one complex line.

// Row pass: one block per (raw row pair, frame of the batch), frames fastest so
// that the blocks reading the same gain rows run together.  Writes Z [nb][Hr][Wb]
// and, per row pair p, (-1)^p times the odd row's alternating odd-sample sum.
template <int M>
__global__ __launch_bounds__(bin_threads(M)) void k_bin2_rows(const void* __restrict__ raw, int dtype, int frame0,
                                                              int nb, int Hr, int Wr, const float* __restrict__ gain,
                                                              float gscale, int use_scale, const float2* __restrict__ G,
                                                              const float2* __restrict__ twM, float* __restrict__ Z,
                                                              double* __restrict__ cpart) {
  constexpr int NT = bin_threads(M);
  extern __shared__ float2 bin_lds[];
  __shared__ double red[NT / 64];
  float2* buf = bin_lds;
  const int tid = threadIdx.x;
  const int Wb = Wr / 2;
  const int fb = blockIdx.x % nb, pr = blockIdx.x / nb;
  const int y0 = 2 * pr, y1 = y0 + 1;
  const size_t b0 = ((size_t)(frame0 + fb) * Hr + y0) * Wr, b1 = b0 + Wr;
  const size_t g0 = (size_t)y0 * Wr, g1 = g0 + Wr;
  const LdsTw tws = bin_twiddles<M, NT>(bin_lds, twM, tid);
  // each pixel pair (2j, 2j+1) of both rows in one load each: the odd samples
  // go to the transform, the even ones stay in registers for the output
  constexpr int PJ = (M / 2 + NT - 1) / NT;     // Wb <= M / 2
  float2 ev[PJ];
  double cs = 0.0;
#pragma unroll
  for (int i = 0; i < PJ; ++i) {
    const int j = tid + i * NT;
    float2 v = make_float2(0.f, 0.f), e = make_float2(0.f, 0.f);
    if (j < Wb) {
      const int x = 2 * j;
      const float2 r0 = raw_pair(raw, dtype, b0 + x), r1 = raw_pair(raw, dtype, b1 + x);
      float2 q0 = make_float2(1.f, 1.f), q1 = make_float2(1.f, 1.f);
      if (gain) {
        q0 = *reinterpret_cast<const float2*>(gain + g0 + x);
        q1 = *reinterpret_cast<const float2*>(gain + g1 + x);
      }
      // gainRef * data, then / gain (float32, as loadData)
      float a0 = gain ? q0.x * r0.x : r0.x, a1 = gain ? q0.y * r0.y : r0.y;
      float c0 = gain ? q1.x * r1.x : r1.x, c1 = gain ? q1.y * r1.y : r1.y;
      if (use_scale) { a0 /= gscale; a1 /= gscale; c0 /= gscale; c1 /= gscale; }
      e = make_float2(a0, c0);
      v = make_float2(a1, c1);
      cs += (j & 1) ? -(double)v.y : (double)v.y;
    }
    ev[i] = e;
    buf[j] = v;
  }
#pragma unroll 4
  for (int j = PJ * NT + tid; j < M; j += NT) buf[j] = make_float2(0.f, 0.f);
  cs = wave_sum_d(cs);
  if ((tid & 63) == 0) red[tid >> 6] = cs;
  __syncthreads();
  if (tid == 0) {
    double c = 0.0;
    for (int i = 0; i < NT / 64; ++i) c += red[i];
    cpart[(size_t)fb * (Hr / 2) + pr] = (pr & 1) ? -c : c;
  }
  wg_fft<M, -1, NT>(buf, tws, tid);
#pragma unroll 4
  for (int k = tid; k < M; k += NT) buf[k] = c_mul(buf[k], G[k]);
  __syncthreads();
  wg_fft<M, 1, NT>(buf, tws, tid);
  float* z0 = Z + ((size_t)fb * Hr + y0) * Wb;
  float* z1 = z0 + Wb;
#pragma unroll
  for (int i = 0; i < PJ; ++i) {
    const int mx = tid + i * NT;
    if (mx < Wb) {
      const float2 w = buf[mx];
      z0[mx] = ev[i].x + w.x;
      z1[mx] = ev[i].y + w.y;
    }
  }
}

// q / (Hb Wb) per frame from the row pairs' signed sums (one block per frame,
// fixed summation order).
__global__ __launch_bounds__(256) void k_bin2_q(const double* __restrict__ cpart, int npairs, double scale,
                                               double* __restrict__ q) {
  __shared__ double red[4];
  const int fb = blockIdx.x, tid = threadIdx.x;
  double a = 0.0;
  for (int i = tid; i < npairs; i += 256) a += cpart[(size_t)fb * npairs + i];
  a = wave_sum_d(a);
  if ((tid & 63) == 0) red[tid >> 6] = a;
  __syncthreads();
  if (tid == 0) q[fb] = ((red[0] + red[1]) + (red[2] + red[3])) * scale;
}

// 32x32 LDS-tiled transpose of float32 matrices [nb][R][C] -> [nb][C][R].
__global__ __launch_bounds__(256) void k_bin2_transpose(const float* __restrict__ in, int R, int C,
                                                       float* __restrict__ out) {
  __shared__ float tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
  const int tilesC = (C + 31) / 32, tilesR = (R + 31) / 32;
  const int b = blockIdx.x / (tilesC * tilesR);
  const int t = blockIdx.x % (tilesC * tilesR);
  const int r0 = (t / tilesC) * 32, c0 = (t % tilesC) * 32;
  const float* src = in + (size_t)b * R * C;
  float* dst = out + (size_t)b * R * C;
  for (int i = ty; i < 32; i += 8) {
    const int r = r0 + i, c = c0 + tx;
    if (r < R && c < C) tile[i][tx] = src[(size_t)r * C + c];
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    const int c = c0 + i, r = r0 + tx;
    if (r < R && c < C) dst[(size_t)c * R + r] = tile[tx][i];
  }
}

// Column pass: one block per (column pair of Z, frame of the batch), on the
// transposed ZT [nb][Wb][Hr]; writes the binned frame transposed, [nb][Wb][Hb].
template <int M>
__global__ __launch_bounds__(bin_threads(M)) void k_bin2_cols(const float* __restrict__ ZT, int nb, int Hr, int Wb,
                                                              const float2* __restrict__ G,
                                                              const float2* __restrict__ twM,
                                                              const double* __restrict__ q, float* __restrict__ outT) {
  constexpr int NT = bin_threads(M);
  extern __shared__ float2 bin_lds[];
  float2* buf = bin_lds;
  const int tid = threadIdx.x;
  const int Hb = Hr / 2;
  const int fb = blockIdx.x % nb, pc = blockIdx.x / nb;
  const int mx0 = 2 * pc;
  const float* c0 = ZT + ((size_t)fb * Wb + mx0) * Hr;
  const float* c1 = c0 + Hr;
  const LdsTw tws = bin_twiddles<M, NT>(bin_lds, twM, tid);
#pragma unroll 4
  for (int j = tid; j < M; j += NT) buf[j] = j < Hb ? make_float2(c0[2 * j + 1], c1[2 * j + 1]) : make_float2(0.f, 0.f);
  __syncthreads();
  wg_fft<M, -1, NT>(buf, tws, tid);
#pragma unroll 4
  for (int k = tid; k < M; k += NT) buf[k] = c_mul(buf[k], G[k]);
  __syncthreads();
  wg_fft<M, 1, NT>(buf, tws, tid);
  const float qs = (float)q[fb];
  float* o0 = outT + ((size_t)fb * Wb + mx0) * Hb;
  float* o1 = o0 + Hb;
#pragma unroll 4
  for (int my = tid; my < Hb; my += NT) {
    const float2 w = buf[my];
    const float t = ((my + mx0) & 1) ? qs : -qs;     // -(-1)^(my + mx0) q / (Hb Wb)
    o0[my] = (c0[2 * my] + w.x) + t;
    o1[my] = (c1[2 * my] + w.y) - t;
  }
}

// Gain product without binning: frames = gain * raw (/ gscale), float32.
__global__ __launch_bounds__(256) void k_gain_convert(const void* __restrict__ raw, int dtype, size_t npix_frame,
                                                      size_t total, const float* __restrict__ gain, float gscale,
                                                      int use_scale, float* __restrict__ frames) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256)
    frames[i] = gained(raw_value(raw, dtype, i), gain, i % npix_frame, gscale, use_scale);
}

// ------------------------------------------------------------------ host
#include <cmath>
#include <complex>
#include <mutex>
#include <vector>

namespace {

// power-of-two DFT in double on the host (iterative radix 2), forward sign
void host_fft(std::vector<std::complex<double>>& a) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    const double ang = -2.0 * M_PI / (double)len;
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < len / 2; ++k) {
        const std::complex<double> w(std::cos(ang * (double)k), std::sin(ang * (double)k));
        const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
  }
}

// e^{sign i pi n^2 / L} from n^2 mod 2L (exact integer phase)
std::complex<double> chirp_phase(long n, long L, int sign) {
  const long r = (long)(((__int128)n * n) % (2 * L));
  const double a = sign * M_PI * (double)r / (double)L;
  return {std::cos(a), std::sin(a)};
}

}  // namespace

// The half-sample delay kernel of 2x binning along an axis of binned length nh
// (even): g[d] = Re sum_{k=-nh/2}^{nh/2-1} e^{i pi k (2d-1)/nh}, from the closed
// form of the geometric sum in FP64, wrapped for a linear convolution through a
// length-M FFT (g[d] at index d mod M for |d| < nh), transformed, and scaled by
// 1/(M nh) (the inverse FFT's length and the binned size).
void zr_bin2_kernel_spectrum(int nh, int M, std::vector<float2>& G) {
  const int K = nh / 2;
  std::vector<double> g(nh);
  for (int d = 0; d < nh; ++d) {
    const double th = M_PI * (double)(2 * d - 1) / (double)nh;
    const std::complex<double> e1(std::cos(th), std::sin(th));
    const std::complex<double> lead(std::cos(-th * K), std::sin(-th * K));
    const std::complex<double> full(std::cos(th * 2 * K), std::sin(th * 2 * K));
    g[d] = (lead * (1.0 - full) / (1.0 - e1)).real();
  }
  std::vector<std::complex<double>> a(M, 0.0);
  for (int d = -(nh - 1); d <= nh - 1; ++d) a[((d % M) + M) % M] = g[((d % nh) + nh) % nh];
  host_fft(a);
  G.resize(M);
  const double sc = 1.0 / ((double)M * (double)nh);
  for (int k = 0; k < M; ++k) G[k] = make_float2((float)(a[k].real() * sc), (float)(a[k].imag() * sc));
}

int zr_bin_conv_size(int L) {
  int M = 4096;             // smallest built transform; any M >= 2L-1 gives the same DFT
  while (M < 2 * L - 1) M <<= 1;
  return M;
}

void zr_bin_chirp_tables(int L, int M, std::vector<float2>& c, std::vector<float2>& B) {
  c.resize(L);
  for (int n = 0; n < L; ++n) {
    const std::complex<double> z = chirp_phase(n, L, -1);
    c[n] = make_float2((float)z.real(), (float)z.imag());
  }
  std::vector<std::complex<double>> b(M, 0.0);
  for (int m = 0; m < L; ++m) {
    const std::complex<double> z = chirp_phase(m, L, +1);
    b[m] = z;
    if (m > 0) b[M - m] = z;
  }
  host_fft(b);
  B.resize(M);
  for (int m = 0; m < M; ++m) B[m] = make_float2((float)(b[m].real() / M), (float)(b[m].imag() / M));
}

template <int M>
static hipError_t launch_bin_rows_t(const void* raw, int dtype, int frame0, int nb, int H, int W, const float* gain,
                                    float gscale, int use_scale, const BinChirp& cw, const float2* twM, int Wb,
                                    float2* Tr, hipStream_t st) {
  static std::once_flag lds_once;
  zr_allow_full_lds((const void*)k_bin_rows<M>, lds_once);
  hipLaunchKernelGGL(k_bin_rows<M>, dim3(nb * H), dim3(bin_threads(M)), bin_lds_bytes(M), st, raw, dtype, frame0, H, W, gain,
                     gscale, use_scale, cw, twM, Wb, Tr);
  return hipGetLastError();
}

template <int M>
static hipError_t launch_bin_cols_t(const float2* Tc, int nb, int H, int Wb, int Hb, const BinChirp& ch,
                                    const BinChirp& chb, const float2* twM, float scale, float2* Gc, hipStream_t st) {
  static std::once_flag lds_once;
  zr_allow_full_lds((const void*)k_bin_cols<M>, lds_once);
  hipLaunchKernelGGL(k_bin_cols<M>, dim3(nb * Wb), dim3(bin_threads(M)), bin_lds_bytes(M), st, Tc, H, Wb, Hb, ch, chb, twM,
                     scale, Gc);
  return hipGetLastError();
}

template <int M>
static hipError_t launch_bin_out_t(const float2* Gr, int nb, int Hb, int Wb, const BinChirp& cwb, const float2* twM,
                                   int frame0, float* frames, hipStream_t st) {
  static std::once_flag lds_once;
  zr_allow_full_lds((const void*)k_bin_out<M>, lds_once);
  hipLaunchKernelGGL(k_bin_out<M>, dim3(nb * Hb), dim3(bin_threads(M)), bin_lds_bytes(M), st, Gr, Hb, Wb, cwb, twM, frame0,
                     frames);
  return hipGetLastError();
}

hipError_t zr_launch_bin_rows(int M, const void* raw, int dtype, int frame0, int nb, int H, int W, const float* gain,
                              float gscale, int use_scale, const BinChirp& cw, const float2* twM, int Wb, float2* Tr,
                              hipStream_t st) {
  switch (M) {
    case 4096: return launch_bin_rows_t<4096>(raw, dtype, frame0, nb, H, W, gain, gscale, use_scale, cw, twM, Wb, Tr, st);
    case 8192: return launch_bin_rows_t<8192>(raw, dtype, frame0, nb, H, W, gain, gscale, use_scale, cw, twM, Wb, Tr, st);
    case 16384: return launch_bin_rows_t<16384>(raw, dtype, frame0, nb, H, W, gain, gscale, use_scale, cw, twM, Wb, Tr, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t zr_launch_bin_cols(int M, const float2* Tc, int nb, int H, int Wb, int Hb, const BinChirp& ch,
                              const BinChirp& chb, const float2* twM, float scale, float2* Gc, hipStream_t st) {
  switch (M) {
    case 4096: return launch_bin_cols_t<4096>(Tc, nb, H, Wb, Hb, ch, chb, twM, scale, Gc, st);
    case 8192: return launch_bin_cols_t<8192>(Tc, nb, H, Wb, Hb, ch, chb, twM, scale, Gc, st);
    case 16384: return launch_bin_cols_t<16384>(Tc, nb, H, Wb, Hb, ch, chb, twM, scale, Gc, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t zr_launch_bin_out(int M, const float2* Gr, int nb, int Hb, int Wb, const BinChirp& cwb, const float2* twM,
                             int frame0, float* frames, hipStream_t st) {
  switch (M) {
    case 4096: return launch_bin_out_t<4096>(Gr, nb, Hb, Wb, cwb, twM, frame0, frames, st);
    case 8192: return launch_bin_out_t<8192>(Gr, nb, Hb, Wb, cwb, twM, frame0, frames, st);
    case 16384: return launch_bin_out_t<16384>(Gr, nb, Hb, Wb, cwb, twM, frame0, frames, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t zr_launch_bin_transpose(const float2* in, int nb, int R, int C, float2* out, hipStream_t st) {
  const int tiles = ((R + 31) / 32) * ((C + 31) / 32);
  hipLaunchKernelGGL(k_bin_transpose, dim3(nb * tiles), dim3(256), 0, st, in, R, C, out);
  return hipGetLastError();
}

template <int M>
static void launch_bin2_t(const void* raw, int dtype, int frame0, int nb, int Hr, int Wr, const float* gain,
                          float gscale, int use_scale, const float2* Gx, const float2* Gy, const float2* twM,
                          float* A, float* B, double* cpart, double* q, float* frames, hipStream_t st) {
  static std::once_flag lr, lc;
  zr_allow_full_lds((const void*)k_bin2_rows<M>, lr);
  zr_allow_full_lds((const void*)k_bin2_cols<M>, lc);
  const int Hb = Hr / 2, Wb = Wr / 2;
  hipLaunchKernelGGL(k_bin2_rows<M>, dim3(nb * (Hr / 2)), dim3(bin_threads(M)), bin_lds_bytes(M), st, raw, dtype,
                     frame0, nb, Hr, Wr, gain, gscale, use_scale, Gx, twM, A, cpart);
  hipLaunchKernelGGL(k_bin2_q, dim3(nb), dim3(256), 0, st, cpart, Hr / 2, 1.0 / ((double)Hb * (double)Wb), q);
  hipLaunchKernelGGL(k_bin2_transpose, dim3(nb * ((Hr + 31) / 32) * ((Wb + 31) / 32)), dim3(256), 0, st, A, Hr, Wb, B);
  hipLaunchKernelGGL(k_bin2_cols<M>, dim3(nb * (Wb / 2)), dim3(bin_threads(M)), bin_lds_bytes(M), st, B, nb, Hr, Wb,
                     Gy, twM, q, A);
  hipLaunchKernelGGL(k_bin2_transpose, dim3(nb * ((Wb + 31) / 32) * ((Hb + 31) / 32)), dim3(256), 0, st, A, Wb, Hb,
                     frames + (size_t)frame0 * Hb * Wb);
}

hipError_t zr_launch_bin2(int M, const void* raw, int dtype, int frame0, int nb, int Hr, int Wr, const float* gain,
                          float gscale, int use_scale, const float2* Gx, const float2* Gy, const float2* twM, float* A,
                          float* B, double* cpart, double* q, float* frames, hipStream_t st) {
  switch (M) {
    case 4096: launch_bin2_t<4096>(raw, dtype, frame0, nb, Hr, Wr, gain, gscale, use_scale, Gx, Gy, twM, A, B, cpart, q, frames, st); break;
    case 8192: launch_bin2_t<8192>(raw, dtype, frame0, nb, Hr, Wr, gain, gscale, use_scale, Gx, Gy, twM, A, B, cpart, q, frames, st); break;
    case 16384: launch_bin2_t<16384>(raw, dtype, frame0, nb, Hr, Wr, gain, gscale, use_scale, Gx, Gy, twM, A, B, cpart, q, frames, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t zr_launch_gain_convert(const void* raw, int dtype, size_t npix_frame, size_t total, const float* gain,
                                  float gscale, int use_scale, float* frames, hipStream_t st) {
  const size_t blocks = std::min<size_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_gain_convert, dim3((unsigned)blocks), dim3(256), 0, st, raw, dtype, npix_frame, total, gain,
                     gscale, use_scale, frames);
  return hipGetLastError();
}
