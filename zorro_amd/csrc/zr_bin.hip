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

// gain product in float32 as loadData does it: gainRef * data, then / gain
ZR_DEV float gained(float v, const float* gain, size_t pix, float gscale, int use_scale) {
  if (gain) v = gain[pix] * v;
  if (use_scale) v = v / gscale;
  return v;
}

// The chirp-z convolution on buf (LDS, M entries, holding x[n] c[n], zero past
// L): on return buf[k] * c[k] is the DFT for k < L.  Synchronised on entry and exit.
template <int M, int NT>
ZR_DEV void bluestein_conv(float2* buf, const float2* __restrict__ B, const float2* __restrict__ twM, int tid) {
  wg_fft<M, -1, NT>(buf, twM, tid);
  for (int e = tid; e < M; e += NT) buf[e] = c_mul(buf[e], B[e]);
  __syncthreads();
  wg_fft<M, 1, NT>(buf, twM, tid);
}

// Row pass: one block per (frame in batch, raw row y).
template <int M>
__global__ __launch_bounds__(1024) void k_bin_rows(const void* __restrict__ raw, int dtype, int frame0, int H, int W,
                                                   const float* __restrict__ gain, float gscale, int use_scale,
                                                   BinChirp cw, const float2* __restrict__ twM, int Wb,
                                                   float2* __restrict__ Tr) {
  constexpr int NT = 1024;
  extern __shared__ float2 bin_lds[];
  float2* buf = bin_lds;
  const int tid = threadIdx.x;
  const int fb = blockIdx.x / H, y = blockIdx.x % H;
  const size_t base = ((size_t)(frame0 + fb) * H + y) * W;
  for (int x = tid; x < M; x += NT) {
    float2 v = make_float2(0.f, 0.f);
    if (x < W) {
      const float a = gained(raw_value(raw, dtype, base + x), gain, (size_t)y * W + x, gscale, use_scale);
      v = c_scale(cw.c[x], a);
    }
    buf[x] = v;
  }
  __syncthreads();
  bluestein_conv<M, NT>(buf, cw.B, twM, tid);
  float2* dst = Tr + ((size_t)fb * H + y) * Wb;
  const int h2 = Wb / 2;
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
__global__ __launch_bounds__(1024) void k_bin_cols(const float2* __restrict__ Tc, int H, int Wb, int Hb, BinChirp ch,
                                                   BinChirp chb, const float2* __restrict__ twM, float scale,
                                                   float2* __restrict__ Gc) {
  constexpr int NT = 1024;
  constexpr int PK = (M / 2 + NT - 1) / NT;    // kept rows per thread (Hb <= M/2)
  extern __shared__ float2 bin_lds[];
  float2* buf = bin_lds;
  const int tid = threadIdx.x;
  const size_t col = blockIdx.x;               // fb * Wb + kx'
  const float2* src = Tc + col * H;
  for (int y = tid; y < M; y += NT) buf[y] = y < H ? c_mul(src[y], ch.c[y]) : make_float2(0.f, 0.f);
  __syncthreads();
  bluestein_conv<M, NT>(buf, ch.B, twM, tid);
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
  for (int e = Hb + tid; e < M; e += NT) buf[e] = make_float2(0.f, 0.f);
  __syncthreads();
  bluestein_conv<M, NT>(buf, chb.B, twM, tid);
  float2* dst = Gc + col * Hb;
  for (int n = tid; n < Hb; n += NT) dst[n] = c_conj(c_mul(chb.c[n], buf[n]));
}

// Final rows: one block per (frame in batch, binned row).  Inverse DFT along x
// (length Wb), real part.
template <int M>
__global__ __launch_bounds__(1024) void k_bin_out(const float2* __restrict__ Gr, int Hb, int Wb, BinChirp cwb,
                                                  const float2* __restrict__ twM, int frame0,
                                                  float* __restrict__ frames) {
  constexpr int NT = 1024;
  extern __shared__ float2 bin_lds[];
  float2* buf = bin_lds;
  const int tid = threadIdx.x;
  const int fb = blockIdx.x / Hb, ny = blockIdx.x % Hb;
  const float2* src = Gr + ((size_t)fb * Hb + ny) * Wb;
  for (int k = tid; k < M; k += NT) buf[k] = k < Wb ? c_mul(c_conj(src[k]), cwb.c[k]) : make_float2(0.f, 0.f);
  __syncthreads();
  bluestein_conv<M, NT>(buf, cwb.B, twM, tid);
  float* dst = frames + ((size_t)(frame0 + fb) * Hb + ny) * Wb;
  for (int n = tid; n < Wb; n += NT) dst[n] = c_mul(cwb.c[n], buf[n]).x;   // Re(conj(z)) = Re(z)
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
  hipLaunchKernelGGL(k_bin_rows<M>, dim3(nb * H), dim3(1024), zr_fft_lds(M) * sizeof(float2), st, raw, dtype, frame0, H, W, gain,
                     gscale, use_scale, cw, twM, Wb, Tr);
  return hipGetLastError();
}

template <int M>
static hipError_t launch_bin_cols_t(const float2* Tc, int nb, int H, int Wb, int Hb, const BinChirp& ch,
                                    const BinChirp& chb, const float2* twM, float scale, float2* Gc, hipStream_t st) {
  static std::once_flag lds_once;
  zr_allow_full_lds((const void*)k_bin_cols<M>, lds_once);
  hipLaunchKernelGGL(k_bin_cols<M>, dim3(nb * Wb), dim3(1024), zr_fft_lds(M) * sizeof(float2), st, Tc, H, Wb, Hb, ch, chb, twM,
                     scale, Gc);
  return hipGetLastError();
}

template <int M>
static hipError_t launch_bin_out_t(const float2* Gr, int nb, int Hb, int Wb, const BinChirp& cwb, const float2* twM,
                                   int frame0, float* frames, hipStream_t st) {
  static std::once_flag lds_once;
  zr_allow_full_lds((const void*)k_bin_out<M>, lds_once);
  hipLaunchKernelGGL(k_bin_out<M>, dim3(nb * Hb), dim3(1024), zr_fft_lds(M) * sizeof(float2), st, Gr, Hb, Wb, cwb, twM, frame0,
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

hipError_t zr_launch_gain_convert(const void* raw, int dtype, size_t npix_frame, size_t total, const float* gain,
                                  float gscale, int use_scale, float* frames, hipStream_t st) {
  const size_t blocks = std::min<size_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_gain_convert, dim3((unsigned)blocks), dim3(256), 0, st, raw, dtype, npix_frame, total, gain,
                     gscale, use_scale, frames);
  return hipGetLastError();
}
