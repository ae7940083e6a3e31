// Internal launcher declarations shared by the kernel files and the C ABI.
#pragma once
#include <hip/hip_runtime.h>

struct PeakParamsHost {
  int m, suppress, bconv, bopti, ry, rx, subpix;
  float scale_y, scale_x;
  int n;
  int disc;   // maxShift disc radius (0: m)
};

// preShift inputs/outputs of the peak stage: per-pair window centres, the
// shifts in cropped pixels for the next pair, and a flag set when a disc holds
// no positive value
struct PeakPreShift {
  const int2* ctr;
  float* pshift;
  int* flag;
};

// LDS available to one workgroup on gfx950
constexpr int ZR_LDS_BYTES = 160 * 1024;

// Dynamic LDS of the peak kernel: a scratch region that holds the
// reflect-padded window of the Gaussian passes and later the sub-pixel
// transform, followed by the filtered 2m x 2m window.
inline size_t zr_peaks_lds_bytes(const PeakParamsHost* hp, int* tmp_floats) {
  const size_t n2 = 2 * (size_t)hp->m, nn = n2 * n2;
  const size_t G = 16 * (size_t)(hp->subpix > 1 ? hp->subpix : 1);
  size_t t = nn > 34 * G ? nn : 34 * G;
  if (hp->bconv || hp->bopti) {
    const size_t r = (size_t)(hp->ry > hp->rx ? hp->ry : hp->rx);
    const size_t padded = (n2 + 2 * r) * n2;
    t = t > padded ? t : padded;
  }
  t = (t + 63) & ~(size_t)63;
  if (tmp_floats) *tmp_floats = (int)t;
  return (t + nn) * sizeof(float);
}

// zr_xcorr.hip
// mode: 0 frame (sums into part), 1 mean-centred masked frame, 2 its square,
// 3 the mask (stats: per-frame (sum, sumsq) for modes 1 and 2); umask: the
// user's binary mask [H][W] multiplying the frame region in modes 1-3, or NULL
// rows per k1_rows block for a padded width (the frame-sum partials come in
// Hp / rows blocks per frame)
int zr_k1_rows_per_block(int Wp);
hipError_t zr_launch_k1_rows(int Wp, const float* frames, int H, int W, int Hp, int Wh, int frame0, int nb,
                             const float2* twL, const float2* tw2L, float2* T, double* part, int mode,
                             const double* stats, const float* umask, int* rb_out, hipStream_t st);
// per-frame FP64 partial sums over H x W; nrb_out = row blocks per frame
hipError_t zr_launch_k1_sums(const float* frames, int n, int H, int W, double* part, int* nrb_out, hipStream_t st);
// masked NCC from the pair and mask-term windows (in place on the P pair windows)
hipError_t zr_launch_mnxc_combine(float* windows, const int4* pairs, int P, int N, int m, hipStream_t st);
hipError_t zr_launch_k1_cols(int Hp, const float2* T, int Wh, int Hc, int frame0, int nb, const float2* twH,
                             float2* S, float2* Sside, int full, hipStream_t st);
hipError_t zr_launch_k1_stats(const double* part, int nrb, int n, double* stats, hipStream_t st);
// ctr (nullable): per-pair window centre (cy, cx), lags [c - m, c + m) (preShift);
// s0 (nullable): per (pair, column) c(kx) Re sum_ky of the product (lag-0 sums)
hipError_t zr_launch_k2_cols(int Lper, const float2* S, const float2* Sside, const int4* work, int npairs, int Wh,
                             int Hc, int Wc, int m, int cols_per_wave, int dc_zero, const float2* twHc, float2* Z,
                             int Wk, int Hs, int hoff, int side_in_col, const int2* ctr, const float2* twWc,
                             float* s0, hipStream_t st);
// column pass for m <= L (L = Hc/64 in {32, 48, 64}): sums over lanes of
// pre-twiddled values with the [L+1][64] table of zr_launch_k2_lane_twiddles
bool zr_k2_cols_r2_ok(int Lper, int m);
hipError_t zr_launch_k2_lane_twiddles(int Lper, int Hc, float2* ltw, hipStream_t st);
// pairs_per_wave > 0 (and ctr == NULL): pair mode, one column per block over
// 4 * pairs_per_wave consecutive pair-list entries; else column groups of
// 4 * cols_per_wave columns per pair.
hipError_t zr_launch_k2_cols_r2(int Lper, const float2* S, const float2* Sside, const int4* work, int npairs, int Wh,
                                int Hc, int Wc, int m, int cols_per_wave, int pairs_per_wave, int dc_zero,
                                const float2* ltw, float2* Z, int Wk, int Hs, int hoff, int side_in_col,
                                const int2* ctr, const float2* twHc, const float2* twWc, float* s0, hipStream_t st);
// btab (non-NULL, 2m a multiple of 16): the matrix-core pass with the table of
// zr_launch_k2_twiddles; otherwise the vector-ALU pass
// sub (nullable): run the npairs pairs sub[0..npairs) (window index .z) instead of 0..npairs
hipError_t zr_launch_k2_rows(const float2* Z, const int4* pairs, int npairs, int Wh, int Wc, int m,
                             const float2* twWc, const float2* btab, const double* den, float nc2, float* windows,
                             const int4* sub, hipStream_t st);
hipError_t zr_launch_k2_twiddles(int Wh, int Wc, int m, float2* btab, hipStream_t st);
// rows of the matrix-core twiddle table: Wh padded to whole pairs of the pass's
// kx chunks (the padding rows are zero)
int zr_k2_btab_rows(int Wh);
hipError_t zr_launch_k2_den(const int4* pairs, int npairs, const double* stats, double npix, double nc2, double* den,
                            hipStream_t st);
hipError_t zr_launch_peaks(const float* windows, const int4* pairs, int npairs, const PeakParamsHost* hp,
                           const float* gy, const float* gx, const float2* ktab, float* shifts, float* corr,
                           float* peaksig, float* origin, float* bsigma, const PeakPreShift* ps, hipStream_t st);
// preShift: window centres of the pairs dpairs[0..cnt) from the previous pairs' shifts
hipError_t zr_launch_preshift_centres(const int4* dpairs, int cnt, const int* pidx, const float* pshift, int N,
                                      int2* ctr, hipStream_t st);
// preShift: originTriMat from the lag-0 column sums
hipError_t zr_launch_origin_sums(const int4* pairs, int npairs, const float* s0, int Wh, const double* den, double nc2,
                                 int N, float* origin, hipStream_t st);
// triMode 'auto': zero the trimat entries after each row's stopping pair
hipError_t zr_launch_auto_prune(float* shifts, float* corr, float* peaksig, float* origin, int N, int auto_max,
                                double thres, hipStream_t st);
// triMode 'refine': sums of spectra slots, closed-form denominators, window combination
hipError_t zr_launch_spec_sum(float2* S, size_t slot_elems, int slot0, int count, int out_slot, hipStream_t st);
hipError_t zr_launch_refine_den(const double* stats, int N, double npix, double nc2, double* den, hipStream_t st);
hipError_t zr_launch_refine_combine(float* windows, int N, int m, int masked, int out0, const double* cm,
                                    double npix, hipStream_t st);
hipError_t zr_launch_common_mean(const double* stats, int N, double* out, hipStream_t st);

// zr_bin.hip: gain product and Fourier binning through Bluestein transforms.
// A chirp table set for length L: c[L] = e^{-i pi n^2/L}, B[M] = FFT_M(conj chirp)/M.
struct BinChirp {
  const float2* c;
  const float2* B;
  int L;
};
#include <vector>
int zr_bin_conv_size(int L);   // power of two >= max(4096, 2L-1)
void zr_bin_chirp_tables(int L, int M, std::vector<float2>& c, std::vector<float2>& B);
hipError_t zr_launch_bin_rows(int M, const void* raw, int dtype, int frame0, int nb, int H, int W, const float* gain,
                              float gscale, int use_scale, const BinChirp& cw, const float2* twM, int Wb, float2* Tr,
                              hipStream_t st);
hipError_t zr_launch_bin_cols(int M, const float2* Tc, int nb, int H, int Wb, int Hb, const BinChirp& ch,
                              const BinChirp& chb, const float2* twM, float scale, float2* Gc, hipStream_t st);
hipError_t zr_launch_bin_out(int M, const float2* Gr, int nb, int Hb, int Wb, const BinChirp& cwb, const float2* twM,
                             int frame0, float* frames, hipStream_t st);
hipError_t zr_launch_bin_transpose(const float2* in, int nb, int R, int C, float2* out, hipStream_t st);
// Binning by exactly two with even binned sizes (Hr = 2 Hb, Wr = 2 Wb): the
// polyphase form (even samples + half-sample-delayed odd samples per axis);
// M: power-of-two convolution size >= 2 max(Hb, Wb) - 1; Gx/Gy from
// zr_bin2_kernel_spectrum; A, B: float scratch of nb * Hr * Wb each; cpart:
// nb * Hr / 2 doubles; q: nb doubles.  Writes frames [frame0 .. frame0 + nb).
void zr_bin2_kernel_spectrum(int nh, int M, std::vector<float2>& G);
hipError_t zr_launch_bin2(int M, const void* raw, int dtype, int frame0, int nb, int Hr, int Wr, const float* gain,
                          float gscale, int use_scale, const float2* Gx, const float2* Gy, const float2* twM, float* A,
                          float* B, double* cpart, double* q, float* frames, hipStream_t st);
hipError_t zr_launch_gain_convert(const void* raw, int dtype, size_t npix_frame, size_t total, const float* gain,
                                  float gscale, int use_scale, float* frames, hipStream_t st);
// zero fill by a kernel (graph-capturable); bytes multiple of 4
hipError_t zr_launch_fill_zero(void* p, size_t bytes, hipStream_t st);

// batches: copy the block [off, off + n)^2 of a batch's trimats into a movie's [n][n] ones
hipError_t zr_launch_gather_trimats(const float* shifts, const float* corr, const float* peaksig, const float* origin,
                                    int Nt, int off, int n, float* d_shifts, float* d_corr, float* d_peaksig,
                                    float* d_origin, hipStream_t st);
