/* zorro_hip.h -- C ABI of the MI355X drift-registration library (libzorro_hip.so).
 *
 * The reference (LBEM-CH/zorro) has no FFI: its seam is the Python object
 * zorro.ImageRegistrator (zorro/zorro.py:78), configured by attributes
 * (zorro/zorro.py:82-243) and run by alignImageStack() (zorro/zorro.py:1569).
 * Each entry point below replaces one stage of that method; the comment on
 * each names the reference function it stands in for.  A ctypes binding that a
 * maintainer would add beside the pyFFTW/numexprz path is shown in
 * INTEGRATION.md; zorro_amd/_lib.py is that binding.
 *
 * Conventions
 *   - every int-returning call returns ZR_OK (0) or a negative ZR_ERR_* code;
 *     zr_last_error() gives the message of the last failure on this thread;
 *   - shapes are row-major [N][H][W] float32, shifts are (y, x);
 *   - host buffers are owned by the caller; device buffers by the movie;
 *   - a zr_movie is bound to one device and one HIP stream and is not re-entrant;
 *     stage calls are asynchronous on that stream, zr_get_* calls synchronise.
 */
#ifndef ZORRO_HIP_H
#define ZORRO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ZR_OK = 0,
  ZR_ERR_ARG = -1,          /* invalid argument / shape (reference: ValueError, IndexError) */
  ZR_ERR_HIP = -2,          /* HIP runtime failure */
  ZR_ERR_UNSUPPORTED = -3,  /* option outside the built path (reference: NotImplemented/broken mode) */
  ZR_ERR_STATE = -4,        /* stage called before its inputs exist */
  ZR_ERR_NOMEM = -5         /* device allocation failed */
};

/* triMode (zorro/zorro.py:207, :559-581): 'tri' all pairs; 'diag' the band
 * diagStart < J-I <= diagWidth + diagStart; 'auto' each row stops after the
 * first pair whose peak significance is <= peaksigThres or J-I >= autoMax
 * (:992-996); 'first' correlates frame 0 with every other frame (no solver,
 * :1628-1635); 'refine' correlates every frame with the sum of the others
 * (no solver, :934-956, :1636-1647). */
enum { ZR_TRI_TRI = 0, ZR_TRI_DIAG = 1, ZR_TRI_AUTO = 2, ZR_TRI_FIRST = 3, ZR_TRI_REFINE = 4 };
/* Bmode (zorro/zorro.py:117, :1104-1141) */
enum { ZR_B_NONE = 0, ZR_B_CONV = 1, ZR_B_OPTI = 2 };
/* shiftMethod (zorro/zorro.py:190, :1734-1808).  FOURIER: the shift theorem on
 * the zero-padded frame's spectrum (the reference's :1748-1772 as intended; it
 * squares the y phase term and cannot run, xcorrnm2_tri failing first at
 * :895-896): aligned = crop(Re(IFFT(FFT(pad0(frame)) exp(-2 pi i (fy ty + fx tx))))
 * / (Hp Wp)) with the fftshifted frequency mesh (Nyquist at -1/2). */
enum { ZR_SHIFT_INTEGER = 0, ZR_SHIFT_LANCZOS = 1, ZR_SHIFT_FOURIER = 2 };
/* weightMode (zorro/zorro.py:199, :1335-1360) */
enum { ZR_W_AUTOLOGISTIC = 0, ZR_W_LOGISTIC = 1, ZR_W_CORR = 2, ZR_W_UNWEIGHTED = 3, ZR_W_NORM = 4 };

/* The hot-path subset of ImageRegistrator's attributes (zorro/zorro.py:82-243). */
typedef struct zr_params {
  int tri_mode, diag_width, diag_start;   /* triMode, diagWidth, diagStart */
  int max_shift;                          /* maxShift (cropped pixels) */
  int b_mode;                             /* Bmode */
  double b_rad;                           /* Brad */
  int sub_pix_reg;                        /* subPixReg */
  int shift_method;                       /* shiftMethod */
  int weight_mode;                        /* weightMode */
  double peaksig_thres, logistic_k, logistic_nu; /* fixed 'logistic' weights */
  int origin_centroid;                    /* originMode == 'centroid' */
  int suppress_origin;                    /* suppressOrigin */
  int fou_crop[2];                        /* fouCrop; {0,0} = none */
  int shape_padded[2];                    /* shapePadded */
  int do_dose;                            /* 'dose' in filterMode */
  double dose_a, dose_b, dose_c, dose_order, dose_missing; /* doseFiltParam[1..5] */
  double pixelsize_nm;                    /* pixelsize (nm) */
  double voltage_kv;                      /* voltage */
  int do_frc;                             /* doLazyFRC */
  int frc_box;                            /* tiledFRC box */
  double frc_overlap;                     /* tiledFRC overlap */
  int auto_max;                           /* autoMax (triMode 'auto') */
  int shape_binned[2];                    /* shapeBinned; {0,0} = no binning (binStack 'fourier') */
  /* 'hot' in filterMode: hotpixFilter (zorro/zorro.py:2497-2662) with hotpixInfo
   * (:139-141); decorrOutliers (random noise) is not built */
  int do_hot;
  double hot_logistic_k, hot_relax, hot_max_sigma;  /* logisticK, relax, maxSigma */
  int hot_psf_k2;                                   /* psf == 'K2' (else the uniform 3x3 kernel) */
  double hot_cutoff_lower, hot_cutoff_upper;        /* cut-offs used when a line fit fails */
  int hot_fixed_cutoffs;                            /* 1: use the two cut-offs as given, no fits
                                                       (hotpixFilter(cutoffLower, cutoffUpper)) */
  double hot_neighbour_thres;                       /* neighbourThres */
  double hot_mu_sigma[2];                           /* both > 0: fixed normal-CDF (mu, sigma) instead of the fit */
  /* preShift (zorro/zorro.py:194, :1148-1163): the maxShift disc of pair (I, J)
   * follows round(shift of (I, J-1)); unmasked square maps, Bmode conv/none,
   * triMode tri/diag/auto/first; windows of maxShift + 8 + Gaussian radius
   * (rounded up to 8, at most 64) around each pair's disc */
  int pre_shift;
  /* filterMode steps applied after the registration, in list order
   * (zorro/zorro.py:1689-1711): ZR_FILTER_* values in filters[0 .. n_filters).
   * n_filters = 0 with do_dose = 1 means the single step ZR_FILTER_DOSE; do_dose
   * is set by zr_movie_create when a DOSE or DOSENORM step is present. */
  int n_filters;
  int filters[4];
  /* launch tuning (no reference counterpart beyond FFTW's planning effort and
   * wisdom, zorro/zorro_util.py:966-1066); 0 picks the built-in default:
   * frames per forward-spectra launch pair (default 8) and lag-window columns
   * per wave of the pair transform's column-group walk (default 8), and
   * pair-list entries per wave of its pair walk (one column per workgroup over
   * 4 * k2_pairs_per_wave pair-list entries, default 16).  The pair walk runs
   * unless k2_cols_per_wave > 0 and k2_pairs_per_wave == 0.  Results do not
   * depend on them. */
  int k1_batch;
  int k2_cols_per_wave;
  int k2_pairs_per_wave;
  /* user_mask = 1: the correlation runs under the frame region times one
   * binary mask [H][W] given by zr_movie_set_mask (ImageRegistrator.masks of
   * shape [1,H,W], zorro/zorro.py:584-589, padded by padStack :1942-1955). */
  int user_mask;
} zr_params;

/* filterMode steps (zorro/zorro.py:1692-1711).  DOSE: doseFilter(normalize=False),
 * DOSENORM: doseFilter(normalize=True) (each rewrites filtSum from the aligned
 * frames); GAUSSLP: filtSum = scipy.ndimage.gaussian_filter(filtSum, 3.0);
 * BACKGROUND: filtSum -= backgroundEstimate(filtSum) (zorro/zorro_util.py:497-549:
 * a 2D Gaussian fitted under an L1 error by Powell's method to the sigma-4
 * low-passed sum sampled every 128 pixels; left unchanged when the fit does not
 * converge).  The last two start from a copy of imageSum when no filtSum exists
 * yet. */
enum { ZR_FILTER_DOSE = 1, ZR_FILTER_DOSENORM = 2, ZR_FILTER_GAUSSLP = 3, ZR_FILTER_BACKGROUND = 4 };

/* Element types of raw frames (MRC modes 2 float32, 0 uint8 counts, 6 uint16,
 * 101 4-bit counts packed two per byte with the first pixel in the low nibble
 * (frame width even), 1 int16). */
enum { ZR_DTYPE_F32 = 0, ZR_DTYPE_U8 = 1, ZR_DTYPE_U16 = 2, ZR_DTYPE_U4 = 3, ZR_DTYPE_I16 = 4 };

typedef struct zr_movie zr_movie;

/* Message of the last failing call on this thread ("" if none). */
const char* zr_last_error(void);

/* Number of HIP devices visible. */
int zr_device_count(int* n);

/* Plan one movie shape: tables, pair list and device workspaces.
 * Stands in for ImageRegistrator.__init_xcorrnm2 (zorro/zorro.py:540-637). */
int zr_movie_create(int device, const zr_params* p, int n_frames, int h, int w, zr_movie** out);

/* zr_movie_create with the pairs given by a boolean triIndices matrix
 * (uint8 [n_frames][n_frames], row-major; pair (I, J) correlated when
 * tri_indices[I][J] != 0) instead of the triMode's pattern.  Replaces
 * xcorrnm2_tri(triIndices=...) (zorro/zorro.py:773, :552-557, :959-975); the
 * trimats and the solver then see exactly these pairs.  Needs tri_mode
 * ZR_TRI_TRI and no pre_shift; entries on or below the diagonal are refused. */
int zr_movie_create_pairs(int device, const zr_params* p, int n_frames, int h, int w, const uint8_t* tri_indices,
                          zr_movie** out);
void zr_movie_destroy(zr_movie* mv);

/* Run the movie's work on a caller-provided HIP stream (NULL: the movie's own). */
int zr_movie_set_stream(zr_movie* mv, void* hip_stream);

/* Provide the frames f32 [N][H][W].  on_device=1: a device pointer that the
 * caller keeps alive (no copy); on_device=0: host memory, copied.
 * Replaces the assignment of ImageRegistrator.images (zorro/zorro.py:116). */
int zr_movie_set_frames(zr_movie* mv, const float* frames, int on_device);

/* The user mask of a plan created with user_mask = 1: H x W float32 values, 0
 * or 1 (host masks are checked), for every frame; replaces the reference's
 * masks [1,H,W] (zorro/zorro.py:584-589, used by mnxc2 :1014-1097).  Returns
 * after the copy.  A mask per frame (masks.shape[0] > 1) has no counterpart:
 * the reference's own path raises (xcorrnm2_tri reads an unbound local
 * baseMaskFFT at :970, mnxc2 an undefined baseMaskFFT_block at :1048). */
int zr_movie_set_mask(zr_movie* mv, const float* mask, int on_device);

/* Provide raw frames [N][h][w] (h, w as given to zr_movie_create) of type
 * dtype, optionally times a gain reference f32 [h][w] (already rotated/flipped
 * per gainInfo) and divided by gain_scalar (0 or 1: none), then Fourier-binned
 * to shape_binned when set.  Replaces loadData's gain product
 * (zorro/zorro.py:3416-3419, :3455-3482) and binStack('fourier') (:1826-1904).
 * Host inputs (on_device / gain_on_device = 0) are copied; the call returns
 * when they may be reused.  ZR_DTYPE_U4 frames occupy N*h*w/2 bytes and need
 * an even w (the MRC reader of the reference, mrcz, unpacks them on the host). */
int zr_movie_set_frames_raw(zr_movie* mv, const void* frames, int dtype, int on_device, const float* gain,
                            int gain_on_device, double gain_scalar);

/* All-pairs masked normalised cross-correlation, peak location, significance
 * and sub-pixel shifts into device trimats.
 * Replaces xcorrnm2_tri + mnxc2 + locatePeak (zorro/zorro.py:773-1231).
 * With do_hot set, the outlier-pixel filter runs first and rewrites the frames
 * in place (hotpixFilter replaces self.images, zorro/zorro.py:1595-1596,
 * :2647); frames given by device pointer are modified. */
int zr_xcorr(zr_movie* mv);

/* Weights and the weighted-L1 shift solution (FP64), translations.
 * Replaces shiftsSolver + cdfLogisticCurve (zorro/zorro.py:1260-1567, :1981-2022). */
int zr_solve(zr_movie* mv);

/* Replace the correlation results with caller-supplied triangular matrices
 * (host, row-major: shifts [N][N][2] (y, x), corr [N][N], peaksig [N][N]) so
 * that zr_solve runs on them.  Replaces the inputs of shiftsSolver
 * (zorro/zorro.py:1260-1261: shiftsTriMat_in, corrTriMat_in, peaksigTriMat_in);
 * equations are the movie's pairs (triMode). */
int zr_set_trimats(zr_movie* mv, const float* shifts, const float* corr, const float* peaksig);

/* Lanczos/integer shift of every frame, imageSum and even/odd half sums;
 * aligned_out (device, [N][H][W], nullable) receives the shifted frames.
 * Replaces applyShifts (zorro/zorro.py:1734-1808) and the half sums of
 * lazyFouRingCorr (zorro/zorro.py:2226-2235). */
int zr_apply_shifts(zr_movie* mv, float* aligned_out);

/* Tiled even/odd Fourier ring correlation.
 * Replaces tiledFRC + rotmean (zorro/zorro.py:2246-2310, zorro/zorro_util.py:147-224). */
int zr_frc(zr_movie* mv);

/* Crop limits and tile counts of the next zr_frc from caller translations
 * f64 [n][2] instead of the movie's own: tiledFRC(Image1, Image2, trans=...)
 * (zorro/zorro.py:2246-2274), as evenOddFouRingCorr passes the even half's
 * translations (:2213-2214).  Call after zr_solve. */
int zr_set_frc_translations(zr_movie* mv, const double* trans, int n);

/* Unblur-style dose-filtered sum (normalize=False).
 * Replaces doseFilter (zorro/zorro.py:2349-2494). */
int zr_dose_filter(zr_movie* mv);

/* One filterMode step (ZR_FILTER_*) on filtSum, as the loop of alignImageStack
 * applies it (zorro/zorro.py:1692-1711): DOSE / DOSENORM run doseFilter with
 * normalize False / True (zorro/zorro.py:2349-2494; the dose stage also forms
 * imageSum and the half sums), GAUSSLP the 3-pixel Gaussian low-pass
 * (scipy.ndimage.gaussian_filter, reflect edges, radius 12; :1707-1711) after
 * zr_apply_shifts or a dose step. */
int zr_filter(zr_movie* mv, int step);

/* applyShifts (when no dose step forms the sums) followed by every filterMode
 * step of the parameters in order: the part of alignImageStack after the
 * solver (zorro/zorro.py:1672-1711) without the FRC. */
int zr_post_filters(zr_movie* mv);

/* Every stage enabled by the parameters, in alignImageStack order
 * (zorro/zorro.py:1614-1712). */
int zr_align(zr_movie* mv);

int zr_synchronize(zr_movie* mv);

/* enable != 0: zr_align captures its whole kernel sequence on the movie's
 * stream into a HIP graph on its next call and replays that graph from then on
 * (three graph launches per movie: spectra, pair transform, the rest, so that
 * stage timing can bracket the pair transform).  The graphs hold the frames pointer: a different
 * device pointer in zr_movie_set_frames, a new stream or a change of timing
 * re-captures it.  Not available for triMode 'auto', whose solver reads the
 * pair list back to the host.  No reference counterpart (launch batching). */
int zr_movie_use_graph(zr_movie* mv, int enable);

/* Result read-back (host buffers, synchronising).
 * shifts f32 [N][N][2], corr/peaksig/origin f32 [N][N] (errorDict trimats).
 * Fails under preShift when a pair's disc held no positive correlation (the
 * reference's argmax then leaves the disc for the first zero of the map). */
int zr_get_trimats(zr_movie* mv, float* shifts, float* corr, float* peaksig, float* origin);
/* The frames the registration works on, f32 [N][H][W] (after the gain product
 * and binning of zr_movie_set_frames_raw: self.images after binStack,
 * zorro/zorro.py:1903). */
int zr_get_frames(zr_movie* mv, float* frames);
/* hotpixFilter results (hotpixInfo, zorro/zorro.py:2578-2602, :2637):
 * info10 = {mu, sigma (normal-CDF fit of the unaligned-sum histogram),
 * cutoffLower, cutoffUpper, guessDeadpix, guessHotpix, neighborPix, frameMean,
 * frameStd, fit objective}.  Fails when the filter has not run or the
 * histogram range was outside 2..65536 unit bins. */
int zr_get_hotpix_info(zr_movie* mv, double* info10);
/* Bmode='opti': the Gaussian width chosen for each pair, f32 [N][N] (the
 * reference only prints it, zorro/zorro.py:1123-1124). */
int zr_get_peak_sigmas(zr_movie* mv, float* sigma);
/* windows f32 [P][2m][2m] in canonical pair order (row-major over I<J). */
int zr_get_windows(zr_movie* mv, float* windows);
int zr_pair_count(zr_movie* mv, int* npairs);
/* Equations of the last zr_solve (the rows of shiftsSolver's Acoeff,
 * zorro/zorro.py:1296-1311): the length of weights and errorXY.  0 for the
 * triModes without a solver. */
int zr_equation_count(zr_movie* mv, int* m);
/* translations f64 [N][2] (after the centroid when originMode='centroid'),
 * relative_est f64 [N-1][2], weights f64 [M], logistic f64 [3] (thr, K, nu),
 * error f64 [M][2] (errorXY), error_stats f64 [2] (mean/std errorNorm),
 * error_xy_frames f64 [N][2] (errorX, errorY).  Any pointer may be NULL.
 * Fails when the equation graph is disconnected or the solver could not prove
 * optimality. */
int zr_get_solution(zr_movie* mv, double* translations, double* relative_est, double* weights, double* logistic,
                    double* error, double* error_stats, double* error_xy_frames);
/* cdfLogisticCurve's histogram of the positive peak significances
 * (zorro/zorro.py:1991-1998, errorDict keys pdfPeaks, cdfPeaks, hSigma):
 * counts as float32, their normalised cumulative sum and the left bin edges
 * (float32, numpy's binning), int(n/7) bins for n positive values.  Any of the
 * three arrays may be NULL; each holds at least `cap` floats; *nbins receives
 * the bin count (0 when weightMode is not 'autologistic').  Call after zr_solve. */
int zr_get_peaksig_hist(zr_movie* mv, float* pdf, float* cdf, float* hsigma, int cap, int* nbins);

/* The last 'background' step's fit: p[6] = (offset, centre x, centre y,
 * amplitude, width x, width y) in the fit's grid coordinates, the L1 error and
 * the status (0 converged, 1 evaluation limit, 2 iteration limit, 3 NaN: the
 * reference then leaves filtSum unchanged).  No counterpart: the reference
 * keeps only the subtracted result. */
int zr_get_background_fit(zr_movie* mv, double* p6, double* fval, int* status);
/* Solver diagnostics: info6 = {positive peak significances, histogram bins,
 * status y, status x (1 optimal vertex, 2 pivot limit, -1 disconnected),
 * simplex pivots y, pivots x}. */
int zr_get_solver_info(zr_movie* mv, int* info6);
/* imageSum, filtSum, even and odd half sums f32 [H][W]; any may be NULL. */
int zr_get_sums(zr_movie* mv, float* image_sum, float* filt_sum, float* even_sum, float* odd_sum);
/* FRC f64 [box/2]; n receives box/2. */
int zr_get_frc(zr_movie* mv, double* frc, int* n);

/* Per-stage timing with HIP events on the movie's stream; stage ids follow the
 * reference's bench dict (zorro/zorro.py:109: xcorr, solve, shifts, dose, frc). */
enum {
  ZR_STAGE_SPECTRA = 0,   /* forward spectra of every frame (xcorr0..2) */
  ZR_STAGE_PAIRS = 1,     /* pair transforms to the lag windows (xcorr2..3) */
  ZR_STAGE_PEAKS = 2,     /* B filter, peak, significance, sub-pixel (xcorr2..3) */
  ZR_STAGE_SOLVE = 3,     /* weights + shift solver (solve0..1) */
  ZR_STAGE_SHIFTS = 4,    /* shift + sums without the dose filter (shifts0..1) */
  ZR_STAGE_DOSE = 5,      /* shift + sums + dose filter (shifts, dose0..1) */
  ZR_STAGE_FRC = 6,       /* even/odd FRC (frc0..1) */
  ZR_STAGE_INPUT = 7,     /* gain product and Fourier binning (bin0..1) */
  ZR_STAGE_TAIL = 8,      /* graph replay: peaks + solve + shifts/dose + FRC as one graph */
  ZR_NSTAGES = 9
};
int zr_set_timing(zr_movie* mv, int enable);
/* ms[ZR_NSTAGES] accumulated milliseconds, counts[ZR_NSTAGES] calls; reset!=0 clears. */
int zr_get_stage_times(zr_movie* mv, double* ms, long* counts, int reset);
/* dims[12] = N, H, W, Hp, Wp, Hc, Wc, Wh, maxShift, pairs, unmasked, K1 batch. */
int zr_movie_info(zr_movie* mv, int* dims);

/* Device memory the movie holds (workspaces, tables, frames copied in by
 * zr_movie_set_frames): what a server admitting movies by HBM budget charges
 * for it.  No reference counterpart (the reference holds one movie per process). */
int zr_movie_device_bytes(zr_movie* mv, size_t* bytes);

/* Asynchronous device-to-device copies of imageSum, filtSum and the even/odd
 * half sums (f32 [H][W] device buffers, any may be NULL) on the movie's stream,
 * so that a reused movie plan can hand its results to a caller-owned buffer.
 * The results of applyShifts/doseFilter (zorro/zorro.py:1804, :2487-2494). */
int zr_copy_sums_device(zr_movie* mv, float* image_sum, float* filt_sum, float* even_sum, float* odd_sum);

/* tiledFRC of two images (zorro/zorro.py:2246-2310): Hann-apodised box x box
 * tiles at 1 - overlap steps (overlap clamped to 0.8) inside the crop limits
 * that translations trans f64 [n][2] (host, NULL: none) give (minCrop 5,
 * :2264-2274); per tile the FP64 sum of Re(E conj(O)) / sqrt(|O|^2 |E|^2) over
 * the float32 tile spectra; normalised by its [0,0] element, fftshifted and
 * rotationally averaged (rotmean, zorro/zorro_util.py:147-224).
 * img1, img2: device f32 [h][w]; frc: device f64 receiving the whole rotmean
 * curve, ceil(sqrt(2) box/2) + 1 values (lazyFouRingCorr keeps the first
 * box/2); *nrings (nullable) receives that count.  box 256, 512 or 1024.
 * Runs on `stream` (NULL: the default stream) and returns when the result is
 * written; allocates its workspace per call.  Fails with ZR_ERR_ARG when no
 * tile fits after the crop. */
int zr_tiled_frc(const float* img1, const float* img2, int h, int w, const double* trans, int n, int box,
                 double overlap, double* frc, int* nrings, void* stream);

/* zr_tiled_frc without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and result as zr_tiled_frc (the same bits), box 256 or 512 only
 * (ZR_ERR_UNSUPPORTED beyond: its workspace is a 28 MiB array that each device
 * the library runs on holds until the process ends).  Asynchronous on
 * `stream`: the result is ready when the stream has passed the call.  Calls
 * that overlap on one device share that workspace and race; serialise them. */
int zr_tiled_frc_noalloc(const float* img1, const float* img2, int h, int w, const double* trans, int n, int box,
                         double overlap, double* frc, int* nrings, void* stream);

/* Batches: several movies of one detector shape and one zr_params registered
 * together.  The reference registers one movie per process and scales by
 * running processes side by side (automator/zorroSkulkManager.py:1089-1113,
 * zorro/scripts/zorro_benchmark.py:174-248); a batch instead runs the forward
 * spectra of every frame, the pair transforms of every movie's pairs (one
 * ragged pair list) and the peak searches as single launches, then each
 * movie's solver, shift/filter and FRC stages on a stream of its own, so
 * that the one-workgroup solvers of the movies overlap.  Every movie's results
 * equal those of a zr_movie of its own.  triMode 'tri' or 'diag', float32
 * frames, no preShift, 'hot' or binning; 2..64 frames per movie. */
typedef struct zr_batch zr_batch;
int zr_batch_create(int device, const zr_params* p, int n_movies, const int* n_frames, int h, int w,
                    zr_batch** out);
void zr_batch_destroy(zr_batch* b);
/* Movie i of the batch as a zr_movie for the zr_get_* calls, zr_movie_info and
 * zr_copy_sums_device (owned by the batch: zr_movie_destroy ignores it). */
int zr_batch_movie(zr_batch* b, int i, zr_movie** view);
/* Device address of movie i's frames f32 [n_i][h][w] inside the batch (a
 * producer may write them there directly). */
int zr_batch_frames(zr_batch* b, int i, float** frames);
/* Copy movie i's frames in (host or device memory; returns when the copy is
 * done).  A device source written on another stream must be complete first. */
int zr_batch_set_frames(zr_batch* b, int i, const float* frames, int on_device);
/* enable != 0: the next zr_batch_align captures the whole batch sequence, forked
 * over the movies' streams and joined back, as one HIP graph and replays it. */
int zr_batch_use_graph(zr_batch* b, int enable);
/* alignImageStack of every movie of the batch (zorro/zorro.py:1569-1716). */
int zr_batch_align(zr_batch* b);
int zr_batch_synchronize(zr_batch* b);
int zr_batch_device_bytes(zr_batch* b, size_t* bytes);

/* Device pointers of results for zero-copy consumers (valid until destroy). */
int zr_device_results(zr_movie* mv, const float** image_sum, const float** filt_sum, const double** translations);

/* ---------------------------------------------------------------- serving
 * A movie server: registration requests of mixed frame counts, shapes and
 * settings on one device, several in flight at once.  Replaces the reference's
 * scale-out by one registration subprocess per movie under the Skulk manager
 * (automator/zorroSkulkManager.py:1089-1113; its per-movie work is
 * alignImageStack, zorro/zorro.py:1569-1716).  Requests queue in arrival order
 * and are admitted when an in-flight slot is free and the movie's plan fits the
 * HBM budget: an idle plan of the same settings and shape is reused, otherwise
 * one is created after evicting idle plans of other keys (least recently used
 * first); a movie that fits may go ahead of a larger head of the queue.  Plans
 * are created and movies run in the server's worker threads; each movie runs
 * on its plan's own stream, so movies overlap on the device.  Results do not
 * depend on what else runs.  A movie larger than the whole budget fails with
 * ZR_ERR_NOMEM. */
typedef struct zr_server zr_server;
enum { ZR_REQ_QUEUED = 0, ZR_REQ_RUNNING = 1, ZR_REQ_DONE = 2, ZR_REQ_FAILED = 3, ZR_REQ_CANCELLED = 4 };

/* One movie as a plan operation sees it (outputs written by run). */
typedef struct zr_serve_job {
  const zr_params* params;
  const void* frames;              /* f32 [n][h][w], device memory */
  int n, h, w;
  float* image_sum;                /* device f32 [h][w] or NULL */
  float* filt_sum;                 /* device f32 [h][w] or NULL */
  double* translations;            /* host f64 [n][2] */
  double* frc;                     /* host f64 [frc_cap] */
  int frc_cap;
  int* frc_n;
} zr_serve_job;

/* The plan operations the scheduler drives: the built-in ones create zr_movie
 * plans of this library (zr_server_create); callers may supply stand-ins (a
 * host-only test, another backend).  create returns the plan and the device
 * bytes it holds while a movie runs on it; estimate gives an upper bound for a
 * key before its first plan exists.  Operations run in the server's threads;
 * they return 0 or a ZR_ERR_* code. */
typedef struct zr_plan_ops {
  void* ctx;
  size_t (*estimate)(void* ctx, const zr_params* p, int n, int h, int w);
  int (*create)(void* ctx, const zr_params* p, int n, int h, int w, void** plan, size_t* bytes);
  int (*run)(void* ctx, void* plan, const zr_serve_job* job);
  void (*destroy)(void* ctx, void* plan);
} zr_plan_ops;

/* hbm_budget 0: 85% of the device's free memory now; max_inflight 1..64. */
int zr_server_create(int device, size_t hbm_budget, int max_inflight, zr_server** out);
/* A server over caller-supplied plan operations (hbm_budget must be given). */
int zr_server_create_ops(const zr_plan_ops* ops, size_t hbm_budget, int max_inflight, zr_server** out);
/* Finishes every queued and running request, frees every plan and the server. */
void zr_server_destroy(zr_server* s);
/* Queue one movie.  frames: device f32 [n][h][w] that the caller keeps alive
 * and unchanged until the request finishes; image_sum / filt_sum: device f32
 * [h][w] buffers the results are copied into (NULL: not wanted); arrival: the
 * arrival time in seconds of CLOCK_MONOTONIC (0: now), from which latency is
 * measured.  *id receives the request's id. */
int zr_server_submit(zr_server* s, const zr_params* p, int n, int h, int w, const void* frames, float* image_sum,
                     float* filt_sum, double arrival, long long* id);
/* Create up to `count` idle plans for movies of this key ahead of the requests;
 * *made receives how many fit the budget. */
int zr_server_warm(zr_server* s, const zr_params* p, int n, int h, int w, int count, int* made);
/* Cancel a request that is still queued (ZR_ERR_STATE once it is admitted). */
int zr_server_cancel(zr_server* s, long long id);
/* *state = ZR_REQ_* of a request. */
int zr_server_poll(zr_server* s, long long id, int* state);
/* Wait until the request has finished (done, failed or cancelled) or timeout_s
 * seconds have passed (< 0: no limit); *state as zr_server_poll. */
int zr_server_wait(zr_server* s, long long id, double timeout_s, int* state);
/* Wait for the next request to finish, in completion order; *id = 0 when none
 * finished within timeout_s. */
int zr_server_wait_any(zr_server* s, double timeout_s, long long* id, int* state);
/* Results of a finished request: translations f64 [n][2] (the request's n),
 * FRC f64 [*nfrc] (at most cap), times3 = arrival, admission and completion
 * (CLOCK_MONOTONIC seconds), the failure message.  Returns the request's own
 * code: 0 done, its ZR_ERR_* code when it failed, ZR_ERR_STATE when cancelled. */
int zr_server_result(zr_server* s, long long id, double* translations, int n, double* frc, int cap, int* nfrc,
                     double* times3, char* error, int error_cap);
/* Forget a finished request. */
int zr_server_release(zr_server* s, long long id);
/* out8 = admitted, plans created, plans evicted, peak bytes held, bytes held,
 * in flight, queued, failed. */
int zr_server_stats(zr_server* s, long long* out8);
/* The server's upper estimate of a movie plan's device bytes. */
size_t zr_server_estimate_bytes(const zr_params* p, int n, int h, int w);

#ifdef __cplusplus
}
#endif
#endif
