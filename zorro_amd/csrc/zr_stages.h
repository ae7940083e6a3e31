// Stages after the pair correlation: FP64 shift solver, shift + sums, tiled
// FRC and the dose filter.  State lives in zr_stage_state (zr_stages.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/zorro_hip.h"

struct zr_stage_state;

const char* zr_stage_error();
int zr_stage_create(zr_stage_state** out, const zr_params& p, int N, int H, int W, int Hp, int Wp,
                    const std::vector<int4>& pairs, int spectral);
void zr_stage_free(zr_stage_state* s);
// Replace the solver's equations with those of a new pair set (triMode 'auto'
// after its row-wise early stop); at most the pair count given at creation.
int zr_stage_set_pairs(zr_stage_state* s, const std::vector<int4>& pairs);
int zr_stage_equation_count(zr_stage_state* s);
size_t zr_stage_device_bytes(const zr_stage_state* s);
// asynchronous device-to-device copies of the sums (any pointer may be NULL)
int zr_stage_copy_sums_device(zr_stage_state* s, float* image_sum, float* filt_sum, float* even_sum, float* odd_sum,
                              hipStream_t st);
int zr_stage_solve(zr_stage_state* s, const float* shifts, const float* corr, const float* peaksig, hipStream_t st);
// 0 when the pairs' equations cannot be solved (zr_stage_solve then fails)
int zr_stage_equations_ok(const zr_stage_state* s);
int zr_stage_shifts(zr_stage_state* s, const float* frames, float* aligned_out, hipStream_t st);
int zr_stage_frc(zr_stage_state* s, hipStream_t st);
// crop limits / tile counts of the next zr_stage_frc from caller translations [n][2]
int zr_stage_set_frc_crop(zr_stage_state* s, const double* trans, int n, hipStream_t st);
// S: kx-major full-resolution half spectra [N][Wk][Hp] (used when the state was
// created with spectral != 0; may be NULL otherwise).
// normalize != 0: doseFilter(normalize=True) (filter sum + 1 in the denominator)
int zr_stage_dose(zr_stage_state* s, const float* frames, const float2* S, int Wk, int normalize, hipStream_t st);
// shiftMethod 'fourier': the aligned frames from the full-height half spectra S
// (slots = frames; mask_slot >= 0: the masked path's mask spectrum, added
// back with the frame means mu[2J] / (H W) that were taken out)
int zr_stage_fourier_align(zr_stage_state* s, const float2* S, int Wk, int Hs, int mask_slot, const double* mu,
                           hipStream_t st);
// 'gausslp' on filtSum (on imageSum when no filter step has run since the solve)
int zr_stage_gausslp(zr_stage_state* s, hipStream_t st);
// filterMode 'background': backgroundEstimate's fitted 2D Gaussian subtracted from filtSum
int zr_stage_background(zr_stage_state* s, hipStream_t st);
int zr_stage_get_background(zr_stage_state* s, double* p6, double* fval, int* status);
// cdfLogisticCurve's histogram (pdf counts, cdf, left edges); nbins = 0 without autologistic weights
int zr_stage_get_hist(zr_stage_state* s, float* pdf, float* cdf, float* hsigma, int cap, int* nbins);
int zr_stage_get_solution(zr_stage_state* s, double* translations, double* relative_est, double* weights,
                          double* logistic, double* error, double* error_stats, double* error_xy_frames);
int zr_stage_get_sums(zr_stage_state* s, float* image_sum, float* filt_sum, float* even_sum, float* odd_sum);
int zr_stage_get_frc(zr_stage_state* s, double* frc, int* n);
int zr_stage_solver_info(zr_stage_state* s, int* out6);
int zr_stage_device_results(zr_stage_state* s, const float** image_sum, const float** filt_sum,
                            const double** translations);

// Standalone tiledFRC of two device images [h][w] (row stride w): crop limits
// and tile counts from translations [n][2] (NULL: none) as tiledFRC derives
// them; workspace layout of one block of zr_frc_work_bytes(box, chunk) bytes,
// tiles transformed `chunk` at a time (the result does not depend on chunk);
// frc: device f64 [zr_frc_rings_count(box)] (rotmean's length).
size_t zr_frc_work_bytes(int box, int chunk);
int zr_frc_rings_count(int box);
int zr_frc_tiles_from_trans(const double* trans, int n, int h, int w, int box, double overlap, int crop[4],
                            int* tiles_y, int* tiles_x);
int zr_frc_tiles_run(const float* img1, const float* img2, int h, int w, const int crop[4], int tiles_y, int tiles_x,
                     int box, double overlap, void* work, int chunk, double* frc, hipStream_t st);
