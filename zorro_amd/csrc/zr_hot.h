// Outlier-pixel filter (hotpixFilter, zorro/zorro.py:2497-2662): arguments,
// device state and the host entry that enqueues its passes.
#pragma once
#include <hip/hip_runtime.h>

#include <stddef.h>

struct HotArgs {
  int N;                      // frames
  float logistic_k, relax;    // hotpixInfo logisticK, relax (float32 as in the reference)
  double max_sigma;           // hotpixInfo maxSigma: histogram half-range in Poisson sigmas
  double psf_k1, psf_k2;      // normalised PSF kernel: edge and corner weights
  double thres;               // neighbourThres
  double cut_lo_default, cut_hi_default;   // cut-offs when a line fit fails
  int fixed_cutoffs;          // 1: the two cut-offs as given, no fits
  double mu_fix, sigma_fix;   // > 0: fixed normal-CDF parameters instead of the fit
  int maxbins;                // histogram capacity
};

// Written by the passes on the device; read back by zr_get_hotpix_info.
struct HotState {
  double sum_mean, pstd, lo;  // mean of the unaligned sum, its sqrt, first bin edge
  int nbins, status;          // status 0: ok; 1: histogram range outside [2, maxbins] bins
  double mu, sigma, objective;
  double cut_lo, cut_hi;      // float32 values
  double best_mean;           // mu / N
  unsigned long long n_dead, n_hot, n_nbr;
};

struct HotBuffers {
  float *usum, *lmask, *nmask, *sp, *nf;   // [H][W] each
  double* part;                             // per-block sums of the unaligned sum
  unsigned* hist;                           // [maxbins]
  float* fitwork;                           // [5 maxbins]
  HotState* state;
};

size_t zr_hot_bytes(int H, int W, int maxbins);
// Enqueue the filter on `st`; rewrites `frames` [N][H][W] in place.
hipError_t zr_hot_filter(float* frames, int N, int H, int W, const HotArgs& a, const HotBuffers& b, hipStream_t st);
