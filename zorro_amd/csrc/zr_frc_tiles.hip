// Standalone tiled Fourier ring correlation of two device images: the C entry
// point of ImageRegistrator.tiledFRC (zorro/zorro.py:2246-2310).  The kernels
// and the workspace layout are those of the movie's FRC stage (zr_stages.hip);
// this file owns the calls that take two images instead of a movie.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/zorro_hip.h"
#include "zr_stages.h"

int zr_capi_fail(int code, const char* msg);   // sets zr_last_error() (zr_capi.hip)

namespace {
int frc_fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return zr_capi_fail(code, buf);
}

// arguments shared by the tiledFRC entry points; fills crop and tile counts
int frc_check(const float* img1, const float* img2, int h, int w, const double* trans, int n, int box,
              double overlap, double* frc, int crop[4], int* ty, int* tx) {
  if (!img1 || !img2 || !frc) return frc_fail(ZR_ERR_ARG, "tiledFRC: NULL image or output");
  if (box != 256 && box != 512 && box != 1024)
    return frc_fail(ZR_ERR_UNSUPPORTED, "tiledFRC: box %d (built sizes 256, 512, 1024)", box);
  if (h < box || w < box) return frc_fail(ZR_ERR_ARG, "tiledFRC: images %dx%d smaller than the box %d", h, w, box);
  if (trans && n < 1) return frc_fail(ZR_ERR_ARG, "tiledFRC: translations given with n = %d", n);
  if (zr_frc_tiles_from_trans(trans, n, h, w, box, overlap, crop, ty, tx) != ZR_OK)
    return frc_fail(ZR_ERR_ARG, "tiledFRC: no %d-pixel tile fits %dx%d after the crop (%d,%d,%d,%d)", box, h, w,
                    crop[0], crop[1], crop[2], crop[3]);
  return ZR_OK;
}
}  // namespace

extern "C" int zr_tiled_frc(const float* img1, const float* img2, int h, int w, const double* trans, int n, int box,
                            double overlap, double* frc, int* nrings, void* stream) {
  int crop[4], ty = 0, tx = 0;
  int rc = frc_check(img1, img2, h, w, trans, n, box, overlap, frc, crop, &ty, &tx);
  if (rc) return rc;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const int chunk = ty * tx;                       // every tile at once
  void* work = nullptr;
  hipError_t e = hipMalloc(&work, zr_frc_work_bytes(box, chunk));
  if (e != hipSuccess) return frc_fail(ZR_ERR_NOMEM, "tiledFRC workspace: %s", hipGetErrorString(e));
  rc = zr_frc_tiles_run(img1, img2, h, w, crop, ty, tx, box, overlap, work, chunk, frc, st);
  e = hipStreamSynchronize(st);                    // the workspace is freed after the kernels ran
  hipFree(work);
  if (rc) return frc_fail(rc, zr_stage_error());
  if (e != hipSuccess) return frc_fail(ZR_ERR_HIP, "tiledFRC: %s", hipGetErrorString(e));
  if (nrings) *nrings = zr_frc_rings_count(box);
  return ZR_OK;
}
