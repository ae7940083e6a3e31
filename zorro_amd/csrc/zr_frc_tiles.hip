// Standalone tiled Fourier ring correlation of two device images: the C entry
// point of ImageRegistrator.tiledFRC (zorro/zorro.py:2246-2310).  The kernels
// and the workspace layout are those of the movie's FRC stage (zr_stages.hip);
// this file owns the calls that take two images instead of a movie.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>
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

// ---------------------------------------------------------------- no allocation
// Workspace of zr_tiled_frc_noalloc: one block per device that the library runs
// on, held until the process ends.  28 MiB holds the box-512 workspace with 4
// tiles per chunk (26.1 MiB: tables, Hann window, 4 tile row-transforms, the
// 16 group planes and the 2D FRC); smaller boxes take more tiles per chunk.
constexpr size_t ZR_FRC_SCRATCH_BYTES = (size_t)28 << 20;
constexpr int ZR_FRC_NOALLOC_MAX_BOX = 512;
static __device__ __align__(256) unsigned char scratch[ZR_FRC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once per
// device; each device holds its own copy of the array at its own address).
int scratch_for_current_device(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return frc_fail(ZR_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return frc_fail(ZR_ERR_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return frc_fail(ZR_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return ZR_OK;
}
}  // namespace

extern "C" int zr_tiled_frc_noalloc(const float* img1, const float* img2, int h, int w, const double* trans, int n,
                                    int box, double overlap, double* frc, int* nrings, void* stream) {
  int crop[4], ty = 0, tx = 0;
  int rc = frc_check(img1, img2, h, w, trans, n, box, overlap, frc, crop, &ty, &tx);
  if (rc) return rc;
  if (box > ZR_FRC_NOALLOC_MAX_BOX)
    return frc_fail(ZR_ERR_UNSUPPORTED, "tiledFRC without allocation: box %d exceeds %d (its %zu-byte workspace)",
                    box, ZR_FRC_NOALLOC_MAX_BOX, ZR_FRC_SCRATCH_BYTES);
  // as many tiles per chunk as the workspace holds (the sums do not depend on it)
  const int nt = ty * tx;
  int chunk = 1;
  while (chunk < nt && zr_frc_work_bytes(box, chunk + 1) <= ZR_FRC_SCRATCH_BYTES) ++chunk;
  if (zr_frc_work_bytes(box, chunk) > ZR_FRC_SCRATCH_BYTES)
    return frc_fail(ZR_ERR_UNSUPPORTED, "tiledFRC without allocation: box %d does not fit %zu bytes", box,
                    ZR_FRC_SCRATCH_BYTES);
  void* work = nullptr;
  if ((rc = scratch_for_current_device(&work))) return rc;
  rc = zr_frc_tiles_run(img1, img2, h, w, crop, ty, tx, box, overlap, work, chunk, frc,
                        static_cast<hipStream_t>(stream));
  if (rc) return frc_fail(rc, zr_stage_error());
  if (nrings) *nrings = zr_frc_rings_count(box);
  return ZR_OK;
}
