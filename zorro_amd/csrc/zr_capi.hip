// C ABI of libzorro_hip.so: movie plans, device workspaces, stage sequencing.
// See include/zorro_hip.h for the contract of every entry point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/zorro_hip.h"
#include "zr_hot.h"
#include "zr_launch.h"
#include "zr_stages.h"

namespace {

thread_local std::string g_err;
thread_local size_t g_alloc_bytes = 0;   // device bytes allocated by this thread's current zr_movie_create

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define ZR_HIP(call)                                                                                \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess) return fail(ZR_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                      __FILE__, __LINE__);                                          \
  } while (0)

bool smooth235(int n) {
  if (n <= 0) return false;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  return n == 1;
}

bool supported_len(int n) {
  static const int ok[] = {256, 512, 1024, 2048, 3072, 3840, 4096, 5760};
  for (int v : ok)
    if (v == n) return true;
  return false;
}

// exp(-2 pi i t / n), t in [0, n)
std::vector<float2> fwd_table(int n) {
  std::vector<float2> t(n);
  for (int i = 0; i < n; ++i) {
    double a = -2.0 * M_PI * (double)i / (double)n;
    t[i] = make_float2((float)cos(a), (float)sin(a));
  }
  return t;
}

template <typename T>
int dev_alloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return ZR_OK;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) return fail(ZR_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
  g_alloc_bytes += count * sizeof(T);
  return ZR_OK;
}

template <typename T>
int dev_upload(T** p, const std::vector<T>& v) {
  int rc = dev_alloc(p, v.size());
  if (rc) return rc;
  if (!v.empty()) {
    hipError_t e = hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(ZR_ERR_HIP, "hipMemcpy upload failed: %s", hipGetErrorString(e));
  }
  return ZR_OK;
}

// Per-stage HIP-event timing on the movie's stream (accumulated on read-out).
struct StageTimer {
  struct Pending { int stage; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  double ms[ZR_NSTAGES] = {0};
  long count[ZR_NSTAGES] = {0};
  bool on = false;
  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
  }
  hipEvent_t begin(hipStream_t s) {
    if (!on) return nullptr;
    hipEvent_t e = get();
    hipEventRecord(e, s);
    return e;
  }
  void end(int stage, hipEvent_t a, hipStream_t s) {
    if (!on || !a) return;
    hipEvent_t b = get();
    hipEventRecord(b, s);
    pending.push_back({stage, a, b});
  }
  void harvest() {
    for (auto& p : pending) {
      hipEventSynchronize(p.b);
      float t = 0.f;
      hipEventElapsedTime(&t, p.a, p.b);
      ms[p.stage] += t;
      count[p.stage] += 1;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  ~StageTimer() {
    for (auto& p : pending) { hipEventDestroy(p.a); hipEventDestroy(p.b); }
    for (auto e : pool) hipEventDestroy(e);
  }
};

}  // namespace

struct zr_movie {
  std::vector<uint8_t> tri_mask;   // user triIndices [N][N] (empty: the triMode's pairs)
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  zr_params p{};
  int N = 0, H = 0, W = 0, Hp = 0, Wp = 0, Hc = 0, Wc = 0, Wh = 0, m = 0, Lper = 0;
  bool crop = false, unmasked = true;
  // spectra layout: Wk kept columns, Hs stored rows per column, row offset of the
  // cropped rows >= Hc/2, index of the +Hc/2 row inside a stored column (-1: Sside)
  bool spectral = false, full_spec = false;
  int Wk = 0, Hs = 0, hoff = 0, side_in_col = -1;
  int batch = 1, rb = 4;
  int k2_cpw = 8;                  // lag-window columns per wave of the pair transform (column-group mode)
  int k2_ppw = 16;                 // pair-list entries per wave (pair mode; 0: column-group mode)
  // frames
  const float* frames = nullptr;
  float* frames_owned = nullptr;
  // tables
  float2 *twWp2 = nullptr, *twWp = nullptr, *twHp = nullptr, *twHc = nullptr, *twWc = nullptr, *ktab = nullptr;
  float2* k2btab = nullptr;        // [zr_k2_btab_rows(Wh)][2m] twiddles of the matrix-core row pass (NULL: vector pass)
  float2* k2ltw = nullptr;         // [L+1][64] lane twiddles of the m <= L column pass (NULL: general pass)
  float *gy = nullptr, *gx = nullptr;
  int ry = 0, rx = 0;
  // pairs: P frame pairs; on the masked path the pair transform also runs the
  // 4N+1 mask-term pairs (Ptot in all), over nspec = 2N+1 spectra (F, Q, M)
  int P = 0, Ptot = 0, nspec = 0;
  int win_out0 = 0, nwin = 0;      // first window of the pairs' results, windows allocated
  int4 *pairs = nullptr, *work = nullptr;
  std::vector<int4> pairs_host;
  // workspaces
  float2 *T = nullptr, *S = nullptr, *Sside = nullptr, *Z = nullptr;
  double *part = nullptr, *stats = nullptr, *den = nullptr, *stats_c = nullptr;
  float* windows = nullptr;
  float *shifts = nullptr, *corr = nullptr, *peaksig = nullptr, *origin = nullptr, *bsigma = nullptr;
  bool have_xcorr = false;
  size_t bytes = 0;                // device memory held (workspaces, tables, owned frames)
  zr_stage_state* st = nullptr;   // solver / shift / FRC / dose state (zr_stages.h)
  // input stage (zr_bin.hip): raw frames Hr x Wr, Fourier-binned to H x W when
  // shape_binned is set; gain product fused with the conversion
  int Hr = 0, Wr = 0, nbb = 1;
  bool binned = false;
  int Mrow = 0, Mcol = 0, Mout = 0;
  float2 *twMrow = nullptr, *twMcol = nullptr, *twMout = nullptr;
  float2 *chW_c = nullptr, *chW_B = nullptr, *chH_c = nullptr, *chH_B = nullptr;
  float2 *chHb_c = nullptr, *chHb_B = nullptr, *chWb_c = nullptr, *chWb_B = nullptr;
  float2 *binA = nullptr, *binB = nullptr;   // [nbb][Hr][W] / [nbb][W][Hr] (then Gc/Gr [nbb][W][H], [nbb][H][W])
  // binning by exactly two (Hr = 2 H, Wr = 2 W, H and W even): the polyphase
  // form (zr_launch_bin2) with its convolution size, kernel spectra and sums
  bool bin2 = false;
  int M2 = 0;
  float2 *bin2Gx = nullptr, *bin2Gy = nullptr, *twM2 = nullptr;
  double *bin2_cpart = nullptr, *bin2_q = nullptr;
  void* raw_owned = nullptr;
  size_t raw_owned_bytes = 0;
  float* gain_owned = nullptr;
  // preShift: pairs in diagonal order (J - I ascending), diagonal offsets/counts,
  // pair index of (I, J), window centres, shifts in cropped pixels, lag-0
  // column sums [P][Wh], and the flag of a disc without positive values
  bool preshift = false;
  int disc = 0;
  std::vector<std::pair<int, int>> diags;
  int4* dpairs = nullptr;
  int* pidx = nullptr;
  int2* ctr = nullptr;
  float *pshift = nullptr, *s0 = nullptr;
  int* pflag = nullptr;
  // outlier-pixel filter (zr_hot.hip), one device block carved into its buffers
  void* hot_block = nullptr;
  HotBuffers hb{};
  HotArgs ha{};
  bool hot_ran = false;
  // zr_align as one HIP graph (zr_movie_use_graph)
  bool use_graph = false;
  hipGraph_t graph[3] = {nullptr, nullptr, nullptr};          // spectra | pair transform | the rest
  hipGraphExec_t graph_exec[3] = {nullptr, nullptr, nullptr};
  StageTimer timer;
  // batches (zr_batch_*): the batch's correlation plan has no solver state;
  // a movie of a batch is a view that borrows the batch's frames and spectra
  bool xcorr_only = false;
  bool view = false;
  bool fourier_aligned = false;          // shiftMethod 'fourier': aligned frames formed since the last solve
  float* umask = nullptr;                // user mask [H][W] (params user_mask), 0 or 1
  bool umask_set = false;
  hipStream_t batch_stream = nullptr;   // a view's results are complete once this stream has passed them
};

// wait for a movie's work (and, for a movie of a batch, the batch's graph)
static hipError_t sync_movie(zr_movie* mv) {
  hipError_t e = hipStreamSynchronize(mv->stream);
  if (e == hipSuccess && mv->batch_stream) e = hipStreamSynchronize(mv->batch_stream);
  return e;
}

extern "C" const char* zr_last_error(void) { return g_err.c_str(); }

// error reporting for the entry points defined in other source files
int zr_capi_fail(int code, const char* msg) { return fail(code, "%s", msg); }

extern "C" int zr_device_count(int* n) {
  if (!n) return fail(ZR_ERR_ARG, "n is NULL");
  ZR_HIP(hipGetDeviceCount(n));
  return ZR_OK;
}

static void drop_graph(zr_movie* mv) {
  for (int i = 0; i < 3; ++i) {
    if (mv->graph_exec[i]) hipGraphExecDestroy(mv->graph_exec[i]);
    if (mv->graph[i]) hipGraphDestroy(mv->graph[i]);
    mv->graph_exec[i] = nullptr;
    mv->graph[i] = nullptr;
  }
}

static void free_movie(zr_movie* mv) {
  if (!mv) return;
  hipSetDevice(mv->device);
  drop_graph(mv);
  if (mv->view) {             // borrowed from the batch
    mv->S = nullptr;
    mv->frames_owned = nullptr;
  }
  void* ptrs[] = {mv->frames_owned, mv->k2btab, mv->k2ltw, mv->twWp2, mv->twWp, mv->twHp, mv->twHc, mv->twWc, mv->ktab, mv->gy, mv->gx,
                  mv->pairs, mv->work, mv->T, mv->S, mv->Sside, mv->Z, mv->part, mv->stats, mv->den, mv->stats_c,
                  mv->windows, mv->shifts, mv->corr, mv->peaksig, mv->origin, mv->bsigma, mv->twMrow, mv->twMcol,
                  mv->twMout, mv->chW_c, mv->chW_B, mv->chH_c, mv->chH_B, mv->chHb_c, mv->chHb_B, mv->chWb_c,
                  mv->chWb_B, mv->binA, mv->binB, mv->raw_owned, mv->gain_owned, mv->hot_block, mv->dpairs,
                  mv->pidx, mv->ctr, mv->pshift, mv->s0, mv->pflag, mv->umask, mv->bin2Gx, mv->bin2Gy, mv->twM2,
                  mv->bin2_cpart, mv->bin2_q};
  for (void* q : ptrs)
    if (q) hipFree(q);
  zr_stage_free(mv->st);
  if (mv->own_stream) hipStreamDestroy(mv->own_stream);
  delete mv;
}

extern "C" void zr_movie_destroy(zr_movie* mv) {
  if (mv && mv->view) return;   // a movie of a batch goes with zr_batch_destroy
  free_movie(mv);
}

// triMode 'refine' (zorro/zorro.py:934-956): frame I (I < N-1) against the
// template (sum_k F_k - F_I) / N.  By linearity its raw window is
// (W[sum, I] - W[I, I]) / N, so the pair transform runs the pairs (sum, I) and
// (I, I) (windows 2I, 2I+1), plus on the masked path the mask terms of a pair
// and the template's Ct, T from the sums (k_refine_combine's layout).  Results
// go to windows win_out0 + I, read by the peak stage as the pair (I, I).
static int build_refine_pairs(zr_movie* mv) {
  const int N = mv->N;
  std::vector<int4> wk, pr;
  const int sumF = mv->unmasked ? N : 2 * N + 1;
  for (int k = 0; k < N - 1; ++k) {
    wk.push_back(make_int4(sumF, k, 2 * k, 0));
    wk.push_back(make_int4(k, k, 2 * k + 1, 0));
  }
  if (mv->unmasked) {
    mv->nspec = N + 1;
  } else {
    const int b = 2 * (N - 1), Ms = 2 * N;
    for (int k = 0; k < N; ++k) wk.push_back(make_int4(k, Ms, b + k, 0));
    for (int k = 0; k < N; ++k) wk.push_back(make_int4(Ms, k, b + N + k, 0));
    for (int k = 0; k < N; ++k) wk.push_back(make_int4(N + k, Ms, b + 2 * N + k, 0));
    for (int k = 0; k < N; ++k) wk.push_back(make_int4(Ms, N + k, b + 3 * N + k, 0));
    wk.push_back(make_int4(Ms, Ms, b + 4 * N, 0));
    wk.push_back(make_int4(2 * N + 1, Ms, b + 4 * N + 1, 0));
    wk.push_back(make_int4(2 * N + 2, Ms, b + 4 * N + 2, 0));
    mv->nspec = 2 * N + 3;
  }
  mv->Ptot = (int)wk.size();
  mv->win_out0 = mv->Ptot;
  for (int k = 0; k < N - 1; ++k) pr.push_back(make_int4(k, k, mv->win_out0 + k, 0));
  mv->P = (int)pr.size();
  mv->nwin = mv->Ptot + mv->P;
  mv->pairs_host = pr;
  int rc = dev_upload(&mv->pairs, pr);
  if (rc) return rc;
  return dev_upload(&mv->work, wk);
}

static int build_pairs(zr_movie* mv) {
  const zr_params& p = mv->p;
  const int N = mv->N;
  std::vector<int4> pr;
  // triIndices of __init_xcorrnm2 (zorro/zorro.py:559-581); 'auto' correlates
  // the band J - I <= autoMax, beyond which every row has stopped (:992-996)
  if (p.tri_mode == ZR_TRI_AUTO && p.auto_max < 1) return fail(ZR_ERR_ARG, "autoMax must be >= 1");
  const bool masked_pairs = !mv->tri_mask.empty();
  for (int I = 0; I < N; ++I)
    for (int J = I + 1; J < N; ++J) {
      int d = J - I;
      bool take = false;
      if (masked_pairs) take = mv->tri_mask[(size_t)I * N + J] != 0;
      else if (p.tri_mode == ZR_TRI_TRI) take = true;
      else if (p.tri_mode == ZR_TRI_DIAG) {
        int wdt = p.diag_width < 0 ? N : p.diag_width;
        take = (d <= wdt + p.diag_start) && (d > p.diag_start);
      } else if (p.tri_mode == ZR_TRI_AUTO) take = d <= p.auto_max;
      else if (p.tri_mode == ZR_TRI_FIRST) take = (I == 0);
      else if (p.tri_mode == ZR_TRI_REFINE) take = false;
      else return fail(ZR_ERR_UNSUPPORTED, "triMode %d is not built (tri, diag, auto, first, refine)", p.tri_mode);
      if (take) pr.push_back(make_int4(I, J, (int)pr.size(), 0));
    }
  if (p.tri_mode == ZR_TRI_REFINE) return build_refine_pairs(mv);
  if (pr.empty()) return fail(ZR_ERR_ARG, "no frame pairs to correlate");
  mv->P = (int)pr.size();
  mv->pairs_host = pr;
  // work order: 8x8 frame tiles so concurrently running blocks share spectra in L2
  std::vector<int4> wk = pr;
  std::stable_sort(wk.begin(), wk.end(), [](const int4& a, const int4& b) {
    int ta = (a.x / 8) * 1024 + a.y / 8, tb = (b.x / 8) * 1024 + b.y / 8;
    return ta < tb;
  });
  mv->Ptot = mv->P;
  mv->nspec = N;
  if (!mv->unmasked) {
    // mask-term pairs over the spectra F_k (slot k), Q_k (slot N+k), M (slot 2N):
    // Ct_k = (F_k, M), Cb_k = (M, F_k), T_k = (Q_k, M), B_k = (M, Q_k), mp = (M, M)
    const int P = mv->P, Ms = 2 * N;
    for (int k = 0; k < N; ++k) wk.push_back(make_int4(k, Ms, P + k, 0));
    for (int k = 0; k < N; ++k) wk.push_back(make_int4(Ms, k, P + N + k, 0));
    for (int k = 0; k < N; ++k) wk.push_back(make_int4(N + k, Ms, P + 2 * N + k, 0));
    for (int k = 0; k < N; ++k) wk.push_back(make_int4(Ms, N + k, P + 3 * N + k, 0));
    wk.push_back(make_int4(Ms, Ms, P + 4 * N, 0));
    mv->Ptot = P + 4 * N + 1;
    mv->nspec = 2 * N + 1;
  }
  mv->nwin = mv->Ptot;
  int rc = dev_upload(&mv->pairs, pr);
  if (rc) return rc;
  return dev_upload(&mv->work, wk);
}

static PeakParamsHost peak_params(const zr_movie* mv) {
  PeakParamsHost hp;
  hp.m = mv->m;
  hp.suppress = mv->p.suppress_origin ? 1 : 0;
  hp.bconv = (mv->p.b_mode == ZR_B_CONV && mv->p.b_rad > 0) ? 1 : 0;
  hp.bopti = (mv->p.b_mode == ZR_B_OPTI) ? 1 : 0;
  // 'opti' searches sigma in [0, 7]: radius up to int(4 * 7 + 0.5) = 28
  hp.ry = hp.bopti ? 28 : mv->ry;
  hp.rx = hp.bopti ? 28 : mv->rx;
  hp.subpix = mv->p.sub_pix_reg;
  hp.disc = mv->disc;
  hp.scale_y = mv->crop ? (float)((double)mv->Hp / (double)mv->Hc) : 1.f;
  hp.scale_x = mv->crop ? (float)((double)mv->Wp / (double)mv->Wc) : 1.f;
  hp.n = mv->N;
  return hp;
}

static int create_movie(int device, const zr_params* p, int n_frames, int h, int w, const uint8_t* tri_indices,
                        zr_movie** out, bool xcorr_only = false) {
  if (!p || !out) return fail(ZR_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (n_frames < 2 || h < 16 || w < 16) return fail(ZR_ERR_ARG, "need >= 2 frames of >= 16x16 (got %d x %d x %d)", n_frames, h, w);
  if (tri_indices) {
    // xcorrnm2_tri(triIndices=...) (zorro/zorro.py:552-557): pairs I < J of the
    // mask, correlated and solved as triMode 'tri' does
    if (p->tri_mode != ZR_TRI_TRI)
      return fail(ZR_ERR_UNSUPPORTED, "a triIndices mask runs with triMode 'tri' (got %d)", p->tri_mode);
    if (p->pre_shift)
      return fail(ZR_ERR_UNSUPPORTED, "preShift with a triIndices mask (the reference warns that it breaks, zorro/zorro.py:549-550)");
    for (int I = 0; I < n_frames; ++I)
      for (int J = 0; J <= I; ++J)
        if (tri_indices[(size_t)I * n_frames + J])
          return fail(ZR_ERR_UNSUPPORTED, "triIndices[%d][%d]: only pairs above the diagonal (I < J) are built", I, J);
  }
  // filterMode steps (zorro/zorro.py:1689-1711): a legacy do_dose alone is the
  // single 'dose' step; do_dose then marks the presence of any dose step
  zr_params pn = *p;
  if (pn.n_filters == 0 && pn.do_dose) { pn.n_filters = 1; pn.filters[0] = ZR_FILTER_DOSE; }
  if (pn.n_filters < 0 || pn.n_filters > 4) return fail(ZR_ERR_ARG, "n_filters %d outside 0..4", pn.n_filters);
  pn.do_dose = 0;
  for (int i = 0; i < pn.n_filters; ++i) {
    const int f = pn.filters[i];
    if (f != ZR_FILTER_DOSE && f != ZR_FILTER_DOSENORM && f != ZR_FILTER_GAUSSLP && f != ZR_FILTER_BACKGROUND)
      return fail(ZR_ERR_UNSUPPORTED, "filter step %d: %d is not a built filterMode step", i, f);
    if (f == ZR_FILTER_DOSE || f == ZR_FILTER_DOSENORM) pn.do_dose = 1;
  }
  for (int i = pn.n_filters; i < 4; ++i) pn.filters[i] = 0;
  p = &pn;
  zr_movie* mv = new zr_movie();
  if (tri_indices) mv->tri_mask.assign(tri_indices, tri_indices + (size_t)n_frames * n_frames);
  g_alloc_bytes = 0;
  mv->device = device;
  mv->p = *p;
  mv->N = n_frames;
  mv->Hr = h;
  mv->Wr = w;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) { free_movie(mv); return fail(ZR_ERR_HIP, "hipSetDevice(%d): %s", device, hipGetErrorString(e)); }
  // binStack (zorro/zorro.py:1591-1592, :1826-1904): everything after it sees
  // the binned frames, and the pixel size grows by the mean bin factor
  mv->binned = p->shape_binned[0] > 0 || p->shape_binned[1] > 0;
  if (mv->binned) {
    const int hb = p->shape_binned[0], wb = p->shape_binned[1];
    if (hb <= 0 || wb <= 0 || hb > h || wb > w || (hb & 1) || (wb & 1)) {
      free_movie(mv);
      return fail(ZR_ERR_ARG, "shapeBinned (%d,%d) must be even and within the frames (%d,%d)", hb, wb, h, w);
    }
    if (h > 8192 || w > 8192) { free_movie(mv); return fail(ZR_ERR_UNSUPPORTED, "binning built for frames up to 8192 (got %dx%d)", h, w); }
    h = hb;
    w = wb;
    mv->p.pixelsize_nm = p->pixelsize_nm * 0.5 * ((double)mv->Hr / hb + (double)mv->Wr / wb);
  }
  mv->H = h;
  mv->W = w;
  if (h < 16 || w < 16) { free_movie(mv); return fail(ZR_ERR_ARG, "binned frames below 16x16"); }
  mv->Hp = p->shape_padded[0] > 0 ? p->shape_padded[0] : h;
  mv->Wp = p->shape_padded[1] > 0 ? p->shape_padded[1] : w;
  if (mv->Hp < h || mv->Wp < w) { free_movie(mv); return fail(ZR_ERR_ARG, "shapePadded (%d,%d) smaller than frames (%d,%d)", mv->Hp, mv->Wp, h, w); }
  // padStack (zorro/zorro.py:1908-1958): frames smaller than shapePadded are
  // zero-padded and correlated under the mask of the frame region
  mv->unmasked = (mv->Hp == h && mv->Wp == w);
  if (p->user_mask) {
    // one user mask [H][W] for every frame (masks.shape[0] == 1): the masked
    // correlation under the frame region times that mask
    if (mv->binned) { free_movie(mv); return fail(ZR_ERR_UNSUPPORTED, "a user mask with shapeBinned is not built"); }
    if (p->shift_method == ZR_SHIFT_FOURIER) { free_movie(mv); return fail(ZR_ERR_UNSUPPORTED, "a user mask with shiftMethod 'fourier' is not built"); }
    mv->unmasked = false;
  }
  mv->crop = (p->fou_crop[0] != 0 || p->fou_crop[1] != 0);
  if (mv->crop) {
    mv->Hc = 2 * (p->fou_crop[0] / 2);
    mv->Wc = 2 * (p->fou_crop[1] / 2);
    if (mv->Hc > mv->Hp || mv->Wc > mv->Wp || mv->Hc <= 0 || mv->Wc <= 0) { free_movie(mv); return fail(ZR_ERR_ARG, "fouCrop must be in (0, shapePadded]"); }
  } else {
    mv->Hc = mv->Hp;
    mv->Wc = mv->Wp;
  }
  if (!supported_len(mv->Wp) || !supported_len(mv->Hp)) { free_movie(mv); return fail(ZR_ERR_UNSUPPORTED, "shapePadded %dx%d: built FFT lengths are 256..4096 (2^a), 3072, 3840, 5760", mv->Hp, mv->Wp); }
  {
    const int l = mv->Hc / 64;
    if (mv->Hc % 64 != 0 || !(l == 4 || l == 8 || l == 16 || l == 32 || l == 48 || l == 64)) { free_movie(mv); return fail(ZR_ERR_UNSUPPORTED, "cropped height %d: built sizes are 256, 512, 1024, 2048, 3072, 4096", mv->Hc); }
  }
  if (!smooth235(mv->Wc)) { free_movie(mv); return fail(ZR_ERR_UNSUPPORTED, "cropped width %d is not 2^a3^b5^c", mv->Wc); }
  mv->Lper = mv->Hc / 64;
  mv->Wh = mv->Wc / 2 + 1;
  mv->m = p->max_shift;
  if (mv->m <= 0 || 2 * mv->m > 128 || 2 * mv->m > mv->Hc || 2 * mv->m > mv->Wc) { free_movie(mv); return fail(ZR_ERR_UNSUPPORTED, "maxShift %d: built range is 1..64 (and 2*maxShift <= cropped size)", mv->m); }
  mv->disc = mv->m;
  mv->preshift = p->pre_shift != 0;
  if (mv->preshift) {
    // the disc of each pair moves with the previous pair's shift: windows of
    // maxShift + 8 (sub-pixel patch) + the Gaussian radius around the disc keep
    // every filtered value the disc and the patch use free of window edges
    if (!mv->unmasked) { free_movie(mv); return fail(ZR_ERR_UNSUPPORTED, "preShift is built for unpadded frames (no mask)"); }
    if (p->tri_mode == ZR_TRI_REFINE) { free_movie(mv); return fail(ZR_ERR_UNSUPPORTED, "preShift with triMode 'refine' is not working in the reference (zorro/zorro.py:1150)"); }
    if (p->b_mode == ZR_B_OPTI) { free_movie(mv); return fail(ZR_ERR_UNSUPPORTED, "preShift with Bmode 'opti' (whole-map objective) is not built"); }
    if (mv->Hc != mv->Wc) { free_movie(mv); return fail(ZR_ERR_UNSUPPORTED, "preShift needs square correlation maps (the reference's mask does not broadcast otherwise, zorro/zorro.py:842-844)"); }
    int rg = 0;
    if (p->b_mode == ZR_B_CONV && p->b_rad > 0)
      for (int ax = 0; ax < 2; ++ax)
        rg = std::max(rg, (int)(4.0 * ((ax == 0 ? mv->Hp : mv->Wp) / (std::sqrt(2.0) * M_PI * p->b_rad)) + 0.5));
    const int mw = ((p->max_shift + 8 + rg + 7) / 8) * 8;
    if (mw > 64 || 2 * mw > mv->Hc) { free_movie(mv); return fail(ZR_ERR_UNSUPPORTED, "preShift needs windows of maxShift + 8 + Gaussian radius = %d <= 64 lags (maxShift %d, radius %d)", p->max_shift + 8 + rg, p->max_shift, rg); }
    mv->m = mw;
  }
  if (p->sub_pix_reg > 1 && (p->sub_pix_reg & (p->sub_pix_reg - 1)) != 0) { free_movie(mv); return fail(ZR_ERR_UNSUPPORTED, "subPixReg must be a power of two"); }
  if (p->sub_pix_reg > 32) { free_movie(mv); return fail(ZR_ERR_UNSUPPORTED, "subPixReg > 32 is not built"); }
  int rc;
  e = hipStreamCreateWithFlags(&mv->own_stream, hipStreamNonBlocking);
  if (e != hipSuccess) { free_movie(mv); return fail(ZR_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
  mv->stream = mv->own_stream;
  // tables
  if ((rc = dev_upload(&mv->twWp2, fwd_table(mv->Wp / 2))) || (rc = dev_upload(&mv->twWp, fwd_table(mv->Wp))) ||
      (rc = dev_upload(&mv->twHp, fwd_table(mv->Hp))) || (rc = dev_upload(&mv->twHc, fwd_table(mv->Hc))) ||
      (rc = dev_upload(&mv->twWc, fwd_table(mv->Wc)))) { free_movie(mv); return rc; }
  {
    int P = p->sub_pix_reg > 1 ? p->sub_pix_reg : 1;
    int G = 16 * P;
    std::vector<float2> kt(G);
    for (int t = 0; t < G; ++t) {
      double re = 0, im = 0;
      for (int f = -8; f < 8; ++f) {
        double a = 2.0 * M_PI * f * t / G;
        re += cos(a);
        im += sin(a);
      }
      kt[t] = make_float2((float)re, (float)im);
    }
    if ((rc = dev_upload(&mv->ktab, kt))) { free_movie(mv); return rc; }
  }
  {
    // scipy.ndimage.gaussian_filter weights: radius int(4 sigma + 0.5), exp(-x^2/(2 sigma^2)) / sum
    std::vector<float> wy(1, 1.f), wx(1, 1.f);
    if (p->b_mode == ZR_B_CONV && p->b_rad > 0) {
      for (int ax = 0; ax < 2; ++ax) {
        double sig = (ax == 0 ? mv->Hp : mv->Wp) / (std::sqrt(2.0) * M_PI * p->b_rad);
        int r = (int)(4.0 * sig + 0.5);
        std::vector<double> k(2 * r + 1);
        double s = 0;
        for (int i = -r; i <= r; ++i) { k[i + r] = std::exp(-0.5 / (sig * sig) * i * i); s += k[i + r]; }
        std::vector<float>& o = ax == 0 ? wy : wx;
        o.resize(2 * r + 1);
        for (int i = 0; i <= 2 * r; ++i) o[i] = (float)(k[i] / s);
        (ax == 0 ? mv->ry : mv->rx) = r;
      }
    }
    if ((rc = dev_upload(&mv->gy, wy)) || (rc = dev_upload(&mv->gx, wx))) { free_movie(mv); return rc; }
  }
  {
    const PeakParamsHost hp = peak_params(mv);
    const size_t lds = zr_peaks_lds_bytes(&hp, nullptr);
    if (lds > (size_t)ZR_LDS_BYTES) {
      free_movie(mv);
      return fail(ZR_ERR_UNSUPPORTED, "peak search needs %zu bytes of LDS (> %d): maxShift %d with a %d-tap Gaussian "
                  "(Brad %g) does not fit; raise Brad or lower maxShift", lds, ZR_LDS_BYTES, mv->m,
                  2 * std::max(mv->ry, mv->rx) + 1, p->b_rad);
    }
  }
  if ((rc = build_pairs(mv))) { free_movie(mv); return rc; }
  if (mv->preshift) {
    std::vector<int4> dp = mv->pairs_host;
    std::stable_sort(dp.begin(), dp.end(), [](const int4& a, const int4& b) { return a.y - a.x < b.y - b.x; });
    for (size_t i = 0; i < dp.size();) {
      size_t j = i;
      while (j < dp.size() && dp[j].y - dp[j].x == dp[i].y - dp[i].x) ++j;
      mv->diags.push_back({(int)i, (int)(j - i)});
      i = j;
    }
    std::vector<int> pi((size_t)mv->N * mv->N, -1);
    for (const int4& q : mv->pairs_host) pi[(size_t)q.x * mv->N + q.y] = q.z;
    if ((rc = dev_upload(&mv->dpairs, dp)) || (rc = dev_upload(&mv->pidx, pi)) ||
        (rc = dev_alloc(&mv->ctr, (size_t)mv->P)) || (rc = dev_alloc(&mv->pshift, (size_t)2 * mv->P)) ||
        (rc = dev_alloc(&mv->s0, (size_t)mv->P * mv->Wh)) || (rc = dev_alloc(&mv->pflag, 1))) {
      free_movie(mv);
      return rc;
    }
  }
  // workspaces
  const size_t N = mv->N;
  mv->batch = (int)std::min<size_t>(N, p->k1_batch > 0 ? p->k1_batch : 8);
  if (p->k2_cols_per_wave < 0 || p->k2_cols_per_wave > 64) { free_movie(mv); return fail(ZR_ERR_ARG, "k2_cols_per_wave %d outside 0..64", p->k2_cols_per_wave); }
  mv->k2_cpw = p->k2_cols_per_wave > 0 ? p->k2_cols_per_wave : 8;
  if (p->k2_pairs_per_wave < 0 || p->k2_pairs_per_wave > 64) { free_movie(mv); return fail(ZR_ERR_ARG, "k2_pairs_per_wave %d outside 0..64", p->k2_pairs_per_wave); }
  // explicit columns per wave select the column-group walk; else pair mode
  mv->k2_ppw = p->k2_pairs_per_wave > 0 ? p->k2_pairs_per_wave : (p->k2_cols_per_wave > 0 ? 0 : 16);
  mv->rb = zr_k1_rows_per_block(mv->Wp);
  const int nrb = mv->Hp / mv->rb;
  const bool fourier = p->shift_method == ZR_SHIFT_FOURIER && p->sub_pix_reg > 1;
  mv->spectral = p->do_dose && mv->unmasked && mv->H == mv->W && mv->Hp <= 4096 && !fourier;
  // full-height half spectra: kept for the spectral dose path and for Fourier shifts
  mv->full_spec = mv->spectral || fourier;
  mv->Wk = mv->full_spec ? mv->Wp / 2 + 1 : mv->Wh;
  mv->Hs = mv->full_spec ? mv->Hp : mv->Hc;
  mv->hoff = mv->full_spec ? mv->Hp - mv->Hc : 0;
  mv->side_in_col = mv->full_spec ? mv->Hc / 2 : -1;
  const size_t nspec = (size_t)mv->nspec;
  const size_t nrb_sums = (size_t)(mv->H + 7) / 8;   // k1_frame_sums row blocks
  // K1 intermediate: column pairs, an odd Wk padded by one zero column
  if ((rc = dev_alloc(&mv->T, (size_t)mv->batch * ((mv->Wk + 1) / 2) * 2 * mv->Hp)) ||
      (rc = dev_alloc(&mv->S, nspec * mv->Wk * mv->Hs)) ||
      (rc = dev_alloc(&mv->Sside, mv->full_spec ? 0 : nspec * mv->Wk)) ||
      (rc = dev_alloc(&mv->Z, (size_t)mv->Ptot * mv->Wh * 2 * mv->m)) ||
      (rc = dev_alloc(&mv->part, N * std::max<size_t>(nrb, nrb_sums) * 2)) ||
      (rc = dev_alloc(&mv->stats, N * 2)) || (rc = dev_alloc(&mv->stats_c, N * 2)) ||
      (rc = dev_alloc(&mv->den, (size_t)mv->Ptot)) ||
      (rc = dev_alloc(&mv->windows, (size_t)mv->nwin * 4 * mv->m * mv->m)) || (rc = dev_alloc(&mv->shifts, N * N * 2)) ||
      (rc = dev_alloc(&mv->corr, N * N)) || (rc = dev_alloc(&mv->peaksig, N * N)) || (rc = dev_alloc(&mv->origin, N * N)) ||
      (rc = dev_alloc(&mv->bsigma, N * N)) ||
      (rc = dev_alloc(&mv->umask, p->user_mask ? (size_t)mv->H * mv->W : 0))) {
    free_movie(mv);
    return rc;
  }
  if ((2 * mv->m) % 16 == 0) {
    if ((rc = dev_alloc(&mv->k2btab, (size_t)zr_k2_btab_rows(mv->Wh) * 2 * mv->m))) { free_movie(mv); return rc; }
    e = zr_launch_k2_twiddles(mv->Wh, mv->Wc, mv->m, mv->k2btab, mv->own_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(mv->own_stream);
    if (e != hipSuccess) { free_movie(mv); return fail(ZR_ERR_HIP, "k2 twiddles: %s", hipGetErrorString(e)); }
  }
  if (p->do_hot) {
    // hotpixFilter's inputs (zorro/zorro.py:2510-2519, :2606-2609): PSF
    // kernel [[p2, p1, p2], [p1, 0, p1], [p2, p1, p2]] / sum in float32
    const float p1 = p->hot_psf_k2 ? 0.173235968f : 1.f, p2 = p1 * p1;
    const float ksum = (((p2 + p1) + (p2 + p1)) + ((0.f + p1) + (p2 + p1))) + p2;
    HotArgs& ha = mv->ha;
    ha.N = mv->N;
    ha.logistic_k = (float)p->hot_logistic_k;
    ha.relax = (float)p->hot_relax;
    ha.max_sigma = p->hot_max_sigma;
    ha.psf_k1 = (double)(p1 / ksum);
    ha.psf_k2 = (double)(p2 / ksum);
    ha.thres = p->hot_neighbour_thres;
    ha.cut_lo_default = (double)(float)p->hot_cutoff_lower;
    ha.cut_hi_default = (double)(float)p->hot_cutoff_upper;
    ha.fixed_cutoffs = p->hot_fixed_cutoffs ? 1 : 0;
    ha.mu_fix = p->hot_mu_sigma[0];
    ha.sigma_fix = p->hot_mu_sigma[1];
    ha.maxbins = 65536;
    if (!(ha.max_sigma > 0.0) || !(p->hot_logistic_k > 0.0)) {
      free_movie(mv);
      return fail(ZR_ERR_ARG, "hotpixInfo: maxSigma and logisticK must be positive");
    }
    const size_t hw = (size_t)mv->H * mv->W, nblk = (hw + 255) / 256;
    const size_t bytes = zr_hot_bytes(mv->H, mv->W, ha.maxbins);
    e = hipMalloc(&mv->hot_block, bytes);
    if (e != hipSuccess) { free_movie(mv); return fail(ZR_ERR_NOMEM, "hipMalloc of %zu bytes failed", bytes); }
    g_alloc_bytes += bytes;
    char* q = (char*)mv->hot_block;
    auto take = [&](size_t n) { char* r = q; q += (n + 255) & ~(size_t)255; return (void*)r; };
    HotBuffers& hb = mv->hb;
    hb.usum = (float*)take(hw * 4);
    hb.lmask = (float*)take(hw * 4);
    hb.nmask = (float*)take(hw * 4);
    hb.sp = (float*)take(hw * 4);
    hb.nf = (float*)take(hw * 4);
    hb.part = (double*)take(nblk * 8);
    hb.hist = (unsigned*)take((size_t)ha.maxbins * 4);
    hb.fitwork = (float*)take((size_t)5 * ha.maxbins * 4);
    hb.state = (HotState*)take(sizeof(HotState));
    if ((size_t)(q - (char*)mv->hot_block) > bytes) { free_movie(mv); return fail(ZR_ERR_ARG, "hot buffer layout"); }
  }
  if (zr_k2_cols_r2_ok(mv->Lper, mv->m) && mv->Hc == 64 * mv->Lper) {
    if ((rc = dev_alloc(&mv->k2ltw, (size_t)(mv->Lper + 1) * 64))) { free_movie(mv); return rc; }
    e = zr_launch_k2_lane_twiddles(mv->Lper, mv->Hc, mv->k2ltw, mv->own_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(mv->own_stream);
    if (e != hipSuccess) { free_movie(mv); return fail(ZR_ERR_HIP, "k2 lane twiddles: %s", hipGetErrorString(e)); }
  }
  if (!mv->unmasked) {
    // the pair transform's own scale is nc2 / max(1, den): den = 0 keeps the
    // raw nc2 * Re R[.] windows that the masked combination needs
    e = hipMemset(mv->den, 0, sizeof(double) * mv->Ptot);
    if (e != hipSuccess) { free_movie(mv); return fail(ZR_ERR_HIP, "hipMemset: %s", hipGetErrorString(e)); }
  }
  if (mv->binned && mv->Hr == 2 * mv->H && mv->Wr == 2 * mv->W && mv->H % 2 == 0 && mv->W % 2 == 0) {
    // binning by two: half-sample delay kernels along x and y, one convolution size
    mv->bin2 = true;
    mv->M2 = zr_bin_conv_size(std::max(mv->H, mv->W));
    mv->nbb = std::min(mv->N, 4);
    std::vector<float2> Gx, Gy;
    zr_bin2_kernel_spectrum(mv->W, mv->M2, Gx);
    zr_bin2_kernel_spectrum(mv->H, mv->M2, Gy);
    const size_t per = (size_t)mv->Hr * mv->W;   // floats of Z and its transpose per frame
    if ((rc = dev_upload(&mv->bin2Gx, Gx)) || (rc = dev_upload(&mv->bin2Gy, Gy)) ||
        (rc = dev_upload(&mv->twM2, fwd_table(mv->M2))) ||
        (rc = dev_alloc(&mv->bin2_cpart, (size_t)mv->nbb * (mv->Hr / 2))) || (rc = dev_alloc(&mv->bin2_q, (size_t)mv->nbb)) ||
        (rc = dev_alloc(&mv->binA, (per * mv->nbb + 1) / 2)) || (rc = dev_alloc(&mv->binB, (per * mv->nbb + 1) / 2)) ||
        (rc = dev_alloc(&mv->frames_owned, (size_t)mv->N * mv->H * mv->W))) {
      free_movie(mv);
      return rc;
    }
  } else if (mv->binned) {
    // Bluestein tables: rows at the raw width, columns at the raw height with the
    // inverse at the binned height (same convolution size), out rows at the binned width
    mv->Mrow = zr_bin_conv_size(mv->Wr);
    mv->Mcol = zr_bin_conv_size(mv->Hr);
    mv->Mout = zr_bin_conv_size(mv->W);
    std::vector<float2> c, B;
    auto tables = [&](int L, int M, float2** dc, float2** dB) {
      zr_bin_chirp_tables(L, M, c, B);
      int r = dev_upload(dc, c);
      return r ? r : dev_upload(dB, B);
    };
    if ((rc = tables(mv->Wr, mv->Mrow, &mv->chW_c, &mv->chW_B)) || (rc = tables(mv->Hr, mv->Mcol, &mv->chH_c, &mv->chH_B)) ||
        (rc = tables(mv->H, mv->Mcol, &mv->chHb_c, &mv->chHb_B)) || (rc = tables(mv->W, mv->Mout, &mv->chWb_c, &mv->chWb_B)) ||
        (rc = dev_upload(&mv->twMrow, fwd_table(mv->Mrow))) || (rc = dev_upload(&mv->twMcol, fwd_table(mv->Mcol))) ||
        (rc = dev_upload(&mv->twMout, fwd_table(mv->Mout)))) {
      free_movie(mv);
      return rc;
    }
    mv->nbb = std::min(mv->N, 4);
    const size_t per = (size_t)mv->Hr * mv->W;   // >= W * H of the column pass
    if ((rc = dev_alloc(&mv->binA, per * mv->nbb)) || (rc = dev_alloc(&mv->binB, per * mv->nbb)) ||
        (rc = dev_alloc(&mv->frames_owned, (size_t)mv->N * mv->H * mv->W))) {
      free_movie(mv);
      return rc;
    }
  }
  mv->xcorr_only = xcorr_only;
  if (!xcorr_only &&
      (rc = zr_stage_create(&mv->st, mv->p, mv->N, mv->H, mv->W, mv->Hp, mv->Wp, mv->pairs_host, mv->spectral ? 1 : 0))) {
    free_movie(mv);
    return fail(rc, "%s", zr_stage_error());
  }
  mv->bytes = g_alloc_bytes + (mv->st ? zr_stage_device_bytes(mv->st) : 0);
  *out = mv;
  return ZR_OK;
}

extern "C" int zr_movie_create(int device, const zr_params* p, int n_frames, int h, int w, zr_movie** out) {
  return create_movie(device, p, n_frames, h, w, nullptr, out);
}

extern "C" int zr_movie_create_pairs(int device, const zr_params* p, int n_frames, int h, int w,
                                     const uint8_t* tri_indices, zr_movie** out) {
  if (!tri_indices) return fail(ZR_ERR_ARG, "NULL triIndices");
  return create_movie(device, p, n_frames, h, w, tri_indices, out);
}

extern "C" int zr_movie_device_bytes(zr_movie* mv, size_t* bytes) {
  if (!mv || !bytes) return fail(ZR_ERR_ARG, "NULL argument");
  *bytes = mv->bytes;
  return ZR_OK;
}

extern "C" int zr_copy_sums_device(zr_movie* mv, float* image_sum, float* filt_sum, float* even_sum, float* odd_sum) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  ZR_HIP(hipSetDevice(mv->device));
  int rc = zr_stage_copy_sums_device(mv->st, image_sum, filt_sum, even_sum, odd_sum, mv->stream);
  return rc ? fail(rc, "%s", zr_stage_error()) : ZR_OK;
}

extern "C" int zr_movie_set_stream(zr_movie* mv, void* s) {
  if (mv && mv->view) return fail(ZR_ERR_STATE, "a movie of a batch takes its frames and stream from the batch");
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  drop_graph(mv);
  mv->stream = s ? (hipStream_t)s : mv->own_stream;
  return ZR_OK;
}

// Input stage: raw frames (device) -> frames_owned: Fourier binning with the
// gain product fused into its row pass, or the gain product alone.
static int run_input_stage(zr_movie* mv, const void* raw, int dtype, const float* gain, double gain_scalar) {
  hipStream_t s = mv->stream;
  const int use_scale = (gain_scalar != 0.0 && std::fabs(gain_scalar - 1.0) > 1e-8) ? 1 : 0;
  const float gs = use_scale ? (float)gain_scalar : 1.f;
  hipEvent_t t0 = mv->timer.begin(s);
  if (!mv->binned) {
    const size_t npix = (size_t)mv->H * mv->W;
    ZR_HIP(zr_launch_gain_convert(raw, dtype, npix, npix * mv->N, gain, gs, use_scale, mv->frames_owned, s));
  } else {
    const int Hr = mv->Hr, Wr = mv->Wr, H = mv->H, W = mv->W;
    if (mv->bin2) {
      for (int f0 = 0; f0 < mv->N; f0 += mv->nbb) {
        const int nb = std::min(mv->nbb, mv->N - f0);
        ZR_HIP(zr_launch_bin2(mv->M2, raw, dtype, f0, nb, Hr, Wr, gain, gs, use_scale, mv->bin2Gx, mv->bin2Gy, mv->twM2,
                              (float*)mv->binA, (float*)mv->binB, mv->bin2_cpart, mv->bin2_q, mv->frames_owned, s));
      }
      mv->timer.end(ZR_STAGE_INPUT, t0, s);
      mv->frames = mv->frames_owned;
      mv->have_xcorr = false;
      return ZR_OK;
    }
    const BinChirp cw{mv->chW_c, mv->chW_B, Wr}, ch{mv->chH_c, mv->chH_B, Hr}, chb{mv->chHb_c, mv->chHb_B, H},
        cwb{mv->chWb_c, mv->chWb_B, W};
    const float scale = (float)(1.0 / ((double)H * (double)W));
    for (int f0 = 0; f0 < mv->N; f0 += mv->nbb) {
      const int nb = std::min(mv->nbb, mv->N - f0);
      ZR_HIP(zr_launch_bin_rows(mv->Mrow, raw, dtype, f0, nb, Hr, Wr, gain, gs, use_scale, cw, mv->twMrow, W, mv->binA, s));
      ZR_HIP(zr_launch_bin_transpose(mv->binA, nb, Hr, W, mv->binB, s));
      ZR_HIP(zr_launch_bin_cols(mv->Mcol, mv->binB, nb, Hr, W, H, ch, chb, mv->twMcol, scale, mv->binA, s));
      ZR_HIP(zr_launch_bin_transpose(mv->binA, nb, W, H, mv->binB, s));
      ZR_HIP(zr_launch_bin_out(mv->Mout, mv->binB, nb, H, W, cwb, mv->twMout, f0, mv->frames_owned, s));
    }
  }
  mv->timer.end(ZR_STAGE_INPUT, t0, s);
  mv->frames = mv->frames_owned;
  mv->have_xcorr = false;
  return ZR_OK;
}

extern "C" int zr_movie_set_frames_raw(zr_movie* mv, const void* frames, int dtype, int on_device, const float* gain,
                                       int gain_on_device, double gain_scalar) {
  if (mv && mv->view) return fail(ZR_ERR_STATE, "a movie of a batch takes its frames from the batch");
  if (!mv || !frames) return fail(ZR_ERR_ARG, "NULL argument");
  if (dtype < ZR_DTYPE_F32 || dtype > ZR_DTYPE_I16) return fail(ZR_ERR_ARG, "dtype %d", dtype);
  if (dtype == ZR_DTYPE_U4 && (mv->Wr & 1))
    return fail(ZR_ERR_ARG, "4-bit packed frames need an even width (w = %d)", mv->Wr);
  ZR_HIP(hipSetDevice(mv->device));
  const size_t npix = (size_t)mv->N * mv->Hr * mv->Wr;
  const size_t raw_bytes = dtype == ZR_DTYPE_F32 ? npix * 4
                         : dtype == ZR_DTYPE_U8 ? npix : dtype == ZR_DTYPE_U4 ? npix / 2 : npix * 2;
  const void* rawd = frames;
  if (!on_device) {
    if (mv->raw_owned_bytes < raw_bytes) {
      if (mv->raw_owned) { hipFree(mv->raw_owned); mv->bytes -= mv->raw_owned_bytes; }
      mv->raw_owned = nullptr;
      mv->raw_owned_bytes = 0;
      ZR_HIP(hipMalloc(&mv->raw_owned, raw_bytes));
      mv->raw_owned_bytes = raw_bytes;
      mv->bytes += raw_bytes;
    }
    ZR_HIP(hipMemcpyAsync(mv->raw_owned, frames, raw_bytes, hipMemcpyHostToDevice, mv->stream));
    rawd = mv->raw_owned;
  }
  const float* gd = gain;
  if (gain && !gain_on_device) {
    const size_t gb = (size_t)mv->Hr * mv->Wr * sizeof(float);
    if (!mv->gain_owned) {
      ZR_HIP(hipMalloc((void**)&mv->gain_owned, gb));
      mv->bytes += gb;
    }
    ZR_HIP(hipMemcpyAsync(mv->gain_owned, gain, gb, hipMemcpyHostToDevice, mv->stream));
    gd = mv->gain_owned;
  }
  if (!mv->frames_owned) {
    const size_t fb = (size_t)mv->N * mv->H * mv->W * sizeof(float);
    ZR_HIP(hipMalloc((void**)&mv->frames_owned, fb));
    mv->bytes += fb;
  }
  // binning by two reads pixel pairs (and gain pairs) with one load each
  if (mv->bin2 && (((uintptr_t)rawd & 7) || ((uintptr_t)gd & 7)))
    return fail(ZR_ERR_ARG, "raw frames and gain reference must be 8-byte aligned device memory");
  int rc = run_input_stage(mv, rawd, dtype, gd, gain_scalar);
  if (rc) return rc;
  if (!on_device || (gain && !gain_on_device)) ZR_HIP(sync_movie(mv));   // host buffers may go away
  return ZR_OK;
}

extern "C" int zr_movie_set_frames(zr_movie* mv, const float* frames, int on_device) {
  if (mv && mv->view) return fail(ZR_ERR_STATE, "a movie of a batch takes its frames and stream from the batch");
  if (!mv || !frames) return fail(ZR_ERR_ARG, "NULL argument");
  if (mv->binned) return zr_movie_set_frames_raw(mv, frames, ZR_DTYPE_F32, on_device, nullptr, 0, 1.0);
  ZR_HIP(hipSetDevice(mv->device));
  if (on_device) {
    if (frames != mv->frames) drop_graph(mv);   // the graph holds the frames pointer
    mv->frames = frames;
  } else {
    size_t bytes = (size_t)mv->N * mv->H * mv->W * sizeof(float);
    if (!mv->frames_owned) {
      ZR_HIP(hipMalloc((void**)&mv->frames_owned, bytes));
      mv->bytes += bytes;
    }
    ZR_HIP(hipMemcpyAsync(mv->frames_owned, frames, bytes, hipMemcpyHostToDevice, mv->stream));
    mv->frames = mv->frames_owned;
  }
  mv->have_xcorr = false;
  return ZR_OK;
}

extern "C" int zr_movie_set_mask(zr_movie* mv, const float* mask, int on_device) {
  if (!mv || !mask) return fail(ZR_ERR_ARG, "NULL argument");
  if (!mv->p.user_mask || !mv->umask) return fail(ZR_ERR_STATE, "the plan was created without user_mask");
  ZR_HIP(hipSetDevice(mv->device));
  const size_t n = (size_t)mv->H * mv->W;
  if (!on_device) {
    size_t ones = 0;
    for (size_t i = 0; i < n; ++i) {
      if (mask[i] != 0.f && mask[i] != 1.f) return fail(ZR_ERR_ARG, "mask values must be 0 or 1 (element %zu is %g)", i, (double)mask[i]);
      ones += mask[i] != 0.f;
    }
    if (ones == 0) return fail(ZR_ERR_ARG, "the mask has no ones");
  }
  ZR_HIP(hipMemcpyAsync(mv->umask, mask, n * sizeof(float), on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                        mv->stream));
  if (!on_device) ZR_HIP(hipStreamSynchronize(mv->stream));   // the caller may reuse its buffer on return
  mv->umask_set = true;
  mv->have_xcorr = false;
  return ZR_OK;
}

static int xcorr_pairs(zr_movie* mv, double nc2, bool refine);
static int xcorr_peaks(zr_movie* mv);

// zr_xcorr in three parts (spectra, pair transform, peaks); a graph-replayed
// movie captures the first two as their own graphs so that the pair transform
// can be timed by events recorded between graph launches.
static int xcorr_part(zr_movie* mv, int part) {
  if (mv->p.user_mask && !mv->umask_set)
    return fail(ZR_ERR_STATE, "the plan takes a user mask: call zr_movie_set_mask before aligning");
  hipStream_t s = mv->stream;
  int rb = 0;
  const double nc2 = 1.0 / std::pow((double)mv->Hc * (double)mv->Wc, 2.0);
  const bool refine = mv->p.tri_mode == ZR_TRI_REFINE;
  if (part == 1) return xcorr_pairs(mv, nc2, refine);
  if (part == 2) return xcorr_peaks(mv);
  if (mv->p.do_hot) {
    // hotpixFilter before the spectra, after binning (zorro/zorro.py:1591-1596)
    hipEvent_t th = mv->timer.begin(s);
    ZR_HIP(zr_hot_filter(const_cast<float*>(mv->frames), mv->N, mv->H, mv->W, mv->ha, mv->hb, s));
    mv->timer.end(ZR_STAGE_INPUT, th, s);
    mv->hot_ran = true;
  }
  hipEvent_t t0 = mv->timer.begin(s);
  const int full = mv->full_spec ? 1 : 0;
  const size_t slot = (size_t)mv->Wk * mv->Hs;
  if (mv->unmasked) {
    // K1: spectra in batches of frames; frame sums on the way
    for (int f0 = 0; f0 < mv->N; f0 += mv->batch) {
      int nb = std::min(mv->batch, mv->N - f0);
      ZR_HIP(zr_launch_k1_rows(mv->Wp, mv->frames, mv->H, mv->W, mv->Hp, mv->Wk, f0, nb, mv->twWp2, mv->twWp, mv->T,
                               mv->part, 0, nullptr, nullptr, &rb, s));
      ZR_HIP(zr_launch_k1_cols(mv->Hp, mv->T, mv->Wk, mv->Hc, f0, nb, mv->twHp, mv->S, mv->Sside, full, s));
    }
    ZR_HIP(zr_launch_k1_stats(mv->part, mv->Hp / rb, mv->N, mv->stats, s));
    if (refine) {
      ZR_HIP(zr_launch_spec_sum(mv->S, slot, 0, mv->N, mv->N, s));
      if (!full) ZR_HIP(zr_launch_spec_sum(mv->Sside, (size_t)mv->Wk, 0, mv->N, mv->N, s));
      ZR_HIP(zr_launch_refine_den(mv->stats, mv->N, (double)mv->Hp * (double)mv->Wp, nc2, mv->den, s));
    } else {
      ZR_HIP(zr_launch_k2_den(mv->pairs, mv->P, mv->stats, (double)mv->Hp * (double)mv->Wp, nc2, mv->den, s));
    }
  } else {
    // masked: frame means first, then the spectra of m(f - mu) (slots 0..N),
    // (m(f - mu))^2 (slots N..2N) and the mask m (slot 2N).  'refine' centres
    // every frame by the mean over all frames instead (k_refine_combine).
    int nrb = 0;
    ZR_HIP(zr_launch_k1_sums(mv->frames, mv->N, mv->H, mv->W, mv->part, &nrb, s));
    ZR_HIP(zr_launch_k1_stats(mv->part, nrb, mv->N, mv->stats, s));
    const double* mu = mv->stats;
    if (refine) {
      ZR_HIP(zr_launch_common_mean(mv->stats, mv->N, mv->stats_c, s));
      mu = mv->stats_c;
    }
    for (int f0 = 0; f0 < mv->N; f0 += mv->batch) {
      int nb = std::min(mv->batch, mv->N - f0);
      for (int sq = 0; sq < 2; ++sq) {
        ZR_HIP(zr_launch_k1_rows(mv->Wp, mv->frames, mv->H, mv->W, mv->Hp, mv->Wk, f0, nb, mv->twWp2, mv->twWp,
                                 mv->T, mv->part, sq ? 2 : 1, mu, mv->umask, &rb, s));
        ZR_HIP(zr_launch_k1_cols(mv->Hp, mv->T, mv->Wk, mv->Hc, f0 + sq * mv->N, nb, mv->twHp, mv->S, mv->Sside,
                                 full, s));
      }
    }
    ZR_HIP(zr_launch_k1_rows(mv->Wp, mv->frames, mv->H, mv->W, mv->Hp, mv->Wk, 0, 1, mv->twWp2, mv->twWp, mv->T,
                             mv->part, 3, mv->stats, mv->umask, &rb, s));
    ZR_HIP(zr_launch_k1_cols(mv->Hp, mv->T, mv->Wk, mv->Hc, 2 * mv->N, 1, mv->twHp, mv->S, mv->Sside, full, s));
    if (refine) {
      for (int q = 0; q < 2; ++q) {
        ZR_HIP(zr_launch_spec_sum(mv->S, slot, q * mv->N, mv->N, 2 * mv->N + 1 + q, s));
        ZR_HIP(zr_launch_spec_sum(mv->Sside, (size_t)mv->Wk, q * mv->N, mv->N, 2 * mv->N + 1 + q, s));
      }
    }
  }
  mv->timer.end(ZR_STAGE_SPECTRA, t0, s);
  return ZR_OK;
}

// preShift: the pairs run diagonal by diagonal (J - I = 1, 2, ...): each pair's
// window is centred on the rounded shift of (I, J-1) found by the previous
// diagonal, then the pair transform and the peak search run for that diagonal.
static int xcorr_preshift(zr_movie* mv, double nc2) {
  hipStream_t s = mv->stream;
  ZR_HIP(zr_launch_fill_zero(mv->shifts, sizeof(float) * 2 * mv->N * mv->N, s));
  ZR_HIP(zr_launch_fill_zero(mv->corr, sizeof(float) * mv->N * mv->N, s));
  ZR_HIP(zr_launch_fill_zero(mv->peaksig, sizeof(float) * mv->N * mv->N, s));
  ZR_HIP(zr_launch_fill_zero(mv->origin, sizeof(float) * mv->N * mv->N, s));
  ZR_HIP(zr_launch_fill_zero(mv->pflag, sizeof(int), s));
  const PeakParamsHost hp = peak_params(mv);
  const PeakPreShift ps{mv->ctr, mv->pshift, mv->pflag};
  for (const auto& dg : mv->diags) {
    const int4* dp = mv->dpairs + dg.first;
    const int cnt = dg.second;
    ZR_HIP(zr_launch_preshift_centres(dp, cnt, mv->pidx, mv->pshift, mv->N, mv->ctr, s));
    if (mv->k2ltw)
      ZR_HIP(zr_launch_k2_cols_r2(mv->Lper, mv->S, mv->Sside, dp, cnt, mv->Wh, mv->Hc, mv->Wc, mv->m, 8, 0, 1, mv->k2ltw,
                                  mv->Z, mv->Wk, mv->Hs, mv->hoff, mv->side_in_col, mv->ctr, mv->twHc, mv->twWc,
                                  mv->s0, s));
    else
      ZR_HIP(zr_launch_k2_cols(mv->Lper, mv->S, mv->Sside, dp, cnt, mv->Wh, mv->Hc, mv->Wc, mv->m, 8, 1, mv->twHc,
                               mv->Z, mv->Wk, mv->Hs, mv->hoff, mv->side_in_col, mv->ctr, mv->twWc, mv->s0, s));
    ZR_HIP(zr_launch_k2_rows(mv->Z, dp, cnt, mv->Wh, mv->Wc, mv->m, mv->twWc, mv->k2btab, mv->den, (float)nc2,
                             mv->windows, dp, s));
    ZR_HIP(zr_launch_peaks(mv->windows, dp, cnt, &hp, mv->gy, mv->gx, mv->ktab, mv->shifts, mv->corr, mv->peaksig,
                           mv->origin, mv->bsigma, &ps, s));
  }
  ZR_HIP(zr_launch_origin_sums(mv->pairs, mv->P, mv->s0, mv->Wh, mv->den, nc2, mv->N, mv->origin, s));
  return ZR_OK;
}

static int xcorr_pairs(zr_movie* mv, double nc2, bool refine) {
  hipStream_t s = mv->stream;
  hipEvent_t t1 = mv->timer.begin(s);
  if (mv->preshift) {
    int rc = xcorr_preshift(mv, nc2);
    if (rc) return rc;
    mv->timer.end(ZR_STAGE_PAIRS, t1, s);
    return ZR_OK;
  }
  if (mv->k2ltw)
    ZR_HIP(zr_launch_k2_cols_r2(mv->Lper, mv->S, mv->Sside, mv->work, mv->Ptot, mv->Wh, mv->Hc, mv->Wc, mv->m,
                                mv->k2_cpw, mv->k2_ppw, mv->unmasked ? 1 : 0, mv->k2ltw, mv->Z, mv->Wk, mv->Hs, mv->hoff,
                                mv->side_in_col, nullptr, nullptr, nullptr, nullptr, s));
  else
    ZR_HIP(zr_launch_k2_cols(mv->Lper, mv->S, mv->Sside, mv->work, mv->Ptot, mv->Wh, mv->Hc, mv->Wc, mv->m, mv->k2_cpw,
                             mv->unmasked ? 1 : 0, mv->twHc, mv->Z, mv->Wk, mv->Hs, mv->hoff, mv->side_in_col, nullptr,
                             nullptr, nullptr, s));
  ZR_HIP(zr_launch_k2_rows(mv->Z, mv->pairs, mv->Ptot, mv->Wh, mv->Wc, mv->m, mv->twWc, mv->k2btab, mv->den,
                           (float)nc2, mv->windows, nullptr, s));
  if (refine)
    ZR_HIP(zr_launch_refine_combine(mv->windows, mv->N, mv->m, mv->unmasked ? 0 : 1, mv->win_out0, mv->stats_c,
                                    (double)mv->H * (double)mv->W, s));
  else if (!mv->unmasked) ZR_HIP(zr_launch_mnxc_combine(mv->windows, mv->pairs, mv->P, mv->N, mv->m, s));
  mv->timer.end(ZR_STAGE_PAIRS, t1, s);
  return ZR_OK;
}

static int xcorr_peaks(zr_movie* mv) {
  hipStream_t s = mv->stream;
  hipEvent_t t2 = mv->timer.begin(s);
  if (!mv->preshift) {     // (preShift ran its peak searches diagonal by diagonal)
    // cleared by a kernel (not hipMemsetAsync) so that the sequence can be captured in a graph
    ZR_HIP(zr_launch_fill_zero(mv->shifts, sizeof(float) * 2 * mv->N * mv->N, s));
    ZR_HIP(zr_launch_fill_zero(mv->corr, sizeof(float) * mv->N * mv->N, s));
    ZR_HIP(zr_launch_fill_zero(mv->peaksig, sizeof(float) * mv->N * mv->N, s));
    ZR_HIP(zr_launch_fill_zero(mv->origin, sizeof(float) * mv->N * mv->N, s));
    const PeakParamsHost hp = peak_params(mv);
    ZR_HIP(zr_launch_peaks(mv->windows, mv->pairs, mv->P, &hp, mv->gy, mv->gx, mv->ktab, mv->shifts, mv->corr,
                           mv->peaksig, mv->origin, mv->bsigma, nullptr, s));
  }
  if (mv->p.tri_mode == ZR_TRI_AUTO)
    ZR_HIP(zr_launch_auto_prune(mv->shifts, mv->corr, mv->peaksig, mv->origin, mv->N, mv->p.auto_max,
                                mv->p.peaksig_thres, s));
  mv->timer.end(ZR_STAGE_PEAKS, t2, s);
  mv->have_xcorr = true;
  return ZR_OK;
}

extern "C" int zr_xcorr(zr_movie* mv) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  if (mv->view) return fail(ZR_ERR_STATE, "a movie of a batch is correlated by zr_batch_align");
  if (!mv->frames) return fail(ZR_ERR_STATE, "frames not set");
  ZR_HIP(hipSetDevice(mv->device));
  int rc;
  for (int part = 0; part < 3; ++part)
    if ((rc = xcorr_part(mv, part))) return rc;
  return ZR_OK;
}

extern "C" int zr_solve(zr_movie* mv) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  if (!mv->have_xcorr) return fail(ZR_ERR_STATE, "zr_solve before zr_xcorr");
  ZR_HIP(hipSetDevice(mv->device));
  if (mv->p.tri_mode == ZR_TRI_AUTO) {
    // the equations follow the pairs that survived the row-wise stop:
    // triIndices = corrTriMat[:-1, 1:] != 0 (zorro/zorro.py:1282)
    std::vector<float> corr((size_t)mv->N * mv->N);
    ZR_HIP(hipMemcpyAsync(corr.data(), mv->corr, corr.size() * sizeof(float), hipMemcpyDeviceToHost, mv->stream));
    ZR_HIP(sync_movie(mv));
    std::vector<int4> pr;
    for (int I = 0; I < mv->N - 1; ++I)
      for (int J = I + 1; J < mv->N; ++J)
        if (corr[(size_t)I * mv->N + J] != 0.f) pr.push_back(make_int4(I, J, (int)pr.size(), 0));
    int rc = zr_stage_set_pairs(mv->st, pr);
    if (rc) return fail(rc, "%s", zr_stage_error());
  }
  hipEvent_t t0 = mv->timer.begin(mv->stream);
  int rc = zr_stage_solve(mv->st, mv->shifts, mv->corr, mv->peaksig, mv->stream);
  mv->timer.end(ZR_STAGE_SOLVE, t0, mv->stream);
  mv->fourier_aligned = false;
  return rc ? fail(rc, "%s", zr_stage_error()) : ZR_OK;
}

extern "C" int zr_set_trimats(zr_movie* mv, const float* shifts, const float* corr, const float* peaksig) {
  if (!mv || !shifts || !corr || !peaksig) return fail(ZR_ERR_ARG, "NULL argument");
  ZR_HIP(hipSetDevice(mv->device));
  const size_t nn = (size_t)mv->N * mv->N;
  ZR_HIP(hipMemcpyAsync(mv->shifts, shifts, nn * 2 * sizeof(float), hipMemcpyHostToDevice, mv->stream));
  ZR_HIP(hipMemcpyAsync(mv->corr, corr, nn * sizeof(float), hipMemcpyHostToDevice, mv->stream));
  ZR_HIP(hipMemcpyAsync(mv->peaksig, peaksig, nn * sizeof(float), hipMemcpyHostToDevice, mv->stream));
  ZR_HIP(hipMemsetAsync(mv->origin, 0, nn * sizeof(float), mv->stream));
  ZR_HIP(sync_movie(mv));   // the host arrays may go away on return
  mv->have_xcorr = true;
  return ZR_OK;
}

// shiftMethod 'fourier': the aligned frames from the resident full-height
// spectra, once per solve, before the stages that read them
static int ensure_fourier(zr_movie* mv) {
  if (mv->p.shift_method != ZR_SHIFT_FOURIER || mv->p.sub_pix_reg <= 1 || mv->fourier_aligned) return ZR_OK;
  if (!mv->have_xcorr || !mv->S) return fail(ZR_ERR_STATE, "Fourier shifts need the spectra of zr_xcorr");
  // masked frames were centred by their means (the common mean for 'refine') before the transforms
  const double* mu = mv->unmasked ? nullptr : (mv->p.tri_mode == ZR_TRI_REFINE ? mv->stats_c : mv->stats);
  int rc = zr_stage_fourier_align(mv->st, mv->S, mv->Wk, mv->Hs, mv->unmasked ? -1 : 2 * mv->N, mu, mv->stream);
  if (rc) return fail(rc, "%s", zr_stage_error());
  mv->fourier_aligned = true;
  return ZR_OK;
}

extern "C" int zr_apply_shifts(zr_movie* mv, float* aligned_out) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  if (!mv->frames) return fail(ZR_ERR_STATE, "frames not set");
  ZR_HIP(hipSetDevice(mv->device));
  if (int rf = ensure_fourier(mv)) return rf;
  hipEvent_t t0 = mv->timer.begin(mv->stream);
  int rc = zr_stage_shifts(mv->st, mv->frames, aligned_out, mv->stream);
  mv->timer.end(ZR_STAGE_SHIFTS, t0, mv->stream);
  return rc ? fail(rc, "%s", zr_stage_error()) : ZR_OK;
}

extern "C" int zr_frc(zr_movie* mv) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  ZR_HIP(hipSetDevice(mv->device));
  hipEvent_t t0 = mv->timer.begin(mv->stream);
  int rc = zr_stage_frc(mv->st, mv->stream);
  mv->timer.end(ZR_STAGE_FRC, t0, mv->stream);
  return rc ? fail(rc, "%s", zr_stage_error()) : ZR_OK;
}

extern "C" int zr_set_frc_translations(zr_movie* mv, const double* trans, int n) {
  if (!mv || !trans || n < 1) return fail(ZR_ERR_ARG, "NULL argument");
  ZR_HIP(hipSetDevice(mv->device));
  int rc = zr_stage_set_frc_crop(mv->st, trans, n, mv->stream);
  return rc ? fail(rc, "%s", zr_stage_error()) : ZR_OK;
}

extern "C" int zr_filter(zr_movie* mv, int step) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  ZR_HIP(hipSetDevice(mv->device));
  int rc = ZR_OK;
  if (step == ZR_FILTER_DOSE || step == ZR_FILTER_DOSENORM) {
    if (mv->spectral && !mv->have_xcorr) return fail(ZR_ERR_STATE, "dose filter needs the spectra of zr_xcorr");
    if (int rf = ensure_fourier(mv)) return rf;
    hipEvent_t t0 = mv->timer.begin(mv->stream);
    rc = zr_stage_dose(mv->st, mv->frames, mv->S, mv->Wk, step == ZR_FILTER_DOSENORM ? 1 : 0, mv->stream);
    mv->timer.end(ZR_STAGE_DOSE, t0, mv->stream);
  } else if (step == ZR_FILTER_GAUSSLP) {
    rc = zr_stage_gausslp(mv->st, mv->stream);
  } else if (step == ZR_FILTER_BACKGROUND) {
    const int ny = (mv->H + 127) / 128, nx = (mv->W + 127) / 128;
    if (ny < 4 || nx < 4 || ny > 80 || nx > 80)
      return fail(ZR_ERR_UNSUPPORTED, "background: %dx%d frames give a %dx%d sample grid (built 4..80 per axis)",
                  mv->H, mv->W, ny, nx);
    rc = zr_stage_background(mv->st, mv->stream);
  } else {
    return fail(ZR_ERR_ARG, "filter step %d is not a ZR_FILTER_* value", step);
  }
  return rc ? fail(rc, "%s", zr_stage_error()) : ZR_OK;
}

extern "C" int zr_dose_filter(zr_movie* mv) { return zr_filter(mv, ZR_FILTER_DOSE); }

extern "C" int zr_get_background_fit(zr_movie* mv, double* p6, double* fval, int* status) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  ZR_HIP(hipSetDevice(mv->device));
  ZR_HIP(sync_movie(mv));
  const int rc = zr_stage_get_background(mv->st, p6, fval, status);
  return rc ? fail(rc, "%s", zr_stage_error()) : ZR_OK;
}

extern "C" int zr_post_filters(zr_movie* mv) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  // a dose step forms the aligned sums and half sums itself; unless the first
  // step is one, the frames are shifted and summed first (applyShifts,
  // zorro/zorro.py:1672), which a Gaussian step needs for its imageSum copy
  int rc;
  const bool dose_first = mv->p.n_filters > 0 &&
                          (mv->p.filters[0] == ZR_FILTER_DOSE || mv->p.filters[0] == ZR_FILTER_DOSENORM);
  if (!dose_first && (rc = zr_apply_shifts(mv, nullptr))) return rc;
  for (int i = 0; i < mv->p.n_filters; ++i)
    if ((rc = zr_filter(mv, mv->p.filters[i]))) return rc;
  return ZR_OK;
}

// the stages after the pair transform (peaks, solve, shifts/filters, FRC)
static int align_tail(zr_movie* mv) {
  int rc;
  if ((rc = xcorr_part(mv, 2))) return rc;
  if ((rc = zr_solve(mv))) return rc;
  if ((rc = zr_post_filters(mv))) return rc;
  if (mv->p.do_frc && (rc = zr_frc(mv))) return rc;
  return ZR_OK;
}

static int align_sequence(zr_movie* mv) {
  int rc;
  if ((rc = zr_xcorr(mv))) return rc;
  if ((rc = zr_solve(mv))) return rc;
  if ((rc = zr_post_filters(mv))) return rc;
  if (mv->p.do_frc && (rc = zr_frc(mv))) return rc;
  return ZR_OK;
}

extern "C" int zr_movie_use_graph(zr_movie* mv, int enable) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  if (mv->view) return fail(ZR_ERR_STATE, "a movie of a batch is correlated by zr_batch_align");
  if (enable && mv->p.tri_mode == ZR_TRI_AUTO)
    return fail(ZR_ERR_UNSUPPORTED, "triMode 'auto' reads the pair list back to the host between stages: no graph");
  if (!enable) drop_graph(mv);
  mv->use_graph = enable != 0;
  return ZR_OK;
}

extern "C" int zr_align(zr_movie* mv) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  if (mv->view) return fail(ZR_ERR_STATE, "a movie of a batch is correlated by zr_batch_align");
  if (!zr_stage_equations_ok(mv->st) && mv->p.tri_mode != ZR_TRI_FIRST && mv->p.tri_mode != ZR_TRI_REFINE)
    return fail(ZR_ERR_UNSUPPORTED, "pair pattern gives equations the solver cannot represent");
  if (!mv->use_graph) return align_sequence(mv);
  if (!mv->frames) return fail(ZR_ERR_STATE, "frames not set");
  ZR_HIP(hipSetDevice(mv->device));
  if (!mv->graph_exec[0]) {
    // capture the stage sequence on the movie's stream once, as three graphs
    // (spectra | pair transform | the rest); replays re-run them on whatever
    // the frames buffer holds.  Timing events stay outside the graphs.
    const bool timing = mv->timer.on;
    mv->timer.on = false;
    int rc = ZR_OK;
    for (int seg = 0; seg < 3 && rc == ZR_OK; ++seg) {
      hipError_t e = hipStreamBeginCapture(mv->stream, hipStreamCaptureModeRelaxed);
      if (e != hipSuccess) { rc = fail(ZR_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e)); break; }
      rc = seg < 2 ? xcorr_part(mv, seg) : align_tail(mv);
      hipGraph_t g = nullptr;
      e = hipStreamEndCapture(mv->stream, &g);
      if (rc == ZR_OK && e != hipSuccess) rc = fail(ZR_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
      if (rc != ZR_OK) { if (g) hipGraphDestroy(g); break; }
      mv->graph[seg] = g;
      e = hipGraphInstantiate(&mv->graph_exec[seg], g, nullptr, nullptr, 0);
      if (e != hipSuccess) rc = fail(ZR_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    mv->timer.on = timing;
    if (rc != ZR_OK) { drop_graph(mv); return rc; }
  }
  static const int seg_stage[3] = {ZR_STAGE_SPECTRA, ZR_STAGE_PAIRS, ZR_STAGE_TAIL};
  for (int seg = 0; seg < 3; ++seg) {
    hipEvent_t t = mv->timer.begin(mv->stream);
    ZR_HIP(hipGraphLaunch(mv->graph_exec[seg], mv->stream));
    mv->timer.end(seg_stage[seg], t, mv->stream);
  }
  return ZR_OK;
}

extern "C" int zr_get_hotpix_info(zr_movie* mv, double* info10) {
  if (!mv || !info10) return fail(ZR_ERR_ARG, "NULL argument");
  if (!mv->p.do_hot || !mv->hot_ran) return fail(ZR_ERR_STATE, "the outlier-pixel filter has not run");
  ZR_HIP(hipSetDevice(mv->device));
  HotState h{};
  ZR_HIP(hipMemcpyAsync(&h, mv->hb.state, sizeof(h), hipMemcpyDeviceToHost, mv->stream));
  ZR_HIP(sync_movie(mv));
  if (h.status != 0)
    return fail(ZR_ERR_UNSUPPORTED, "hotpixFilter: the histogram of the unaligned sum (mean %g) needs %g unit bins, "
                "outside 2..%d", h.sum_mean, 2.0 * mv->p.hot_max_sigma * std::sqrt(std::max(h.sum_mean, 0.0)) + 1.0,
                mv->ha.maxbins);
  const double n = (double)mv->N;
  const double v[10] = {h.mu, h.sigma, h.cut_lo, h.cut_hi, (double)h.n_dead, (double)h.n_hot, (double)h.n_nbr,
                        h.mu / n, h.sigma / std::sqrt(n), h.objective};
  std::memcpy(info10, v, sizeof(v));
  return ZR_OK;
}

extern "C" int zr_synchronize(zr_movie* mv) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  ZR_HIP(hipSetDevice(mv->device));
  ZR_HIP(sync_movie(mv));
  return ZR_OK;
}

extern "C" int zr_equation_count(zr_movie* mv, int* m) {
  if (!mv || !m) return fail(ZR_ERR_ARG, "NULL argument");
  *m = zr_stage_equation_count(mv->st);
  return ZR_OK;
}

extern "C" int zr_pair_count(zr_movie* mv, int* npairs) {
  if (!mv || !npairs) return fail(ZR_ERR_ARG, "NULL argument");
  *npairs = mv->P;
  return ZR_OK;
}

extern "C" int zr_get_trimats(zr_movie* mv, float* shifts, float* corr, float* peaksig, float* origin) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  if (!mv->have_xcorr) return fail(ZR_ERR_STATE, "no correlation results");
  ZR_HIP(hipSetDevice(mv->device));
  ZR_HIP(sync_movie(mv));
  if (mv->preshift) {
    int flag = 0;
    ZR_HIP(hipMemcpy(&flag, mv->pflag, sizeof(int), hipMemcpyDeviceToHost));
    if (flag)
      return fail(ZR_ERR_UNSUPPORTED, "preShift: a pair's disc holds no positive correlation; the reference's argmax "
                  "then takes the first zero of the whole map (zorro/zorro.py:1159-1161), which no window holds");
  }
  size_t nn = (size_t)mv->N * mv->N;
  if (shifts) ZR_HIP(hipMemcpy(shifts, mv->shifts, nn * 2 * sizeof(float), hipMemcpyDeviceToHost));
  if (corr) ZR_HIP(hipMemcpy(corr, mv->corr, nn * sizeof(float), hipMemcpyDeviceToHost));
  if (peaksig) ZR_HIP(hipMemcpy(peaksig, mv->peaksig, nn * sizeof(float), hipMemcpyDeviceToHost));
  if (origin) ZR_HIP(hipMemcpy(origin, mv->origin, nn * sizeof(float), hipMemcpyDeviceToHost));
  return ZR_OK;
}

extern "C" int zr_get_frames(zr_movie* mv, float* frames) {
  if (!mv || !frames) return fail(ZR_ERR_ARG, "NULL argument");
  if (!mv->frames) return fail(ZR_ERR_STATE, "frames not set");
  ZR_HIP(hipSetDevice(mv->device));
  ZR_HIP(sync_movie(mv));
  ZR_HIP(hipMemcpy(frames, mv->frames, (size_t)mv->N * mv->H * mv->W * sizeof(float), hipMemcpyDeviceToHost));
  return ZR_OK;
}

extern "C" int zr_get_peak_sigmas(zr_movie* mv, float* sigma) {
  if (!mv || !sigma) return fail(ZR_ERR_ARG, "NULL argument");
  if (!mv->have_xcorr) return fail(ZR_ERR_STATE, "no correlation results");
  if (mv->p.b_mode != ZR_B_OPTI) return fail(ZR_ERR_STATE, "per-pair sigmas exist only for Bmode='opti'");
  if (!mv->bsigma) return fail(ZR_ERR_UNSUPPORTED, "per-pair sigmas of a movie in a batch are not kept");
  ZR_HIP(hipSetDevice(mv->device));
  ZR_HIP(sync_movie(mv));
  ZR_HIP(hipMemcpy(sigma, mv->bsigma, (size_t)mv->N * mv->N * sizeof(float), hipMemcpyDeviceToHost));
  return ZR_OK;
}

extern "C" int zr_get_windows(zr_movie* mv, float* windows) {
  if (!mv || !windows) return fail(ZR_ERR_ARG, "NULL argument");
  if (mv->view) return fail(ZR_ERR_UNSUPPORTED, "the windows of a batch are not kept per movie");
  if (!mv->have_xcorr) return fail(ZR_ERR_STATE, "no correlation results");
  ZR_HIP(hipSetDevice(mv->device));
  ZR_HIP(sync_movie(mv));
  const size_t nn = 4 * (size_t)mv->m * mv->m;
  ZR_HIP(hipMemcpy(windows, mv->windows + mv->win_out0 * nn, (size_t)mv->P * nn * sizeof(float), hipMemcpyDeviceToHost));
  return ZR_OK;
}

extern "C" int zr_get_solution(zr_movie* mv, double* translations, double* relative_est, double* weights,
                               double* logistic, double* error, double* error_stats, double* error_xy_frames) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  ZR_HIP(hipSetDevice(mv->device));
  ZR_HIP(sync_movie(mv));
  int rc = zr_stage_get_solution(mv->st, translations, relative_est, weights, logistic, error, error_stats,
                                 error_xy_frames);
  return rc ? fail(rc, "%s", zr_stage_error()) : ZR_OK;
}

extern "C" int zr_get_peaksig_hist(zr_movie* mv, float* pdf, float* cdf, float* hsigma, int cap, int* nbins) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  ZR_HIP(hipSetDevice(mv->device));
  ZR_HIP(sync_movie(mv));
  int rc = zr_stage_get_hist(mv->st, pdf, cdf, hsigma, cap, nbins);
  return rc ? fail(rc, "%s", zr_stage_error()) : ZR_OK;
}

extern "C" int zr_get_solver_info(zr_movie* mv, int* info6) {
  if (!mv || !info6) return fail(ZR_ERR_ARG, "NULL argument");
  ZR_HIP(hipSetDevice(mv->device));
  ZR_HIP(sync_movie(mv));
  int rc = zr_stage_solver_info(mv->st, info6);
  return rc ? fail(rc, "%s", zr_stage_error()) : ZR_OK;
}

extern "C" int zr_get_sums(zr_movie* mv, float* image_sum, float* filt_sum, float* even_sum, float* odd_sum) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  ZR_HIP(hipSetDevice(mv->device));
  ZR_HIP(sync_movie(mv));
  int rc = zr_stage_get_sums(mv->st, image_sum, filt_sum, even_sum, odd_sum);
  return rc ? fail(rc, "%s", zr_stage_error()) : ZR_OK;
}

extern "C" int zr_get_frc(zr_movie* mv, double* frc, int* n) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  ZR_HIP(hipSetDevice(mv->device));
  ZR_HIP(sync_movie(mv));
  int rc = zr_stage_get_frc(mv->st, frc, n);
  return rc ? fail(rc, "%s", zr_stage_error()) : ZR_OK;
}

extern "C" int zr_device_results(zr_movie* mv, const float** image_sum, const float** filt_sum,
                                 const double** translations) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  int rc = zr_stage_device_results(mv->st, image_sum, filt_sum, translations);
  return rc ? fail(rc, "%s", zr_stage_error()) : ZR_OK;
}

extern "C" int zr_set_timing(zr_movie* mv, int enable) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  mv->timer.on = enable != 0;
  return ZR_OK;
}

extern "C" int zr_get_stage_times(zr_movie* mv, double* ms, long* counts, int reset) {
  if (!mv) return fail(ZR_ERR_ARG, "NULL movie");
  ZR_HIP(hipSetDevice(mv->device));
  mv->timer.harvest();
  for (int i = 0; i < ZR_NSTAGES; ++i) {
    if (ms) ms[i] = mv->timer.ms[i];
    if (counts) counts[i] = mv->timer.count[i];
    if (reset) { mv->timer.ms[i] = 0.0; mv->timer.count[i] = 0; }
  }
  return ZR_OK;
}

extern "C" int zr_movie_info(zr_movie* mv, int* dims) {
  if (!mv || !dims) return fail(ZR_ERR_ARG, "NULL argument");
  const int v[12] = {mv->N, mv->H, mv->W, mv->Hp, mv->Wp, mv->Hc, mv->Wc, mv->Wh, mv->m, mv->P, mv->unmasked ? 1 : 0, mv->batch};
  for (int i = 0; i < 12; ++i) dims[i] = v[i];
  return ZR_OK;
}

// ==================================================================== batches
// Several movies of one detector shape and one set of settings registered
// together (the reference registers one movie per process and scales by
// running processes side by side, automator/zorroSkulkManager.py:1089-1113):
// the forward spectra of every frame of every movie, the pair transforms of
// all their pairs (one ragged pair list) and the peak searches run as single
// launches over the batch; each movie's solver, shift/filter and FRC stages
// then run on a stream of their own, so that the one-workgroup solvers of the
// movies overlap.  With graphs on, the whole sequence (forked over the movie
// streams and joined back) is one captured graph.
struct zr_batch {
  int device = 0;
  zr_movie* big = nullptr;               // correlation plan over all frames (block-diagonal pairs)
  std::vector<zr_movie*> views;          // one per movie
  std::vector<int> off, nfr;             // first frame and frame count per movie
  std::vector<hipEvent_t> done;          // per movie: its stages finished
  hipEvent_t xc = nullptr;               // the batch's correlation results are ready
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
};

static void free_batch(zr_batch* b) {
  if (!b) return;
  hipSetDevice(b->device);
  if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
  if (b->graph) hipGraphDestroy(b->graph);
  for (zr_movie* v : b->views) free_movie(v);
  for (hipEvent_t e : b->done)
    if (e) hipEventDestroy(e);
  if (b->xc) hipEventDestroy(b->xc);
  free_movie(b->big);
  delete b;
}

extern "C" void zr_batch_destroy(zr_batch* b) { free_batch(b); }

// a movie of the batch: frames and spectra are slices of the batch's, the
// trimats, solver/shift/filter/FRC state and the stream are its own
static int create_view(zr_batch* b, int i, const zr_params& p, zr_movie** out) {
  const zr_movie* big = b->big;
  const int n = b->nfr[i], off = b->off[i];
  zr_movie* v = new zr_movie();
  v->view = true;
  v->device = big->device;
  v->p = big->p;
  v->p.tri_mode = p.tri_mode;
  v->N = n;
  v->Hr = big->Hr; v->Wr = big->Wr; v->H = big->H; v->W = big->W; v->Hp = big->Hp; v->Wp = big->Wp;
  v->Hc = big->Hc; v->Wc = big->Wc; v->Wh = big->Wh; v->m = big->m; v->Lper = big->Lper; v->disc = big->disc;
  v->crop = big->crop; v->unmasked = big->unmasked; v->spectral = big->spectral; v->full_spec = big->full_spec;
  v->Wk = big->Wk; v->Hs = big->Hs; v->hoff = big->hoff; v->side_in_col = big->side_in_col;
  v->batch = big->batch; v->rb = big->rb; v->k2_cpw = big->k2_cpw; v->k2_ppw = big->k2_ppw;
  v->frames_owned = big->frames_owned + (size_t)off * big->H * big->W;
  v->frames = v->frames_owned;
  v->S = big->S + (size_t)off * big->Wk * big->Hs;     // frame f's spectrum is slot f (unmasked and masked)
  for (const int4& q : big->pairs_host)
    if (q.x >= off && q.y < off + n) v->pairs_host.push_back(make_int4(q.x - off, q.y - off, (int)v->pairs_host.size(), 0));
  v->P = v->Ptot = (int)v->pairs_host.size();
  v->nspec = big->unmasked ? n : 2 * n + 1;
  g_alloc_bytes = 0;
  const size_t nn = (size_t)n * n;
  int rc;
  hipError_t e = hipStreamCreateWithFlags(&v->own_stream, hipStreamNonBlocking);
  if (e != hipSuccess) { free_movie(v); return fail(ZR_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
  v->stream = v->own_stream;
  v->batch_stream = big->stream;
  if ((rc = dev_alloc(&v->shifts, nn * 2)) || (rc = dev_alloc(&v->corr, nn)) || (rc = dev_alloc(&v->peaksig, nn)) ||
      (rc = dev_alloc(&v->origin, nn))) {
    free_movie(v);
    return rc;
  }
  if ((rc = zr_stage_create(&v->st, v->p, n, v->H, v->W, v->Hp, v->Wp, v->pairs_host, v->spectral ? 1 : 0))) {
    free_movie(v);
    return fail(rc, "%s", zr_stage_error());
  }
  v->bytes = g_alloc_bytes + zr_stage_device_bytes(v->st);
  *out = v;
  return ZR_OK;
}

extern "C" int zr_batch_create(int device, const zr_params* p, int n_movies, const int* n_frames, int h, int w,
                               zr_batch** out) {
  if (!p || !n_frames || !out || n_movies < 1) return fail(ZR_ERR_ARG, "NULL argument or no movies");
  *out = nullptr;
  if (p->tri_mode != ZR_TRI_TRI && p->tri_mode != ZR_TRI_DIAG)
    return fail(ZR_ERR_UNSUPPORTED, "batches run triMode 'tri' or 'diag' (got %d)", p->tri_mode);
  if (p->pre_shift || p->do_hot || p->shape_binned[0] > 0 || p->shape_binned[1] > 0)
    return fail(ZR_ERR_UNSUPPORTED, "batches take float32 frames without preShift, 'hot' or binning");
  if (p->shift_method == ZR_SHIFT_FOURIER)
    return fail(ZR_ERR_UNSUPPORTED, "batches shift by Lanczos kernels or integers ('fourier' runs per movie)");
  if (p->user_mask) return fail(ZR_ERR_UNSUPPORTED, "batches correlate under the frame region (user masks run per movie)");
  zr_batch* b = new zr_batch();
  b->device = device;
  int total = 0;
  for (int i = 0; i < n_movies; ++i) {
    if (n_frames[i] < 2 || n_frames[i] > 64) { free_batch(b); return fail(ZR_ERR_ARG, "movie %d: %d frames outside 2..64", i, n_frames[i]); }
    b->off.push_back(total);
    b->nfr.push_back(n_frames[i]);
    total += n_frames[i];
  }
  // the pairs of every movie, in frame numbers of the whole batch
  std::vector<uint8_t> mask((size_t)total * total, 0);
  for (int i = 0; i < n_movies; ++i)
    for (int I = 0; I < b->nfr[i]; ++I)
      for (int J = I + 1; J < b->nfr[i]; ++J) {
        const int d = J - I;
        const int wdt = p->diag_width < 0 ? b->nfr[i] : p->diag_width;
        const bool take = p->tri_mode == ZR_TRI_TRI || (d <= wdt + p->diag_start && d > p->diag_start);
        if (take) mask[(size_t)(b->off[i] + I) * total + b->off[i] + J] = 1;
      }
  zr_params pb = *p;
  pb.tri_mode = ZR_TRI_TRI;
  int rc = create_movie(device, &pb, total, h, w, mask.data(), &b->big, true);
  if (rc) { delete b; return rc; }
  zr_movie* big = b->big;
  g_alloc_bytes = 0;
  if ((rc = dev_alloc(&big->frames_owned, (size_t)total * big->H * big->W))) { free_batch(b); return rc; }
  big->frames = big->frames_owned;
  big->bytes += g_alloc_bytes;
  for (int i = 0; i < n_movies; ++i) {
    zr_movie* v = nullptr;
    if ((rc = create_view(b, i, *p, &v))) { free_batch(b); return rc; }
    b->views.push_back(v);
    hipEvent_t e = nullptr;
    hipError_t he = hipEventCreateWithFlags(&e, hipEventDisableTiming);
    if (he != hipSuccess) { free_batch(b); return fail(ZR_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(he)); }
    b->done.push_back(e);
  }
  hipError_t he = hipEventCreateWithFlags(&b->xc, hipEventDisableTiming);
  if (he != hipSuccess) { free_batch(b); return fail(ZR_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(he)); }
  *out = b;
  return ZR_OK;
}

extern "C" int zr_batch_movie(zr_batch* b, int i, zr_movie** view) {
  if (!b || !view || i < 0 || i >= (int)b->views.size()) return fail(ZR_ERR_ARG, "batch movie index");
  *view = b->views[i];
  return ZR_OK;
}

extern "C" int zr_batch_frames(zr_batch* b, int i, float** frames) {
  if (!b || !frames || i < 0 || i >= (int)b->views.size()) return fail(ZR_ERR_ARG, "batch movie index");
  *frames = b->views[i]->frames_owned;
  return ZR_OK;
}

extern "C" int zr_batch_set_frames(zr_batch* b, int i, const float* frames, int on_device) {
  if (!b || !frames || i < 0 || i >= (int)b->views.size()) return fail(ZR_ERR_ARG, "batch movie index");
  zr_movie* v = b->views[i];
  ZR_HIP(hipSetDevice(b->device));
  const size_t bytes = (size_t)v->N * v->H * v->W * sizeof(float);
  ZR_HIP(hipMemcpyAsync(v->frames_owned, frames, bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                        b->big->stream));
  ZR_HIP(hipStreamSynchronize(b->big->stream));   // the source may be reused or freed on return
  return ZR_OK;
}

extern "C" int zr_batch_device_bytes(zr_batch* b, size_t* bytes) {
  if (!b || !bytes) return fail(ZR_ERR_ARG, "NULL argument");
  size_t t = b->big->bytes;
  for (zr_movie* v : b->views) t += v->bytes;
  *bytes = t;
  return ZR_OK;
}

static int batch_sequence(zr_batch* b) {
  zr_movie* big = b->big;
  int rc;
  for (int part = 0; part < 3; ++part)
    if ((rc = xcorr_part(big, part))) return rc;
  ZR_HIP(hipEventRecord(b->xc, big->stream));
  for (size_t i = 0; i < b->views.size(); ++i) {
    zr_movie* v = b->views[i];
    ZR_HIP(hipStreamWaitEvent(v->stream, b->xc, 0));
    ZR_HIP(zr_launch_gather_trimats(big->shifts, big->corr, big->peaksig, big->origin, big->N, b->off[i], v->N,
                                    v->shifts, v->corr, v->peaksig, v->origin, v->stream));
    v->have_xcorr = true;
    if ((rc = zr_solve(v))) return rc;
    if ((rc = zr_post_filters(v))) return rc;
    if (v->p.do_frc && (rc = zr_frc(v))) return rc;
    ZR_HIP(hipEventRecord(b->done[i], v->stream));
  }
  for (hipEvent_t e : b->done) ZR_HIP(hipStreamWaitEvent(big->stream, e, 0));
  return ZR_OK;
}

extern "C" int zr_batch_use_graph(zr_batch* b, int enable) {
  if (!b) return fail(ZR_ERR_ARG, "NULL batch");
  if (!enable && b->graph_exec) {
    hipGraphExecDestroy(b->graph_exec);
    hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
  }
  b->use_graph = enable != 0;
  return ZR_OK;
}

extern "C" int zr_batch_align(zr_batch* b) {
  if (!b) return fail(ZR_ERR_ARG, "NULL batch");
  ZR_HIP(hipSetDevice(b->device));
  if (!b->use_graph) return batch_sequence(b);
  hipStream_t st = b->big->stream;
  if (!b->graph_exec) {
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) return fail(ZR_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    int rc = batch_sequence(b);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(st, &g);
    if (rc == ZR_OK && e != hipSuccess) rc = fail(ZR_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    if (rc != ZR_OK) { if (g) hipGraphDestroy(g); return rc; }
    b->graph = g;
    e = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      hipGraphDestroy(g);
      b->graph = nullptr;
      return fail(ZR_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
  }
  ZR_HIP(hipGraphLaunch(b->graph_exec, st));
  return ZR_OK;
}

extern "C" int zr_batch_synchronize(zr_batch* b) {
  if (!b) return fail(ZR_ERR_ARG, "NULL batch");
  ZR_HIP(hipSetDevice(b->device));
  ZR_HIP(hipStreamSynchronize(b->big->stream));
  for (zr_movie* v : b->views) ZR_HIP(hipStreamSynchronize(v->stream));
  return ZR_OK;
}
