"""Thin object layer over the C ABI: one ``MoviePlan`` per movie shape.

``MoviePlan`` owns a ``zr_movie`` (device workspaces for one frame count and
shape) and exposes the stages of ``ImageRegistrator.alignImageStack``
(``zorro/zorro.py:1569-1716``) one by one, plus result read-back.  Frames may be
a NumPy array (copied to the device) or a torch CUDA tensor (borrowed, no
copy).  Nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

TRI = {"tri": L.ZR_TRI_TRI, "diag": L.ZR_TRI_DIAG, "auto": L.ZR_TRI_AUTO, "first": L.ZR_TRI_FIRST,
       "refine": L.ZR_TRI_REFINE}
BMODE = {None: L.ZR_B_NONE, "none": L.ZR_B_NONE, "conv": L.ZR_B_CONV, "convolution": L.ZR_B_CONV,
         "opti": L.ZR_B_OPTI}
WMODE = {"autologistic": L.ZR_W_AUTOLOGISTIC, "logistic": L.ZR_W_LOGISTIC, "corr": L.ZR_W_CORR, None: L.ZR_W_CORR,
         "unweighted": L.ZR_W_UNWEIGHTED, "norm": L.ZR_W_NORM, "normalized": L.ZR_W_NORM}


# filterMode steps after the registration (zorro/zorro.py:1692-1711)
FILTER_STEPS = {"dose": L.ZR_FILTER_DOSE, "dosenorm": L.ZR_FILTER_DOSENORM, "gausslp": L.ZR_FILTER_GAUSSLP,
                "background": L.ZR_FILTER_BACKGROUND}


def parse_filter_mode(filterMode):
    """(steps, hot) of a filterMode string as alignImageStack reads it: 'hot'
    anywhere in the string runs hotpixFilter before the registration
    (zorro/zorro.py:1595-1596); the comma-separated tokens, lower-cased with
    spaces removed, are then applied in order to filtSum (:1692-1711).  A
    token that is none of dose, dosenorm, background, gausslp or a 'hot'
    token raises ValueError (the reference skips it silently, which would hide
    a misspelt filter)."""
    if not filterMode:
        return [], False
    text = str(filterMode).lower()
    steps = []
    for tok in text.replace(" ", "").split(","):
        if tok in FILTER_STEPS:
            steps.append(tok)
        elif tok == "" or "hot" in tok:
            continue
        else:
            raise ValueError(f"filterMode token {tok!r} is not one of 'dose', 'dosenorm', 'background', "
                             "'gausslp', 'hot'")
    if len(steps) > 4:
        raise NotImplementedError(f"at most 4 filterMode steps are built (got {len(steps)})")
    return steps, "hot" in text


def make_params(triMode="diag", diagWidth=5, diagStart=0, maxShift=100, Bmode="opti", Brad=512.0, subPixReg=16,
                shiftMethod="lanczos", weightMode="autologistic", peaksigThres=6.0, logisticK=5.0,
                logisticNu=0.15, originMode="centroid", suppressOrigin=True, fouCrop=(3072, 3072),
                shapePadded=(4096, 4096), filterMode=None, doseFiltParam=(None, 0.24499, -1.6649, 2.8141, 32, 0),
                pixelsize=None, voltage=300.0, doLazyFRC=True, frc_box=512, frc_overlap=0.5,
                autoMax=10, shapeBinned=None, hotpixInfo=None, neighbourThres=0.01,
                hotMuSigma=None, hotCutoffs=None, preShift=False, k1_batch=0, k2_cols_per_wave=0, k2_pairs_per_wave=0,
                user_mask=False) -> L.ZrParams:
    """ZrParams from ImageRegistrator attribute values (defaults of zorro/zorro.py:82-243).
    ``hotpixInfo`` (defaults of zorro/zorro.py:139-141) and ``neighbourThres``
    configure the 'hot' filter mode; ``hotMuSigma`` fixes the normal-CDF
    parameters that the filter otherwise fits; ``hotCutoffs`` = (lower, upper)
    skips the cut-off fits (``hotpixFilter(cutoffLower, cutoffUpper)``).
    ``user_mask``: the plan correlates under a binary mask given by
    ``MoviePlan.set_mask`` (the reference's ``masks`` [1,H,W]).
    ``k1_batch`` / ``k2_cols_per_wave`` / ``k2_pairs_per_wave``: launch tuning (0: built-in default;
    ``zorro_amd.autotune`` picks them per shape); results do not depend on them."""
    p = L.ZrParams()
    if triMode not in TRI:
        raise NotImplementedError(f"triMode {triMode!r} is not built")
    p.tri_mode = TRI[triMode]
    p.diag_width = -1 if diagWidth is None else int(diagWidth)
    p.diag_start = int(diagStart)
    p.auto_max = int(autoMax)
    if not maxShift:
        raise NotImplementedError("maxShift=None is not built")
    p.max_shift = int(maxShift)
    if Bmode == "fourier":
        raise NotImplementedError("Bmode='fourier' is broken in the reference (zorro/zorro.py:1125-1133)")
    if not Brad and Bmode in ("conv",):
        Bmode = None
    p.b_mode = BMODE.get(Bmode, L.ZR_B_NONE)
    p.b_rad = float(Brad or 0.0)
    p.sub_pix_reg = int(subPixReg or 1)
    # 'fourier' is broken in the reference (zorro/zorro.py:891-898, :1762-1763); here it is
    # the shift theorem on the zero-padded frames' spectra
    p.shift_method = {"lanczos": L.ZR_SHIFT_LANCZOS, "fourier": L.ZR_SHIFT_FOURIER}.get(shiftMethod,
                                                                                       L.ZR_SHIFT_INTEGER)
    p.weight_mode = WMODE.get(weightMode, L.ZR_W_CORR)
    p.peaksig_thres, p.logistic_k, p.logistic_nu = float(peaksigThres), float(logisticK), float(logisticNu)
    p.origin_centroid = 1 if originMode == "centroid" else 0
    p.suppress_origin = 1 if suppressOrigin else 0
    fc = (0, 0) if fouCrop is None or not np.any(fouCrop) else tuple(int(v) for v in fouCrop)
    p.fou_crop[0], p.fou_crop[1] = fc
    p.shape_padded[0], p.shape_padded[1] = (int(shapePadded[0]), int(shapePadded[1]))
    steps, hot = parse_filter_mode(filterMode)
    p.n_filters = len(steps)
    for i, st in enumerate(steps):
        p.filters[i] = FILTER_STEPS[st]
    p.do_dose = 1 if any(st in ("dose", "dosenorm") for st in steps) else 0
    if p.do_dose and doseFiltParam[0] is not None:
        raise NotImplementedError("doseFiltParam[0] must be None (the reference raises NameError otherwise, Q11)")
    p.dose_a, p.dose_b, p.dose_c, p.dose_order, p.dose_missing = (float(v) for v in doseFiltParam[1:6])
    p.pixelsize_nm = float(pixelsize) if pixelsize else 0.0
    p.voltage_kv = float(voltage or 0.0)
    p.do_frc = 1 if doLazyFRC else 0
    p.frc_box = int(frc_box)
    p.frc_overlap = float(frc_overlap)
    if shapeBinned is not None and np.any(shapeBinned):
        p.shape_binned[0], p.shape_binned[1] = (int(shapeBinned[0]), int(shapeBinned[1]))
    hp = {**HOTPIX_DEFAULTS, **(hotpixInfo or {})}
    p.do_hot = 1 if hot else 0
    if p.do_hot and hp.get("decorrOutliers"):
        raise NotImplementedError("hotpixInfo['decorrOutliers'] adds random noise (zorro/zorro.py:2651-2666): not built")
    p.hot_logistic_k, p.hot_relax, p.hot_max_sigma = (float(hp[k]) for k in ("logisticK", "relax", "maxSigma"))
    p.hot_psf_k2 = 1 if hp["psf"] == "K2" else 0
    p.hot_cutoff_lower, p.hot_cutoff_upper = float(hp["cutoffLower"]), float(hp["cutoffUpper"])
    p.hot_neighbour_thres = float(neighbourThres)
    p.pre_shift = 1 if preShift else 0
    p.k1_batch = int(k1_batch or 0)
    p.k2_cols_per_wave = int(k2_cols_per_wave or 0)
    p.k2_pairs_per_wave = int(k2_pairs_per_wave or 0)
    p.user_mask = 1 if user_mask else 0
    if hotCutoffs is not None:
        p.hot_cutoff_lower, p.hot_cutoff_upper = (float(v) for v in hotCutoffs)
        p.hot_fixed_cutoffs = 1
    if hotMuSigma is not None:
        p.hot_mu_sigma[0], p.hot_mu_sigma[1] = (float(v) for v in hotMuSigma)
    return p


# hotpixInfo defaults (zorro/zorro.py:139-141)
HOTPIX_DEFAULTS = {"logisticK": 6.0, "relax": 0.925, "maxSigma": 8.0, "psf": "K2", "cutoffLower": -4.0,
                   "cutoffUpper": 3.25, "decorrOutliers": False}


class MoviePlan:
    """Device plan for one movie of ``n`` frames of ``h`` x ``w`` (the raw size:
    with ``shapeBinned`` the frames are binned on input and every result has the
    binned size ``hb`` x ``wb``)."""

    def __init__(self, params: L.ZrParams, n: int, h: int, w: int, device: int = 0, triIndices=None, _handle=None):
        """``triIndices``: optional boolean [n, n] matrix of the pairs (I < J) to
        correlate, as ``xcorrnm2_tri(triIndices=...)`` takes (triMode 'tri').
        ``_handle``: a movie of a batch (owned by the batch, see batch.py)."""
        self.lib = L.load()
        self.params = params
        self.n, self.h, self.w = int(n), int(h), int(w)
        self._mv = C.c_void_p()
        self._owned = _handle is None
        if _handle is not None:
            self._mv = _handle
        elif triIndices is None:
            L.check(self.lib.zr_movie_create(int(device), C.byref(params), self.n, self.h, self.w,
                                             C.byref(self._mv)))
        else:
            t = np.asarray(triIndices)
            if t.shape != (self.n, self.n):
                raise IndexError(f"triIndices is wrong size, should be of length: {self.n}")
            t = np.ascontiguousarray(t != 0, dtype=np.uint8)
            L.check(self.lib.zr_movie_create_pairs(int(device), C.byref(params), self.n, self.h, self.w,
                                                   C.c_void_p(t.ctypes.data), C.byref(self._mv)))
        npairs = C.c_int()
        L.check(self.lib.zr_pair_count(self._mv, C.byref(npairs)))
        self.npairs = npairs.value
        self.m = params.max_shift
        self._frames_ref = None
        inf = self.info()
        self.hb, self.wb = inf["H"], inf["W"]

    def close(self):
        if self._mv:
            if self._owned:
                self.lib.zr_movie_destroy(self._mv)
            self._mv = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_handle: int | None):
        L.check(self.lib.zr_movie_set_stream(self._mv, C.c_void_p(stream_handle or 0)))

    def set_frames(self, frames):
        """NumPy float32 [n,h,w] (copied) or a contiguous torch float32 CUDA tensor (borrowed)."""
        if hasattr(frames, "data_ptr") and getattr(frames, "is_cuda", False):
            if frames.dtype.__str__() != "torch.float32" or not frames.is_contiguous():
                raise ValueError("frames must be a contiguous float32 CUDA tensor")
            if tuple(frames.shape) != (self.n, self.h, self.w):
                raise ValueError(f"frames shape {tuple(frames.shape)} != plan {(self.n, self.h, self.w)}")
            self._frames_ref = frames
            L.check(self.lib.zr_movie_set_frames(self._mv, C.c_void_p(frames.data_ptr()), 1))
        else:
            a = np.ascontiguousarray(frames, dtype=np.float32)
            if a.shape != (self.n, self.h, self.w):
                raise ValueError(f"frames shape {a.shape} != plan {(self.n, self.h, self.w)}")
            self._frames_ref = a
            L.check(self.lib.zr_movie_set_frames(self._mv, C.c_void_p(a.ctypes.data), 0))

    def set_mask(self, mask):
        """The user mask of a plan made with ``make_params(user_mask=True)``:
        [h, w] (or [1, h, w]) of 0 and 1, numpy (copied) or a float32 CUDA tensor."""
        if hasattr(mask, "data_ptr") and getattr(mask, "is_cuda", False):
            m = mask.reshape(self.h, self.w)
            if str(m.dtype) != "torch.float32" or not m.is_contiguous():
                raise ValueError("mask must be a contiguous float32 CUDA tensor")
            L.check(self.lib.zr_movie_set_mask(self._mv, C.c_void_p(m.data_ptr()), 1))
            return
        a = np.asarray(mask)
        if a.size != self.h * self.w:
            raise ValueError(f"mask shape {a.shape} != frame {(self.h, self.w)}")
        a = np.ascontiguousarray(a.reshape(self.h, self.w), dtype=np.float32)
        L.check(self.lib.zr_movie_set_mask(self._mv, C.c_void_p(a.ctypes.data), 0))

    def set_frames_raw(self, frames, gain=None, gain_scalar=None, packed4=False):
        """Raw frames (uint8, uint16, int16 or float32; numpy or a CUDA tensor) and
        an optional gain reference f32 [h, w] already oriented as loadData does
        (``zorro/zorro.py:3486-3503``); gain product and binning run on the GPU.
        ``packed4``: frames are uint8 [n, h, w/2] holding two 4-bit counts per
        byte, the first pixel in the low nibble (MRC mode 101)."""
        dt = {"uint8": L.ZR_DTYPE_U8, "uint16": L.ZR_DTYPE_U16, "float32": L.ZR_DTYPE_F32,
              "int16": L.ZR_DTYPE_I16}
        def ptr(a, kind):
            if a is None:
                return C.c_void_p(0), 0, None
            if hasattr(a, "data_ptr") and getattr(a, "is_cuda", False):
                if not a.is_contiguous():
                    raise ValueError(f"{kind} must be contiguous")
                return C.c_void_p(a.data_ptr()), 1, a
            a = np.ascontiguousarray(a)
            return C.c_void_p(a.ctypes.data), 0, a
        name = str(frames.dtype).replace("torch.", "")
        if name not in dt:
            raise ValueError(f"raw frames must be uint8, uint16, int16 or float32, not {frames.dtype}")
        want = (self.n, self.h, self.w)
        code = dt[name]
        if packed4:
            if name != "uint8" or self.w % 2:
                raise ValueError("4-bit packed frames are uint8 [n, h, w/2] with an even w")
            want, code = (self.n, self.h, self.w // 2), L.ZR_DTYPE_U4
        if tuple(frames.shape) != want:
            raise ValueError(f"frames shape {tuple(frames.shape)} != {want}")
        if gain is not None:
            if not hasattr(gain, "data_ptr"):
                gain = np.asarray(gain, np.float32)
            if tuple(gain.shape) != (self.h, self.w):
                raise ValueError(f"gain shape {tuple(gain.shape)} != frame shape {(self.h, self.w)}")
        fp, fdev, fref = ptr(frames, "frames")
        gp, gdev, gref = ptr(gain, "gain")
        self._frames_ref = (fref, gref)
        L.check(self.lib.zr_movie_set_frames_raw(self._mv, fp, code, fdev, gp, gdev,
                                                 float(gain_scalar) if gain_scalar else 0.0))

    # ---- stages
    def xcorr(self):
        L.check(self.lib.zr_xcorr(self._mv))

    def solve(self):
        L.check(self.lib.zr_solve(self._mv))

    def set_trimats(self, shiftsTriMat, corrTriMat, peaksigTriMat):
        """Use caller-supplied triangular matrices as the correlation results
        (the inputs of the reference's shiftsSolver, ``zorro/zorro.py:1260``)."""
        n = self.n
        sh = np.ascontiguousarray(shiftsTriMat, np.float32).reshape(n, n, 2)
        co = np.ascontiguousarray(corrTriMat, np.float32).reshape(n, n)
        ps = np.ascontiguousarray(peaksigTriMat, np.float32).reshape(n, n)
        L.check(self.lib.zr_set_trimats(self._mv, L.fptr(sh), L.fptr(co), L.fptr(ps)))

    def apply_shifts(self, aligned_out=None):
        ptr = C.c_void_p(aligned_out.data_ptr()) if aligned_out is not None else C.c_void_p(0)
        L.check(self.lib.zr_apply_shifts(self._mv, ptr))

    def frc(self):
        L.check(self.lib.zr_frc(self._mv))

    def set_frc_translations(self, trans):
        """Crop limits of the next ``frc()`` from ``trans`` [n, 2] (tiledFRC's
        ``trans`` argument, zorro/zorro.py:2246-2274)."""
        t = np.ascontiguousarray(np.asarray(trans, np.float64)[:, :2])
        L.check(self.lib.zr_set_frc_translations(self._mv, t.ctypes.data_as(C.POINTER(C.c_double)), t.shape[0]))

    def dose_filter(self):
        L.check(self.lib.zr_dose_filter(self._mv))

    def filter(self, step: str):
        """One filterMode step ('dose', 'dosenorm', 'gausslp', 'background') on filtSum."""
        L.check(self.lib.zr_filter(self._mv, FILTER_STEPS[step]))

    def post_filters(self):
        """applyShifts (unless a dose step forms the sums) and the parameters'
        filterMode steps in order (zorro/zorro.py:1672-1711)."""
        L.check(self.lib.zr_post_filters(self._mv))

    def align(self):
        L.check(self.lib.zr_align(self._mv))

    def use_graph(self, on: bool = True):
        """Replay zr_align as one captured HIP graph (see zr_movie_use_graph)."""
        L.check(self.lib.zr_movie_use_graph(self._mv, 1 if on else 0))

    def synchronize(self):
        L.check(self.lib.zr_synchronize(self._mv))

    # ---- timing / info
    def set_timing(self, on: bool = True):
        L.check(self.lib.zr_set_timing(self._mv, 1 if on else 0))

    def stage_times(self, reset: bool = False):
        ms = np.zeros(len(L.STAGES))
        cnt = (C.c_long * len(L.STAGES))()
        L.check(self.lib.zr_get_stage_times(self._mv, L.dptr(ms), cnt, 1 if reset else 0))
        return {k: (ms[i], cnt[i]) for i, k in enumerate(L.STAGES)}

    def info(self):
        d = (C.c_int * 12)()
        L.check(self.lib.zr_movie_info(self._mv, d))
        keys = ["N", "H", "W", "Hp", "Wp", "Hc", "Wc", "Wh", "maxShift", "pairs", "unmasked", "batch"]
        return dict(zip(keys, list(d)))

    # ---- results
    def trimats(self):
        n = self.n
        sh = np.zeros((n, n, 2), np.float32)
        co = np.zeros((n, n), np.float32)
        ps = np.zeros((n, n), np.float32)
        og = np.zeros((n, n), np.float32)
        L.check(self.lib.zr_get_trimats(self._mv, L.fptr(sh), L.fptr(co), L.fptr(ps), L.fptr(og)))
        return dict(shiftsTriMat=sh, corrTriMat=co, peaksigTriMat=ps, originTriMat=og)

    def frames(self):
        """The working frames [n, hb, wb] (after gain and binning)."""
        f = np.zeros((self.n, self.hb, self.wb), np.float32)
        L.check(self.lib.zr_get_frames(self._mv, L.fptr(f)))
        return f

    def hotpix_info(self):
        """Results of the 'hot' filter: the hotpixInfo keys the reference fills
        (zorro/zorro.py:2578-2602, :2637) plus the fitted (mu, sigma) and the
        fit objective."""
        v = np.zeros(10, np.float64)
        L.check(self.lib.zr_get_hotpix_info(self._mv, v.ctypes.data_as(C.POINTER(C.c_double))))
        return dict(mu=float(v[0]), sigma=float(v[1]), cutoffLower=float(v[2]), cutoffUpper=float(v[3]),
                    guessDeadpix=int(v[4]), guessHotpix=int(v[5]), neighborPix=int(v[6]), frameMean=float(v[7]),
                    frameStd=float(v[8]), objective=float(v[9]))

    def peak_sigmas(self):
        """Bmode='opti': Gaussian width chosen per pair, [N, N]."""
        sg = np.zeros((self.n, self.n), np.float32)
        L.check(self.lib.zr_get_peak_sigmas(self._mv, L.fptr(sg)))
        return sg

    def windows(self):
        w = np.zeros((self.npairs, 2 * self.m, 2 * self.m), np.float32)
        L.check(self.lib.zr_get_windows(self._mv, L.fptr(w)))
        return w

    def equation_count(self):
        m = C.c_int()
        L.check(self.lib.zr_equation_count(self._mv, C.byref(m)))
        return m.value

    def solution(self):
        n = self.n
        M = max(self.equation_count(), 1)
        t = np.zeros((n, 2))
        rel = np.zeros((n - 1, 2))
        w = np.zeros(M)
        lg = np.zeros(3)
        err = np.zeros((M, 2))
        es = np.zeros(2)
        exy = np.zeros((n, 2))
        L.check(self.lib.zr_get_solution(self._mv, L.dptr(t), L.dptr(rel), L.dptr(w), L.dptr(lg), L.dptr(err),
                                         L.dptr(es), L.dptr(exy)))
        M = self.equation_count()
        w, err = w[:M], err[:M]
        return dict(translations=t, relativeEst=rel, weights=w, logistic=lg, errorXY=err,
                    mean_errorNorm=es[0], std_errorNorm=es[1], errorX=exy[:, 1], errorY=exy[:, 0])

    def peaksig_hist(self):
        """cdfLogisticCurve's histogram (errorDict pdfPeaks, cdfPeaks, hSigma,
        zorro/zorro.py:1991-1998) or None without autologistic weights."""
        nb = C.c_int()
        L.check(self.lib.zr_get_peaksig_hist(self._mv, None, None, None, 0, C.byref(nb)))
        if nb.value < 1:
            return None
        pdf, cdf, hs = (np.zeros(nb.value, np.float32) for _ in range(3))
        L.check(self.lib.zr_get_peaksig_hist(self._mv, L.fptr(pdf), L.fptr(cdf), L.fptr(hs), nb.value,
                                             C.byref(nb)))
        return dict(pdfPeaks=pdf, cdfPeaks=cdf, hSigma=hs)

    def background_fit(self):
        """The last 'background' step's fitted surface: dict(p=(offset, centre
        x, centre y, amplitude, width x, width y), error, status) (status 0:
        converged and subtracted)."""
        p = np.zeros(6)
        fv = C.c_double()
        st = C.c_int()
        L.check(self.lib.zr_get_background_fit(self._mv, p.ctypes.data_as(L._D), C.byref(fv), C.byref(st)))
        return dict(p=p, error=fv.value, status=st.value)

    def solver_info(self):
        d = (C.c_int * 6)()
        L.check(self.lib.zr_get_solver_info(self._mv, d))
        return dict(zip(["n_peaksig", "bins", "status_y", "status_x", "pivots_y", "pivots_x"], list(d)))

    def sums(self):
        """imageSum, evenSum, oddSum and (when the dose filter ran) filtSum."""
        h, w = self.hb, self.wb
        out = [np.zeros((h, w), np.float32) for _ in range(4)]
        ptrs = [L.fptr(a) for a in out]
        filt = self.params.n_filters > 0
        if not filt:
            ptrs[1] = None
        L.check(self.lib.zr_get_sums(self._mv, *ptrs))
        res = dict(imageSum=out[0], evenSum=out[2], oddSum=out[3])
        if filt:
            res["filtSum"] = out[1]
        return res

    def frc_curve(self):
        n = C.c_int()
        buf = np.zeros(4096)
        L.check(self.lib.zr_get_frc(self._mv, L.dptr(buf), C.byref(n)))
        return buf[:n.value].copy()


def tiled_frc(img1, img2, trans=None, box=512, overlap=0.5, device=0):
    """tiledFRC of two images on the GPU (``zorro/zorro.py:2246-2310``):
    returns the whole rotational-mean FRC curve (float64, ceil(sqrt(2) box/2) + 1
    values).  ``img1``/``img2``: float32 [h, w] NumPy arrays or CUDA tensors;
    ``trans``: [n, 2] translations for the crop limits (None: none)."""
    import torch
    lib = L.load()
    dev = torch.device("cuda", device)

    def on_dev(a):
        t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a, np.float32))
        return t.to(dev, torch.float32).contiguous()

    a, b = on_dev(img1), on_dev(img2)
    if a.shape != b.shape or a.dim() != 2:
        raise ValueError(f"tiledFRC needs two images of one shape (got {tuple(a.shape)}, {tuple(b.shape)})")
    h, w = (int(v) for v in a.shape)
    t = None if trans is None else np.ascontiguousarray(np.asarray(trans, np.float64)[:, :2])
    nr = int(np.ceil(np.sqrt(2.0) * (box // 2))) + 2
    out = torch.zeros(nr, dtype=torch.float64, device=dev)
    n = C.c_int()
    stream = torch.cuda.current_stream(dev).cuda_stream
    L.check(lib.zr_tiled_frc(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), h, w,
                             None if t is None else t.ctypes.data_as(C.POINTER(C.c_double)),
                             0 if t is None else t.shape[0], int(box), float(overlap),
                             C.c_void_p(out.data_ptr()), C.byref(n), C.c_void_p(stream)))
    return out[:n.value].cpu().numpy()
