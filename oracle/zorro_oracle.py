"""FP64 NumPy/SciPy restatement of Zorro's drift-registration path (the oracle).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.  The product package
never imports this module.

Every public function cites the reference lines it restates (paths relative
to the reference repository root).  Quirks of the reference that decide
parity are carried on purpose and named where they occur:

* raw (unnormalised) inverse transforms -- pyFFTW ``execute()``
  (``zorro/zorro.py:1214-1216`` compensates for it)
* the origin-suppression stencil with duplicated terms (``zorro/zorro.py:1079-1083``)
* the ``sign(round_pos)`` sub-pixel bias and the asymmetric 16x16 patch
  (``zorro/zorro.py:1180-1216``)
* the rot90 construction of the equation matrix (``zorro/zorro.py:1296-1311``)
* Lanczos ``applyShifts`` = banker's-rounded roll + 3x3 reflect convolution
  (``zorro/zorro_util.py:296-379``)
* the dose-filter Nyquist-corner replacement (``zorro/zorro.py:2422``) and the
  raw, modulus inverse (``zorro/zorro.py:2487-2488``)
* the FRC crop limits computed with ``min(floor(.), 5)`` (``zorro/zorro.py:2266-2274``)

Two places of the reference are stochastic (``scipy.optimize.basinhopping`` in
``shiftsSolver`` ``zorro/zorro.py:1407-1425`` and in ``cdfLogisticCurve``
``zorro/zorro.py:1981-2022``).  The oracle replaces them with deterministic
optimisers of the same objectives:

* the weighted-L1 shift problem is solved exactly (``scipy.optimize.linprog``,
  HiGHS) and the optimal vertex is then re-evaluated exactly by summing the
  measured shifts along a spanning tree of the exactly-satisfied equations;
* the logistic CDF fit uses a fixed coarse grid followed by a shrinking 3x3x3
  pattern search (``logistic_fit``), whose objective is checked in the tests
  to be no worse than the reference's basinhopping on the same data.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import scipy.ndimage
import scipy.optimize

__all__ = [
    "Params", "tri_indices", "pad_stack", "shape_cropped", "fourier_crop",
    "raw_ifft2", "frame_spectra", "mask_spectrum", "mnxc2_window", "background_estimate",
    "mnxc2_window_unmasked", "window_from_full", "b_filter", "locate_peak",
    "xcorrnm2_tri", "build_equations", "logistic", "logistic_objective",
    "logistic_fit", "cdf_logistic_curve", "l1_solve", "exact_vertex",
    "shifts_solver", "lanczos_kernel", "apply_shifts", "fourier_shift", "rotmean",
    "hann_apodization", "frc_crop_limits", "tiled_frc", "lazy_frc",
    "dose_filter", "gauss_lowpass", "filter_steps", "orient_gain", "apply_gain", "bin_stack", "align_image_stack", "PINS",
    "HOTPIX_DEFAULTS", "hotpix_psf_kernel", "hotpix_norm_objective", "hotpix_stats", "hotpix_filter",
    "even_odd_frc",
]

# Which oracle function is pinned by what (mirrored in tests/test_oracle_pins.py).
PINS = {
    "raw_ifft2": "closed forms: delta, constant; Parseval",
    "background_estimate": "started at a 2D Gaussian surface drawn in the fit's coordinates, the fit keeps it "
                           "and the full-grid surface reproduces the frame (square); the y-centring quirk on "
                           "non-square frames; a constant frame fails the fit (NaN) and gives no background",
    "frame_spectra/fourier_crop": "crop of a band-limited image == decimation; DC == sum",
    "mnxc2_window": "brute-force direct masked NCC on tiny frames; roll KAT (peak at d, C=1)",
    "mnxc2_window_unmasked": "equals mnxc2_window with an all-ones mask (Q2 closed form)",
    "b_filter": "scipy.ndimage.gaussian_filter == explicit reflect-padded separable sum",
    "locate_peak": "brute-force 256x256 DFT of the 16x16 patch; KAT on a band-limited peak",
    "build_equations": "tri/diag KAT: row (I,J) sums x_I..x_{J-1}",
    "l1_solve/exact_vertex": "noise-free consistent system -> exact truth; objective <= random restarts",
    "logistic_fit": "objective <= scipy basinhopping (seeded) on the same CDF",
    "lanczos_kernel/apply_shifts": "integer shifts == np.roll; kernel probe values (t=0.7 -> +1.0095 px)",
    "rotmean": "radially symmetric input returns its profile; constant -> constant",
    "tiled_frc": "identical halves -> FRC == 1; independent noise -> ~0",
    "dose_filter": "N=1 closed form per frequency; huge critical dose -> filtSum == |image| (scaled)",
    "bin_stack": "band-limited image -> 4 x its even samples (2x crop); constant c -> 4c; brute-force DFT",
    "orient_gain/apply_gain": "numpy rot90/flip identities on an index ramp; element-wise product",
    "locate_peak (preShift roll)": "two-peak whole-map KAT: unrolled vs rolled disc; brute-force disc statistics",
    "hotpix_filter/hotpix_stats": "PSF kernel sum/ratios; fit of a binned normal = (mu - 1/2, sigma); planted "
                                  "singletons/clusters vs explicit PSF and neighbour means; cut-offs on a large stack",
    "xcorrnm2_tri (triIndices)": "a mask equal to the 'diag' band reproduces triMode 'diag' exactly "
                                 "(tests/test_tri_indices.py)",
    "shifts_solver (error analysis)": "consistent system -> every error 0; one planted equation: errorXY, mean/std "
                                      "and per-frame errorX counted over the pair list (row J, column J-1)",
    "hann_apodization": "closed-form values at r = 0, R/4, R/2, R and outside the circle; radial, decreasing",
    "cdf_logistic_curve": "hand-counted 3-bin histogram, CDF and left edges of a fixed sample",
    "dose_filter (normalize)": "constant frames: DC divided by N+1; cosine gain g of the plain filter -> g/(g+1)",
    "gauss_lowpass": "explicit symmetric-padded separable sum of the normalised radius-12 taps; constant -> constant",
    "fourier_shift": "integer shift == np.roll (of the zero-padded frame); shift then unshift == identity on a "
                     "band-limited image; fractional shift of a cosine == its closed form; Nyquist row -> cos(pi t)",
}


@dataclass
class Params:
    """Hot-path knobs of ``ImageRegistrator`` (``zorro/zorro.py:82-243``).

    Defaults follow the reference except where it cannot run:
    ``shapePadded`` is an explicit pair (the reference divides the *list*
    ``[4096,4096]`` by a float for ``Bmode='conv'``, ``zorro/zorro.py:1136``).
    """
    triMode: str = "diag"                # :207
    diagWidth: int = 5                   # :211
    diagStart: int = 0                   # :210
    autoMax: int = 10                    # :212
    maxShift: int = 100                  # :191
    preShift: bool = False               # :194
    Bmode: str = "opti"                  # :117
    Brad: float = 512.0                  # :116
    subPixReg: int = 16                  # :187
    shiftMethod: str = "lanczos"         # :190
    weightMode: str = "autologistic"     # :199
    peaksigThres: float = 6.0            # :200
    logisticK: float = 5.0               # :201
    logisticNu: float = 0.15             # :202
    originMode: str = "centroid"         # :203
    suppressOrigin: bool = True          # :204
    fouCrop: tuple = (3072, 3072)        # :120
    shapeBinned: Optional[tuple] = None  # :185
    shapePadded: tuple = (4096, 4096)    # :184
    filterMode: Optional[str] = None     # :135
    doseFiltParam: tuple = (None, 0.24499, -1.6649, 2.8141, 32, 0)  # :137
    pixelsize: Optional[float] = None    # nm, :103
    voltage: float = 300.0               # :104
    doLazyFRC: bool = True               # :164
    frc_box: int = 512                   # :2226
    frc_overlap: float = 0.5             # :2226
    hotpixInfo: Optional[dict] = None    # :139-141 ('hot' in filterMode)
    hotMuSigma: Optional[tuple] = None   # fixed normal-CDF fit of hotpixFilter (test override)


# ---------------------------------------------------------------------------
# Frame bookkeeping
# ---------------------------------------------------------------------------

def tri_indices(N: int, triMode: str = "tri", diagWidth: int = 5, diagStart: int = 0) -> np.ndarray:
    """Boolean [N,N] pair mask.  ``zorro/zorro.py:559-581``."""
    xm, ym = np.meshgrid(np.arange(N), np.arange(N))
    d = xm - ym
    if triMode == "first":
        t = np.ones([1, N], dtype=bool)
        t[0, 0] = False
        return t
    if triMode == "diag":
        if diagWidth is None or diagWidth < 0:
            diagWidth = N
        return (d <= diagWidth + diagStart) & (d > diagStart)
    if triMode in ("autocorr", "refine"):
        return d == 0
    return d >= 1                                    # 'tri' and 'auto'


def pad_stack(images: np.ndarray, shapePadded, interiorPad: int = 0):
    """Zero-pad frames into the top-left corner and build the single mask.

    ``zorro/zorro.py:1908-1958``: mask is bool [1,Hp,Wp], ones over the
    original frame (inset by ``interiorPad``).
    """
    n, h, w = images.shape
    hp, wp = int(shapePadded[0]), int(shapePadded[1])
    if hp < h or wp < w:
        raise ValueError("shapePadded %s smaller than frames %s" % ((hp, wp), (h, w)))
    out = np.zeros([n, hp, wp], dtype=np.float64)
    out[:, :h, :w] = images
    mask = np.zeros([1, hp, wp], dtype=bool)
    if interiorPad > 0:
        mask[0, interiorPad:h - interiorPad, interiorPad:w - interiorPad] = True
    else:
        mask[0, :h, :w] = True
    return out, mask


def shape_cropped(fouCrop, shapePadded):
    """``zorro/zorro.py:598-608``: cropped size is ``2*floor(fouCrop/2)``, or the
    padded size when ``fouCrop`` is all zero."""
    if fouCrop is None or not np.any(fouCrop):
        return int(shapePadded[0]), int(shapePadded[1])
    c = 2 * np.floor(np.asarray(fouCrop, dtype=float) / 2.0).astype(int)
    if c[0] > shapePadded[0] or c[1] > shapePadded[1]:
        raise ValueError("fouCrop must not exceed shapePadded")
    return int(c[0]), int(c[1])


def fourier_crop(F: np.ndarray, shape) -> np.ndarray:
    """Four-quadrant Fourier crop.  ``zorro/zorro.py:821-824``, ``:901-904``."""
    hc, wc = shape
    h2, w2 = hc // 2, wc // 2
    out = np.empty((hc, wc), dtype=F.dtype)
    out[:h2, :w2] = F[:h2, :w2]
    out[:h2, -w2:] = F[:h2, -w2:]
    out[-h2:, :w2] = F[-h2:, :w2]
    out[-h2:, -w2:] = F[-h2:, -w2:]
    return out


def raw_ifft2(X: np.ndarray) -> np.ndarray:
    """pyFFTW ``execute()`` backward transform: the raw sum, no 1/N (Q1)."""
    return np.fft.ifft2(X) * (X.shape[0] * X.shape[1])


# ---------------------------------------------------------------------------
# Correlation (xcorrnm2_tri / mnxc2)
# ---------------------------------------------------------------------------

def mask_spectrum(mask: np.ndarray, shapeC):
    """Cropped mask spectrum and ``maskProduct``.  ``zorro/zorro.py:814-835``."""
    hc, wc = shapeC
    nc2 = 1.0 / (float(hc) * float(wc)) ** 2
    M = fourier_crop(np.fft.fft2(mask.astype(np.float64)), shapeC)
    mp = nc2 * np.real(raw_ifft2(np.conj(M) * M))
    return M, mp


def frame_spectra(images: np.ndarray, mask: np.ndarray, shapeC):
    """Cropped spectra of m*f and (m*f)**2 per frame.  ``zorro/zorro.py:870-927``."""
    n = images.shape[0]
    F = np.empty((n,) + tuple(shapeC), dtype=np.complex128)
    Q = np.empty_like(F)
    for i in range(n):
        t = mask * images[i]
        F[i] = fourier_crop(np.fft.fft2(t), shapeC)
        Q[i] = fourier_crop(np.fft.fft2(t * t), shapeC)
    return F, Q


def window_from_full(C: np.ndarray, m, suppressOrigin: bool = True):
    """Origin value, Q3 origin suppression, ifftshift and crop to the 2m x 2m lag
    window (index i <-> lag i-m); m=None keeps the whole shifted map (preShift,
    ``:1089-1092``).  ``zorro/zorro.py:1078-1092``."""
    C = C.copy()
    origin = C[0, 0]
    if suppressOrigin:
        # duplicated C[-1,0] and C[-1,1]; C[0,-1] and C[1,-1] absent (Q3)
        C[0, 0] = 0.125 * (C[1, 0] + C[0, 1] + C[-1, 0] + C[-1, 0] +
                           C[1, 1] + C[-1, 1] + C[-1, 1] + C[-1, -1])
    C = np.fft.ifftshift(C)
    if m is None:
        return C, origin
    h2, w2 = C.shape[0] // 2, C.shape[1] // 2
    return C[h2 - m:h2 + m, w2 - m:w2 + m].copy(), origin


def mnxc2_window(FI, QI, FJ, QJ, M, mp, shapeC, m, suppressOrigin=True):
    """Masked, intensity-normalised cross-correlation of template I and base J,
    general (single-mask) form.  ``zorro/zorro.py:1014-1097``."""
    hc, wc = shapeC
    nc2 = 1.0 / (float(hc) * float(wc)) ** 2
    Ct = nc2 * np.real(raw_ifft2(M * np.conj(FI)))          # :1034-1037
    Cb = nc2 * np.real(raw_ifft2(np.conj(M) * FJ))          # :1040-1044
    t1 = np.real(raw_ifft2(M * np.conj(QI)))                # :1049
    t2 = np.real(raw_ifft2(np.conj(M) * QJ))                # :1057
    with np.errstate(invalid="ignore", divide="ignore"):
        den = np.sqrt((t2 * nc2 - Cb * (Cb / mp)) * (t1 * nc2 - Ct * (Ct / mp)))  # :1061-1062
        den = np.clip(den, 1, np.inf)                                               # :1065
        num = np.real(raw_ifft2(FJ * np.conj(FI)))                                  # :1069-1070
        C = (num * nc2 - Ct * Cb / mp) / den                                        # :1074
    return window_from_full(C, m, suppressOrigin)


def mnxc2_window_unmasked(FI, FJ, sum_i, sumsq_i, sum_j, sumsq_j, shape_padded, shapeC, m,
                          suppressOrigin=True):
    """Closed form of ``mnxc2`` for one all-ones mask without padding (Q2):
    C(d) = nc2 * Re R[conj(F_I) F_J with DC removed](d) / max(1, D), with
    D = nc2 * sqrt((Np*SQ_J - S_J^2)(Np*SQ_I - S_I^2)), Np = Hp*Wp."""
    hc, wc = shapeC
    nc2 = 1.0 / (float(hc) * float(wc)) ** 2
    npix = float(shape_padded[0]) * float(shape_padded[1])
    X = np.conj(FI) * FJ
    X[0, 0] = 0.0
    with np.errstate(invalid="ignore"):
        den = nc2 * math.sqrt(max(npix * sumsq_j - sum_j * sum_j, 0.0) * max(npix * sumsq_i - sum_i * sum_i, 0.0))
    den = max(den, 1.0)
    C = nc2 * np.real(raw_ifft2(X)) / den
    return window_from_full(C, m, suppressOrigin)


def b_filter(C: np.ndarray, Bmode: str, Brad: float, shapePadded):
    """B-factor low-pass of the lag window.  ``zorro/zorro.py:1104-1141``."""
    if Bmode == "opti":
        def inverse_peak_contrast(s):
            cf = scipy.ndimage.gaussian_filter(C, s)
            return np.std(cf) / (np.max(cf) - np.mean(cf))
        res = scipy.optimize.minimize_scalar(inverse_peak_contrast, bounds=[0.0, 7.0], method="bounded",
                                             options={"maxiter": 15, "xatol": 0.01})
        return scipy.ndimage.gaussian_filter(C, res.x)
    if Bmode == "fourier" and bool(Brad):
        raise NotImplementedError("Bmode='fourier' is broken in the reference (zorro/zorro.py:1125-1133)")
    if (bool(Brad) and Bmode == "conv") or Bmode == "convolution":
        sigma = np.asarray(shapePadded, dtype=float) / (np.sqrt(2) * np.pi * Brad)
        return scipy.ndimage.gaussian_filter(C, sigma)
    return C


def _py_slice(c, half, n):
    """Python slice ``[c-half : c+half]`` on an axis of length n -> (start, length)."""
    s, e = c - half, c + half
    if s < 0:
        s += n
        if s < 0:
            s = 0
    e = min(e, n)
    return s, max(0, e - s)


def locate_peak(C: np.ndarray, p: Params, shapePadded, roll=None):
    """Peak, significance and Fourier-interpolated sub-pixel shift.
    ``zorro/zorro.py:1099-1231``.  Returns (shift_yx, corr, peaksig, cmaxpos).
    ``roll`` (preShift): C is the whole shifted map and the maxShift disc,
    centred on the map, is rolled by ``roll`` = (dy, dx) (``:1148-1163``)."""
    Cf = b_filter(C, p.Bmode, p.Brad, shapePadded)
    m = p.maxShift
    if not m:
        raise NotImplementedError("maxShift=None is outside the supported path")
    if roll is None:
        xm, ym = np.meshgrid(np.arange(-m, m), np.arange(-m, m))
        disc = (xm * xm + ym * ym) < m ** 2                             # :841-851
    else:
        h2, w2 = C.shape[0] // 2, C.shape[1] // 2
        xm, ym = np.meshgrid(np.arange(-h2, h2), np.arange(-w2, w2))      # :842-844 (square maps)
        disc = np.roll(np.roll((xm * xm + ym * ym) < m ** 2, int(roll[0]), axis=0), int(roll[1]), axis=1)
    Cm = Cf * disc
    cmax = np.unravel_index(np.argmax(Cm), Cm.shape)                     # :1167
    vals = Cf[disc]
    peaksig = (Cm[cmax] - np.mean(vals)) / np.std(vals)                  # :1168
    half = np.array(C.shape) / 2.0
    if p.subPixReg > 1:
        subR = 8
        sy, ly = _py_slice(cmax[0], subR, C.shape[0])
        sx, lx = _py_slice(cmax[1], subR, C.shape[1])
        if ly == 2 * subR and lx == 2 * subR:
            P = int(p.subPixReg)
            Csub = Cf[sy:sy + 2 * subR, sx:sx + 2 * subR]
            S = np.fft.fftshift(np.fft.fft2(Csub))                          # :1186, :1196-1197
            pad = (P - 1) * subR
            Sp = np.pad(S, pad, mode="constant")
            over = raw_ifft2(np.fft.ifftshift(Sp))                           # :1198-1199
            a = np.abs(over)
            smax = np.unravel_index(np.argmax(a), a.shape)
            round_pos = np.array(cmax, dtype=float) - half
            rem = (np.array(smax, dtype=float) - np.array(a.shape) / 2.0 + np.sign(round_pos)) / P  # Q4
            shift = round_pos + rem
            corr = a[smax] / (subR * 2.0) ** 2
        else:                                                                # :1217-1220
            shift = np.array(cmax, dtype=float) - half
            corr = C[cmax]
    else:
        shift = np.array(cmax, dtype=float) - half
        corr = C[cmax]
    return shift, float(corr), float(peaksig), cmax


def xcorrnm2_tri(images: np.ndarray, mask: np.ndarray, p: Params, return_windows: bool = False, triIndices=None):
    """All-pairs registration.  ``zorro/zorro.py:773-1011``.

    ``images`` are already padded to ``p.shapePadded`` and ``mask`` is [1,Hp,Wp].
    ``triIndices``: boolean [n, n] pairs to correlate in place of the triMode's
    pattern (``:552-557``); the rows are then walked as for 'tri' (``:959-975``).
    """
    n = images.shape[0]
    shapeP = images.shape[1:]
    shapeC = shape_cropped(p.fouCrop, shapeP)
    if triIndices is not None:
        tri = np.asarray(triIndices).astype(bool)
        if tri.shape != (n, n):
            raise IndexError("triIndices is wrong size, should be of length: " + str(n))
    else:
        tri = tri_indices(n, p.triMode, p.diagWidth, p.diagStart)
    if p.triMode not in ("tri", "diag", "auto", "first", "refine"):
        raise NotImplementedError("triMode %r is outside the supported path" % p.triMode)
    unmasked = bool(mask.all())
    F, Q = frame_spectra(images, mask[0], shapeC)
    if not unmasked:
        M, mp = mask_spectrum(mask[0], shapeC)
    sums = images.reshape(n, -1).sum(axis=1)
    sumsq = (images.reshape(n, -1) ** 2).sum(axis=1)
    # trimats are float32 as in the reference (float_dtype, zorro/zorro.py:592-595)
    shifts = np.zeros([n, n, 2], dtype=np.float32)
    corr = np.zeros([n, n], dtype=np.float32)
    peaksig = np.zeros([n, n], dtype=np.float32)
    origin = np.zeros([n, n], dtype=np.float32)
    windows = {}
    if p.triMode == "refine":
        # template of frame I: (sum_k F_k - F_I)/N and (sum_k Q_k - Q_I)/N, base frame I,
        # for I < N-1 (zorro/zorro.py:934-956)
        sF, sQ = F.sum(axis=0), Q.sum(axis=0)
        for I in range(n - 1):
            TF, TQ = (sF - F[I]) / n, (sQ - Q[I]) / n
            if unmasked:
                s_t, q_t = (sums.sum() - sums[I]) / n, (sumsq.sum() - sumsq[I]) / n
                W, o = mnxc2_window_unmasked(TF, F[I], s_t, q_t, sums[I], sumsq[I], shapeP, shapeC, p.maxShift,
                                             p.suppressOrigin)
            else:
                W, o = mnxc2_window(TF, TQ, F[I], Q[I], M, mp, shapeC, p.maxShift, p.suppressOrigin)
            origin[I, I] = o
            s, c, ps, _ = locate_peak(W, p, shapeP)
            shifts[I, I] = s
            corr[I, I] = c
            peaksig[I, I] = ps
            if return_windows:
                windows[(I, I)] = W
        tri = np.zeros((0, n), bool)
    if p.preShift and (p.triMode == "refine" or shapeC[0] != shapeC[1]):
        raise NotImplementedError("preShift: 'refine' is not working in the reference (:1150) and non-square maps "
                                  "do not broadcast against its mask (:842-844)")
    mwin = None if p.preShift else p.maxShift
    for I in range(tri.shape[0]):          # one row for triMode first
        for J in np.nonzero(tri[I])[0]:
            if unmasked:
                W, o = mnxc2_window_unmasked(F[I], F[J], sums[I], sumsq[I], sums[J], sumsq[J],
                                             shapeP, shapeC, mwin, p.suppressOrigin)
            else:
                W, o = mnxc2_window(F[I], Q[I], F[J], Q[J], M, mp, shapeC, mwin, p.suppressOrigin)
            origin[I, J] = o
            # preShift: the disc follows the previous pair of the row (:1151-1153);
            # shiftsTriMat still holds cropped pixels here (rescaled after the loop)
            roll = np.round(shifts[I, J - 1]).astype(int) if p.preShift else None
            s, c, ps, _ = locate_peak(W, p, shapeP, roll)
            shifts[I, J] = s
            corr[I, J] = c
            peaksig[I, J] = ps
            if return_windows:
                windows[(I, J)] = W
            if p.triMode == "auto" and (ps <= p.peaksigThres or J - I >= p.autoMax):   # :992-996
                break
    if np.any(p.fouCrop):                                                  # :1004-1006 (float32 multiply)
        shifts[:, :, 0] *= np.float32(shapeP[0] / shapeC[0])
        shifts[:, :, 1] *= np.float32(shapeP[1] / shapeC[1])
    out = dict(shiftsTriMat=shifts, corrTriMat=corr, peaksigTriMat=peaksig, originTriMat=origin,
               shapeCropped=shapeC)
    if return_windows:
        out["windows"] = windows
    return out


# ---------------------------------------------------------------------------
# Shift solver
# ---------------------------------------------------------------------------

def build_equations(tri: np.ndarray):
    """Equation matrix from the boolean [N-1,N-1] pair mask via the reference's
    rot90 construction.  ``zorro/zorro.py:1290-1311``."""
    N = tri.shape[0]
    last_col = np.zeros(N, dtype=int)
    M = 0
    for I in range(N):
        if tri[I].any():
            last_col[I] = np.argwhere(tri[I])[-1][0] + 1
            M += last_col[I] - I
    A = np.zeros([M, N])
    pos = 0
    for I in range(N):
        rot = np.rot90(tri[I:last_col[I], I:last_col[I]], k=2)
        A[pos:pos + rot.shape[0], I:I + rot.shape[1]] = rot
        pos += rot.shape[0]
    return A


def logistic(h, thr, K, nu):
    """``zorro/zorro_util.py:858-859``."""
    return 1.0 - 1.0 / (1.0 + np.exp(-K * (-h + thr))) ** nu


def logistic_objective(x, hSigma, cdf):
    """``minLogistic`` without the float32 cast, summed in index order.
    ``zorro/zorro_util.py:862-863``."""
    return float(np.cumsum(np.abs(np.asarray(cdf, np.float64) - logistic(np.asarray(hSigma, np.float64), x[0], x[1], x[2])))[-1])


# Deterministic replacement for basinhopping in cdfLogisticCurve (documented in
# the module docstring).  The same constants are used by the HIP solver kernel.
LOGI_GRID = (32, 32, 32)
LOGI_ROUNDS = 30
LOGI_MAX_ITERS = 400
LOGI_K_BOUNDS = (0.1, 20.0)
LOGI_NU_BOUNDS = (0.05, 5.0)


def logistic_fit(hSigma, cdf, lo, hi):
    """Bounded fit of (thr, K, nu) minimising ``logistic_objective``.

    Stage 1: grid over thr (linear in [lo,hi]), K and nu (log-spaced in their
    bounds); first minimum in (thr,K,nu) C order wins.  Stage 2: a 3x3x3 pattern
    search in (thr, log K, log nu), candidates clamped to the bounds: move to the
    best candidate on strict improvement (first in C order on ties), otherwise
    halve the steps; stop after LOGI_ROUNDS halvings or LOGI_MAX_ITERS steps.
    """
    hSigma = np.asarray(hSigma).astype(np.float64)
    cdf = np.asarray(cdf).astype(np.float64)
    g0, g1, g2 = LOGI_GRID
    lk0, lk1 = math.log(LOGI_K_BOUNDS[0]), math.log(LOGI_K_BOUNDS[1])
    ln0, ln1 = math.log(LOGI_NU_BOUNDS[0]), math.log(LOGI_NU_BOUNDS[1])
    thr = lo + (hi - lo) * np.arange(g0) / (g0 - 1)
    lk = lk0 + (lk1 - lk0) * np.arange(g1) / (g1 - 1)
    ln = ln0 + (ln1 - ln0) * np.arange(g2) / (g2 - 1)
    T, LK, LN = np.meshgrid(thr, lk, ln, indexing="ij")
    obj = np.cumsum(np.abs(cdf[None, :] - logistic(hSigma[None, :], T.reshape(-1, 1), np.exp(LK).reshape(-1, 1),
                                                   np.exp(LN).reshape(-1, 1))), axis=1)[:, -1]
    best = int(np.argmin(obj))
    cur = np.array([T.reshape(-1)[best], LK.reshape(-1)[best], LN.reshape(-1)[best]])
    fcur = float(obj[best])
    step = np.array([(hi - lo) / (g0 - 1), (lk1 - lk0) / (g1 - 1), (ln1 - ln0) / (g2 - 1)])
    lob = np.array([lo, lk0, ln0])
    hib = np.array([hi, lk1, ln1])
    offs = np.array([[a, b, c] for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)], dtype=float)
    halvings = 0
    for _ in range(LOGI_MAX_ITERS):
        cand = np.clip(cur[None, :] + offs * step[None, :], lob, hib)
        vals = np.cumsum(np.abs(cdf[None, :] - logistic(hSigma[None, :], cand[:, :1], np.exp(cand[:, 1:2]),
                                                        np.exp(cand[:, 2:3]))), axis=1)[:, -1]
        k = int(np.argmin(vals))
        if vals[k] < fcur:
            cur, fcur = cand[k], float(vals[k])
        else:
            step = step * 0.5
            halvings += 1
            if halvings >= LOGI_ROUNDS:
                break
    return float(cur[0]), float(math.exp(cur[1])), float(math.exp(cur[2])), fcur


def cdf_logistic_curve(peaksigTriMat: np.ndarray):
    """Histogram CDF of positive peak significances and its logistic fit.
    ``zorro/zorro.py:1981-2022``."""
    ps = peaksigTriMat[peaksigTriMat > 0.0]          # float32, as the reference trimat
    bins = int(ps.size / 7.0)
    pdf, edges = np.histogram(ps, bins=bins)          # float32 edges (numpy dtype rules)
    hSigma = edges[:-1]
    pdf = pdf.astype(np.float32)                      # :1996-1998
    cdf = np.cumsum(pdf)
    cdf /= cdf[-1]
    thr, K, nu, f = logistic_fit(hSigma, cdf, float(ps.min()), float(ps.max()))
    return dict(peaksigThres=thr, logisticK=K, logisticNu=nu, pdfPeaks=pdf, cdfPeaks=cdf,
                hSigma=hSigma, fit_objective=f)


def l1_solve(A, b, w, bound):
    """Exact minimiser of sum(w*|A x - b|) with |x| <= bound (LP, HiGHS)."""
    M, n = A.shape
    c = np.concatenate([np.zeros(n), w])
    I = np.eye(M)
    A_ub = np.block([[A, -I], [-A, -I]])
    b_ub = np.concatenate([b, -b])
    bounds = [(-bound, bound)] * n + [(0, None)] * M
    res = scipy.optimize.linprog(c, A_ub=A_ub, b_ub=b_ub, bounds=bounds, method="highs")
    if res.status != 0:
        raise RuntimeError("L1 solve failed: %s" % res.message)
    return res.x[:n]


def exact_vertex(A, b, x, tol=1e-7):
    """Re-evaluate an optimal vertex exactly: pick a spanning tree of the
    equations with |residual| < tol (smallest first) in position space
    (p_k = sum_{j<k} x_j, a contiguous row [lo,hi) is the edge lo-hi) and sum
    the measured shifts along it.  Returns x or None if no spanning tree."""
    M, n = A.shape
    r = np.abs(A @ x - b)
    edges = []
    for i in range(M):
        nz = np.nonzero(A[i])[0]
        if nz.size == 0 or nz[-1] - nz[0] + 1 != nz.size or not np.all(A[i, nz] == 1):
            return None
        edges.append((nz[0], nz[-1] + 1))
    order = np.argsort(r, kind="stable")
    parent = list(range(n + 1))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    adj = [[] for _ in range(n + 1)]
    used = 0
    for i in order:
        if r[i] >= tol:
            break
        lo, hi = edges[i]
        ra, rb = find(lo), find(hi)
        if ra != rb:
            parent[ra] = rb
            adj[lo].append((hi, b[i]))
            adj[hi].append((lo, -b[i]))
            used += 1
            if used == n:
                break
    if used != n:
        return None
    p = np.full(n + 1, np.nan)
    p[0] = 0.0
    stack = [0]
    while stack:
        u = stack.pop()
        for v, d in adj[u]:
            if np.isnan(p[v]):
                p[v] = p[u] + d
                stack.append(v)
    return np.diff(p)


def shifts_solver(shiftsTriMat, corrTriMat, peaksigTriMat, p: Params, shape_image):
    """Weighted-L1 shift solver with the reference's error analysis.
    ``zorro/zorro.py:1260-1567`` (+ ``cdfLogisticCurve`` ``:1981-2022``)."""
    S = shiftsTriMat[:-1, 1:, :]
    Cr = corrTriMat[:-1, 1:]
    Ps = peaksigTriMat[:-1, 1:]
    tri = Cr.astype(bool)
    ed = dict(corrTriMat=corrTriMat, peaksigTriMat=peaksigTriMat)
    N = tri.shape[0]
    A = build_equations(tri)
    M = A.shape[0]
    vec = np.arange(tri.size)[np.ravel(tri)]
    unr = np.unravel_index(vec, [N, N])
    b_x = np.ravel(S[tri, 1]).astype(np.float64)
    b_y = np.ravel(S[tri, 0]).astype(np.float64)
    if M != b_x.size:
        raise ValueError("equation count %d != measured pairs %d (reference would fail in np.dot)" % (M, b_x.size))
    ps = np.ravel(Ps[tri]).astype(np.float64)
    wm = p.weightMode
    if wm is None or wm == "corr":
        weights = ps.copy()
    elif wm == "unweighted":
        weights = np.ones_like(ps)
    elif wm in ("norm", "normalized"):
        weights = (ps - ps.min()) / (ps.max() - ps.min())
    elif wm == "autologistic":
        fit = cdf_logistic_curve(peaksigTriMat)
        ed.update({k: fit[k] for k in ("pdfPeaks", "cdfPeaks", "hSigma")})
        ed.update(logisticNu=fit["logisticNu"], logisticK=fit["logisticK"], peaksigThres=fit["peaksigThres"])
        weights = logistic(ps, fit["peaksigThres"], fit["logisticK"], fit["logisticNu"])
    elif wm == "logistic":
        weights = logistic(ps, p.peaksigThres, p.logisticK, p.logisticNu)
    else:
        weights = ps.copy()
    accepted = np.ones_like(b_x, dtype=bool)
    bound = np.min([shape_image[0] / 2.0, shape_image[1] / 2.0, N * p.maxShift]) if p.maxShift else \
        np.min([shape_image[0] / 2.0, shape_image[1] / 2.0])
    rel = np.zeros([N, 2])
    for ax, b in ((1, b_x), (0, b_y)):
        x = l1_solve(A, b, weights * accepted, bound)
        xe = exact_vertex(A, b, x)
        if xe is not None and np.sum(weights * np.abs(A @ xe - b)) <= np.sum(weights * np.abs(A @ x - b)) + 1e-9:
            x = xe
        rel[:, ax] = x
    # weighted error analysis :1483-1547
    w = np.clip(weights, 1e-6, np.inf)
    exy = np.zeros([M, 2])
    exy[:, 1] = A @ rel[:, 1] - b_x
    exy[:, 0] = A @ rel[:, 0] - b_y
    en = np.sqrt(exy[:, 0] ** 2 + exy[:, 1] ** 2)
    aen = en[accepted]
    mean_en = np.sum(w * aen) / np.sum(w)
    std_en = np.sqrt(np.sum(w * aen ** 2) / np.sum(w) - mean_en ** 2)
    eun = np.zeros([N, N]); wun = np.zeros([N, N]); exun = np.zeros([N, N]); eyun = np.zeros([N, N])
    eun[unr] = en; wun[unr] = w; exun[unr] = np.abs(exy[:, 1]); eyun[unr] = np.abs(exy[:, 0])
    acc_un = np.zeros([N, N]); acc_un[unr] = accepted
    pad = ((0, 1), (1, 0))
    acc_un = np.pad(acc_un, pad); exun = np.pad(exun, pad) * acc_un; eyun = np.pad(eyun, pad) * acc_un
    wun = np.pad(wun, pad)
    errorX = np.zeros(N + 1); errorY = np.zeros(N + 1)
    for J in range(N + 1):
        den = np.sum(wun[J, :]) + np.sum(wun[:, J - 1])
        errorX[J] = (np.sum(exun[J, :] * wun[J, :]) + np.sum(exun[:, J - 1] * wun[:, J - 1])) / den
        errorY[J] = (np.sum(eyun[J, :] * wun[J, :]) + np.sum(eyun[:, J - 1] * wun[:, J - 1])) / den
    ed.update(translations=-np.vstack((np.zeros([1, 2]), np.cumsum(rel, axis=0))), relativeEst=rel,
              acceptedEqns=accepted, weights=w, errorXY=exy, shiftsTriMat=shiftsTriMat, errorX=errorX,
              errorY=errorY, errorUnraveled=eun, mean_errorNorm=mean_en, std_errorNorm=std_en, M=M,
              Maccepted=M, Acoeff=A)
    return ed


# ---------------------------------------------------------------------------
# applyShifts, sums, FRC, dose filter
# ---------------------------------------------------------------------------

def lanczos_kernel(rem):
    """3x3 'Lanczos' kernel (lobes=2).  ``zorro/zorro_util.py:296-317``."""
    ks = 3
    lobes = (ks + 1) / 2
    xr = (2.0 / ks) * (np.arange(-ks / 2, ks / 2) + 1.0 - rem[1])
    yr = (2.0 / ks) * (np.arange(-ks / 2, ks / 2) + 1.0 - rem[0])
    xm, ym = np.meshgrid(xr, yr)
    k = np.sinc(xm * lobes) * np.sinc(xm) * np.sinc(ym * lobes) * np.sinc(ym)
    return k / np.sum(k)


def apply_shifts(images: np.ndarray, translations: np.ndarray, p: Params):
    """``zorro/zorro.py:1734-1808`` with ``lanczosIndexedShift``
    (``zorro/zorro_util.py:341-379``).  Returns (aligned, imageSum, translations)."""
    t = np.array(translations, dtype=np.float64)
    if p.originMode == "centroid":
        t = t - np.mean(t, axis=0)
    out = np.empty_like(images, dtype=np.float64)
    if p.subPixReg > 1 and p.shiftMethod == "fourier":
        for J in range(images.shape[0]):
            out[J] = fourier_shift(images[J], t[J], p.shapePadded)
        return out, out.sum(axis=0), t
    for J in range(images.shape[0]):
        it = np.round(t[J]).astype(int)
        g = np.roll(np.roll(images[J].astype(np.float64), it[0], axis=0), it[1], axis=1)
        if p.subPixReg > 1 and p.shiftMethod == "lanczos":
            rem = np.remainder(t[J], 1)
            if not (np.isclose(rem[0], 0.0) and np.isclose(rem[1], 0.0)):
                g = scipy.ndimage.convolve(g, lanczos_kernel(rem), mode="reflect")
        out[J] = g
    return out, out.sum(axis=0), t


def fourier_shift(img: np.ndarray, t, shapePadded):
    """shiftMethod='fourier' with the shift theorem (``zorro/zorro.py:1748-1772``
    as intended; the reference squares the y phase term and its xcorrnm2_tri
    fails first, Q10): the frame zero-padded to shapePadded (mask x frame, as
    the stored spectra of ``:884-895``), times exp(-2 pi i (fy ty + fx tx)) on
    the fftshifted frequency mesh (Nyquist at -1/2), the real part of the
    normalised inverse, cropped back to the frame."""
    h, w = img.shape
    hp, wp = (int(shapePadded[0]), int(shapePadded[1])) if shapePadded is not None else (h, w)
    pad = np.zeros((hp, wp))
    pad[:h, :w] = img
    xm, ym = np.meshgrid(np.arange(-wp / 2, wp / 2) / wp, np.arange(-hp / 2, hp / 2) / hp)
    ramp = np.fft.fftshift(np.exp(-2j * np.pi * (xm * t[1] + ym * t[0])))
    return np.real(np.fft.ifft2(np.fft.fft2(pad) * ramp))[:h, :w]


def rotmean(mage: np.ndarray):
    """Bilinear rotational mean about the centre.  ``zorro/zorro_util.py:147-224``."""
    M2, N2 = mage.shape[0] // 2, mage.shape[1] // 2
    rmax = int(np.ceil(np.sqrt(N2 ** 2 + M2 ** 2)) + 1)
    xm, ym = np.meshgrid(np.arange(-N2, N2), np.arange(-M2, M2))
    r = np.sqrt(xm ** 2 + ym ** 2)
    rf = np.floor(r)
    rem = (r - rf).ravel()
    rf = rf.ravel().astype(int)
    rc = rf + 1
    weights = np.bincount(rc, rem) + np.bincount(rf, 1.0 - rem, minlength=rmax)
    v = mage.ravel().astype(np.float64)
    rmean = np.bincount(rc, v * rem) + np.bincount(rf, v * (1.0 - rem), minlength=rmax)
    return rmean / weights, np.arange(weights.size)


def hann_apodization(box):
    """``apodization(name='hann', shape=box)``.  ``zorro/zorro_util.py:62-104``."""
    shape = np.asarray(box)
    radius = shape / 2.0
    xm, ym = np.meshgrid(np.arange(-shape[1] / 2, shape[1] / 2), np.arange(-shape[0] / 2, shape[0] / 2))
    rr = (xm / radius[1]) ** 2 + (ym / radius[0]) ** 2
    return (rr <= 1.0) * 0.5 * (1.0 + np.cos(np.pi * np.sqrt(rr)))


def frc_crop_limits(trans, minCrop=5):
    """``zorro/zorro.py:2264-2274`` (note ``min(floor(.),5)`` on the end crops)."""
    if not np.any(trans):
        return np.array([minCrop] * 4, dtype=float)
    yend = -np.minimum(np.floor(trans[:, 0].min()), minCrop)
    xend = -np.minimum(np.floor(trans[:, 1].min()), minCrop)
    ystart = np.maximum(np.ceil(trans[:, 0].max()), minCrop)
    xstart = np.maximum(np.ceil(trans[:, 1].max()), minCrop)
    return np.array([ystart, xstart, yend, xend], dtype=float)


def tiled_frc(img1, img2, trans, box=512, overlap=0.5):
    """Tiled, Hann-apodised 2D FRC and its rotational mean.
    ``zorro/zorro.py:2246-2310``.  Returns (FRC, FRC2D, n_tiles)."""
    if overlap > 0.8:
        overlap = 0.8
    c = frc_crop_limits(trans)
    hann = hann_apodization([box, box])
    tx = int(np.floor((img1.shape[1] - c[1] - c[3] - box) / box / (1.0 - overlap)))
    ty = int(np.floor((img1.shape[0] - c[0] - c[2] - box) / box / (1.0 - overlap)))
    frc2 = np.zeros([box, box])
    for I in range(ty):
        for J in range(tx):
            oy = int(I * box * (1.0 - overlap) + c[0])
            ox = int(J * box * (1.0 - overlap) + c[1])
            E = np.fft.fft2(hann * img1[oy:oy + box, ox:ox + box])
            O = np.fft.fft2(hann * img2[oy:oy + box, ox:ox + box])
            frc2 += np.real(E * np.conj(O)) / np.sqrt(np.real(O * np.conj(O)) * np.real(E * np.conj(E)))
    frc2 /= frc2[0, 0]
    frc2 = np.fft.fftshift(frc2)
    r, _ = rotmean(frc2)
    return r, frc2, tx * ty


def lazy_frc(aligned, trans, box=512, overlap=0.5):
    """Even/odd half-sum FRC of the aligned stack.  ``zorro/zorro.py:2226-2243``."""
    even = aligned[0::2].sum(axis=0)
    odd = aligned[1::2].sum(axis=0)
    r, frc2, nt = tiled_frc(even, odd, trans, box, overlap)
    return r[:int(box / 2)], dict(evenSum=even, oddSum=odd, FRC2D=frc2, n_tiles=nt)


def dose_filter(aligned, imageSum, p: Params, shapePadded, normalize: bool = False):
    """Unblur-style dose filter of the aligned frames.  ``zorro/zorro.py:2349-2494``;
    ``normalize=True`` divides the filtered sum by (sum of the per-frame filters
    + 1) before the inverse (``:2472-2486``).  Returns (filtSum, info)."""
    A, B, Cc, order = (float(p.doseFiltParam[i]) for i in (1, 2, 3, 4))
    if p.doseFiltParam[0] is not None:
        raise NotImplementedError("doseFiltParam[0] != None raises NameError in the reference (Q11)")
    if not p.voltage:
        raise ValueError("Accelerating voltage not set in advance for dose filtering")
    vs = np.sqrt(p.voltage / 300.0)
    m = aligned.shape[0]
    N, Mw = int(shapePadded[0]), int(shapePadded[1])
    ipx = 1.0 / (Mw * (p.pixelsize * 10))
    ipy = 1.0 / (N * (p.pixelsize * 10))
    xm, ym = np.meshgrid(np.arange(-Mw / 2, Mw / 2), np.arange(-N / 2, N / 2))
    q = np.fft.fftshift(np.sqrt(xm * xm * ipx ** 2 + ym * ym * ipy ** 2))
    dpf = np.mean(imageSum) / m
    missing = dpf * float(p.doseFiltParam[5])
    acc_dose = np.zeros(m + 1)
    acc_dose[1:] = np.cumsum(np.ones(m) * dpf)
    acc_dose += missing
    with np.errstate(divide="ignore"):
        crit = vs * (A * q ** B + Cc)
    crit[N // 2, Mw // 2] = crit[N // 2, Mw // 2 - 1] ** 2                 # Q: Nyquist corner
    qv = np.arange(0, N // 2) * np.sqrt(ipx * ipy)
    opti = np.zeros(N // 2)
    with np.errstate(divide="ignore"):
        opti[1:] = 2.51284 * vs * (A * qv[1:] ** B + Cc)
    opti[0] = opti[1] ** 2
    h, w = imageSum.shape
    acc = np.zeros([N, Mw], dtype=np.complex128)
    mag = np.zeros([N, Mw])
    thresQs = []
    for J in range(m):
        d1, d0 = acc_dose[J + 1], acc_dose[J]
        filt = np.exp(-0.5 * d1 / crit)
        idx = np.argwhere(np.abs(d1 - opti) < np.abs(d0 - opti))
        if idx.size == 0:
            raise IndexError("no optimal-dose radius for frame %d" % J)
        tq = qv[idx[-1][0]]
        thresQs.append(tq)
        thres = np.exp(-(q / tq) ** order)
        padded = np.pad(aligned[J].astype(np.float64), ((0, N - h), (0, Mw - w)), mode="symmetric")
        acc += np.fft.fft2(padded) * thres * filt
        mag += thres * filt                                                # :2479-2480
    if normalize:
        acc = acc * (1.0 / (mag + 1.0))                                    # :2482-2486
    filt_sum = np.abs(raw_ifft2(acc)[:h, :w])
    return filt_sum, dict(dosePerFrame=dpf, thresQ=np.array(thresQs), accumDose=acc_dose)


def gauss_lowpass(img):
    """filterMode 'gausslp': ``scipy.ndimage.gaussian_filter(filtSum, 3.0)``
    (reflect edges, truncate 4).  ``zorro/zorro.py:1707-1711``."""
    return scipy.ndimage.gaussian_filter(np.asarray(img, np.float64), 3.0)


def background_estimate(img, binFact: int = 128, lpSigma: float = 4.0, p_start=None):
    """``backgroundEstimate(fitType='gauss2')`` (``zorro/zorro_util.py:497-549``):
    the float32 image low-passed by ``gaussian_filter(sigma 4)``, sampled every
    ``binFact`` pixels (``interp2_nn`` at integer positions, ``:482-495``),
    cropped [1:-2, 1:-2], and a 2D Gaussian p0 + p3 exp(-(x-p1)^2/p4^2 -
    (y-p2)^2/p5^2) fitted under an L1 error with ``minimize(method='Powell')``
    from (mean, 1, 1, mean, mean*std, mean*std).  The fit's y coordinates are
    centred by the x sample count (``ycrop.shape[1]``, ``:512``).  Returns
    (background f64 [H, W] evaluated at ((x - W/2)/binFact, (y - H/2)/binFact),
    or zeros when the fit fails; the scipy result).  ``p_start`` replaces the
    reference's starting point (tests only)."""
    im = np.asarray(img, np.float32)
    h, w = im.shape
    xcrop, ycrop = np.meshgrid(np.arange(0, w, binFact), np.arange(0, h, binFact))
    g = scipy.ndimage.gaussian_filter(im, lpSigma)
    nn = g[np.clip(np.floor(ycrop).astype(int), 0, h - 1), np.clip(np.floor(xcrop).astype(int), 0, w - 1)]
    xmesh = (xcrop.astype(np.float32) / binFact - xcrop.shape[1] / 2.0)[1:-2, 1:-2]
    ymesh = (ycrop.astype(np.float32) / binFact - ycrop.shape[1] / 2.0)[1:-2, 1:-2]
    nn = nn[1:-2, 1:-2]

    def gauss2(q, x, y):
        xc, yc = x - q[1], y - q[2]
        return q[0] + q[3] * np.exp(-xc * xc / q[4] ** 2 - yc * yc / q[5] ** 2)

    def err(q, c):
        return np.sum(np.abs(c - gauss2(q, xmesh, ymesh)))

    p0 = np.ones(6, np.float64)
    p0[0] = np.mean(nn)
    p0[3] = np.mean(nn)
    p0[4] = np.mean(nn) * np.std(nn)
    p0[5] = p0[4]
    if p_start is not None:
        p0 = np.asarray(p_start, np.float64)
    fit = scipy.optimize.minimize(err, p0, method="Powell", args=(nn,))
    fit.samples, fit.xmesh, fit.ymesh = nn, xmesh, ymesh      # for checks of other fits on the same samples
    xback, yback = np.meshgrid(np.arange(w), np.arange(h))
    xback = (xback.astype(np.float32) - w / 2.0) / np.float32(binFact)
    yback = (yback.astype(np.float32) - h / 2.0) / np.float32(binFact)
    if fit.success:
        return gauss2(fit.x, xback, yback), fit
    return np.zeros((h, w)), fit


def filter_steps(filterMode):
    """The filterMode tokens alignImageStack applies after the registration, in
    order (``zorro/zorro.py:1692-1711``): lower-cased, spaces removed, split on
    commas; tokens other than dose/dosenorm/background/gausslp do nothing there."""
    if not filterMode:
        return []
    toks = filterMode.lower().replace(" ", "").split(",")
    return [t for t in toks if t in ("dose", "dosenorm", "gausslp", "background")]


# ---------------------------------------------------------------------------
# Gain reference and Fourier binning (the C4 input path)
# ---------------------------------------------------------------------------

def orient_gain(gainRef: np.ndarray, gainInfo: dict | None):
    """Rotations/flips of a loaded gain reference.  ``zorro/zorro.py:3486-3503``:
    'Diagonal' -> rot90 (k=1); then 'Horizontal' and 'Vertical' together ->
    rot90 (k=2), else 'Horizontal' -> fliplr, 'Vertical' -> flipud."""
    g = np.asarray(gainRef)
    info = gainInfo or {}
    if info.get("Diagonal"):
        g = np.rot90(g, k=1)
    if info.get("Horizontal") and info.get("Vertical"):
        g = np.rot90(g, k=2)
    elif info.get("Horizontal"):
        g = np.fliplr(g)
    elif info.get("Vertical"):
        g = np.flipud(g)
    return g


def apply_gain(raw: np.ndarray, gainRef: np.ndarray | None, gain: float | None = None):
    """``loadData`` for a stack (``zorro/zorro.py:3416-3419``, ``:3455-3482``):
    integer data cast to float32, times the (oriented) gain reference, divided
    by the scalar gain when it is set and not 1.  float32 arithmetic."""
    t = np.asarray(raw).astype(np.float32)
    if gainRef is not None:
        t = np.asarray(gainRef, np.float32) * t
    if gain and not np.isclose(gain, 1.0):
        t = t / np.float32(gain)
    return t


def bin_stack(images: np.ndarray, shapeBinned):
    """``binStack(binKernel='fourier')`` (``zorro/zorro.py:1866-1901``): per frame a
    forward transform at the full size, the four-quadrant crop to ``shapeBinned``
    (corner sizes shapeBinned // 2), scaling by 1/(Hb Wb), the raw inverse at the
    binned size and the real part.  Returns (binned, binScale)."""
    n, h, w = images.shape
    hb, wb = int(shapeBinned[0]), int(shapeBinned[1])
    b2 = (hb // 2, wb // 2)
    out = np.empty((n, hb, wb), dtype=np.float64)
    for J in range(n):
        F = np.fft.fft2(np.asarray(images[J], np.float64))
        Fb = np.zeros((hb, wb), dtype=np.complex128)
        Fb[:b2[0], :b2[1]] = F[:b2[0], :b2[1]]
        Fb[:b2[0], -b2[1]:] = F[:b2[0], -b2[1]:]
        Fb[-b2[0]:, :b2[1]] = F[-b2[0]:, :b2[1]]
        Fb[-b2[0]:, -b2[1]:] = F[-b2[0]:, -b2[1]:]
        Fb *= 1.0 / (hb * wb)
        out[J] = np.real(raw_ifft2(Fb))
    return out, np.array([h, w], float) / np.array([hb, wb], float)


# ---------------------------------------------------------------------------
# Outlier-pixel filter (filterMode containing 'hot')
# ---------------------------------------------------------------------------

HOTPIX_DEFAULTS = {"logisticK": 6.0, "relax": 0.925, "maxSigma": 8.0, "psf": "K2",
                   "cutoffLower": -4.0, "cutoffUpper": 3.25}          # zorro/zorro.py:139-141


def hotpix_psf_kernel(psf: str = "K2") -> np.ndarray:
    """3x3 point-spread kernel with a zero centre, normalised to sum 1
    (``zorro/zorro.py:2510-2519``)."""
    p1 = np.float32(0.173235968) if psf == "K2" else np.float32(1.0)
    k = np.array([[p1 * p1, p1, p1 * p1], [p1, 0.0, p1], [p1 * p1, p1, p1 * p1]], dtype=np.float32)
    return k / np.sum(k)


def hotpix_norm_objective(params, counts, cdf):
    """sum |cdf - Phi((c - mu)/sigma)| (``zorro/zorro.py:2550-2552``)."""
    import scipy.stats
    return float(np.sum(np.abs(cdf - scipy.stats.norm.cdf(counts, loc=params[0], scale=params[1]))))


def hotpix_stats(unaligned_sum: np.ndarray, info: dict, cutoffLower=None, cutoffUpper=None):
    """Histogram of the unaligned sum, its CDF, the bounded normal-CDF fit and the
    two sigma cut-offs (``zorro/zorro.py:2529-2596``).  ``info['mu_sigma']``, when
    present, replaces the L-BFGS-B fit (the GPU path takes the same override).
    Returns a dict with edges, hist, counts, cdf, mu, sigma, cutoffLower,
    cutoffUpper and the objective at (mu, sigma)."""
    import scipy.special
    s = np.asarray(unaligned_sum, np.float32)
    sum_mean = float(np.mean(s, dtype=np.float64))
    pstd = math.sqrt(sum_mean)
    max_sigma = float(info.get("maxSigma", 8.0))
    edges = np.arange(math.floor(sum_mean - max_sigma * pstd) - 0.5, math.ceil(sum_mean + max_sigma * pstd) + 0.5, 1)
    hist, edges = np.histogram(s, edges)
    hist = hist.astype(np.float32)
    counts = edges[:-1].astype(np.float32)
    counts += np.float32(0.5) * (counts[1] - counts[0])
    cdf = np.cumsum(hist, dtype=np.float32)
    cdf /= cdf[-1]
    if info.get("mu_sigma") is not None:
        mu, sigma = (float(v) for v in info["mu_sigma"])
    else:
        res = scipy.optimize.minimize(hotpix_norm_objective, (sum_mean, pstd), args=(counts, cdf), method="L-BFGS-B",
                                      bounds=((sum_mean - 0.5 * pstd, sum_mean + 0.5 * pstd),
                                              (0.7 * pstd, 1.3 * pstd)))
        mu, sigma = float(res.x[0]), float(res.x[1])
    sig_cdf = (np.float32(math.sqrt(2.0)) * scipy.special.erfinv(np.float32(2.0) * cdf - np.float32(1.0))).astype(np.float32)
    normal_sigma = ((counts - np.float32(mu)) / np.float32(sigma)).astype(np.float32)
    err = (normal_sigma - sig_cdf).astype(np.float32)
    keep = ~np.isinf(err)
    err, ns = err[keep], normal_sigma[keep]

    def fit(xs, ys):
        if xs.size == 0:
            raise ValueError("empty fit")
        A = np.stack([xs.astype(np.float64), np.ones(xs.size)]).T
        sol = np.linalg.lstsq(A, ys.astype(np.float64), rcond=None)[0]
        return np.float32(-sol[1] / sol[0])

    # given cut-offs skip their fits (hotpixFilter(cutoffLower, cutoffUpper), :2570, :2583);
    # an empty fit raised LinAlgError in the numpy of the reference's time, which
    # the reference catches and replaces by the default (:2579-2581, :2590-2592)
    if cutoffLower:
        cut_lo = np.float32(cutoffLower)
    else:
        try:
            lower_index = np.where(err > -0.5)[0][0]
            cut_lo = fit(ns[:lower_index], err[:lower_index])
        except (IndexError, ValueError, np.linalg.LinAlgError):
            cut_lo = np.float32(info.get("cutoffLower", -4.0))
    if cutoffUpper:
        cut_hi = np.float32(cutoffUpper)
    else:
        try:
            upper_index = np.where(err < 0.5)[0][-1]
            cut_hi = fit(ns[upper_index:], err[upper_index:])
        except (IndexError, ValueError, np.linalg.LinAlgError):
            cut_hi = np.float32(info.get("cutoffUpper", 3.25))
    return dict(edges=edges, hist=hist, counts=counts, cdf=cdf, mu=mu, sigma=sigma, cutoffLower=float(cut_lo),
                cutoffUpper=float(cut_hi), objective=hotpix_norm_objective((mu, sigma), counts, cdf),
                sumMean=sum_mean)


def hotpix_filter(images: np.ndarray, info: dict | None = None, neighbourThres: float = 0.01, cutoffLower=None,
                  cutoffUpper=None):
    """``hotpixFilter`` (``zorro/zorro.py:2497-2662``): logistic weights from the
    sigma of each pixel of the unaligned sum, singleton outliers replaced by the
    PSF-weighted mean of their neighbours, clustered outliers by the mean of
    their non-outlier 3x3 neighbours (or the fitted frame mean).  float32 per
    pixel as in the reference; ``decorrOutliers`` (random noise) is not
    restated.  Returns (filtered float32 stack, stats dict)."""
    info = {**HOTPIX_DEFAULTS, **(info or {})}
    imgs = np.asarray(images, np.float32)
    n, h, w = imgs.shape
    psf = hotpix_psf_kernel(info["psf"])
    usum = np.sum(imgs, axis=0, dtype=np.float32)
    st = hotpix_stats(usum, info, cutoffLower, cutoffUpper)
    mu, sigma = st["mu"], st["sigma"]
    usig = ((usum - np.float32(mu)) / np.float32(sigma)).astype(np.float32)
    lk, rel = np.float32(info["logisticK"]), np.float32(info["relax"])
    cl, cu = st["cutoffLower"], st["cutoffUpper"]
    # numexpr types (:2610-2611): the exponent and exp in float32 (float32
    # operands), the 1.0 constants make the rest float64, cast at the end
    with np.errstate(over="ignore"):
        e1 = np.exp(lk * (usig - np.float32(cl) * rel)).astype(np.float64)
        e2 = np.exp(lk * (usig - np.float32(cu) * rel)).astype(np.float64)
        lm = ((1.0 - 1.0 / (1.0 + e1)) / (1.0 + e2)).astype(np.float32)
    conv_lm = (np.float32(1.0) - lm).astype(np.float32)
    nmask = (conv_lm * scipy.ndimage.convolve(conv_lm, psf)).astype(np.float32)    # :2616
    single = (conv_lm * (nmask <= neighbourThres)).astype(np.float32)              # :2622
    umean = (usum / np.float32(n)).astype(np.float32)                              # :2624
    pfm = scipy.ndimage.convolve(umean, psf).astype(np.float32)                    # :2625
    nb = nmask > neighbourThres                                                    # :2631
    best_mean = mu / n
    nfilt = np.zeros_like(pfm)
    for ny, nx in zip(*np.nonzero(nb)):                                            # :2637-2645
        hood = nb[ny - 1:ny + 2, nx - 1:nx + 2]
        with np.errstate(invalid="ignore", divide="ignore"):
            ratio = np.sum(hood) / hood.size
        if ratio > 0.66 or ratio <= 0.001 or np.isnan(ratio):
            nfilt[ny, nx] = best_mean
        else:
            nfilt[ny, nx] = conv_lm[ny, nx] * np.mean(umean[ny - 1:ny + 2, nx - 1:nx + 2][~hood])
    out = (lm * imgs + single * pfm) + nfilt                                       # :2647
    st.update(guessDeadpix=int(np.sum(usig < np.float32(cl))), guessHotpix=int(np.sum(usig > np.float32(cu))),
              neighborPix=int(np.count_nonzero(nb)), frameMean=mu / n, frameStd=sigma / math.sqrt(n),
              logisticMask=lm, unalignedSum=usum)
    return out.astype(np.float32), st


# ---------------------------------------------------------------------------
# Orchestration
# ---------------------------------------------------------------------------

def even_odd_frc(images: np.ndarray, p: Params, box: int = 512, overlap: float = 0.5, edges: int = 64):
    """``evenOddFouRingCorr`` made runnable (``zorro/zorro.py:2099-2224``; the
    reference recurses through its saved configuration, Q12): even and odd
    frames registered on their own under ``util.edge_mask`` (rows/columns
    ``edges .. size - edges`` inclusive, ``zorro/zorro_util.py:123-134``), their
    sums aligned with triMode 'first', then ``tiledFRC`` of the two aligned sums
    with ``trans = hstack([transEven, transOdd])`` (its first two columns: the
    even half's translations).  Returns (FRC, transEven, transOdd)."""
    imgs = np.asarray(images, np.float64)
    n, h, w = imgs.shape
    ph = Params(**{**p.__dict__, "filterMode": None, "doLazyFRC": False})
    em = np.zeros((1,) + tuple(int(v) for v in p.shapePadded), bool)
    em[0, edges:h - edges + 1, edges:w - edges + 1] = True

    def half(frames):
        r = align_image_stack(frames, ph, mask=em)
        return r["translations"], r["imageSum"]

    te, se = half(imgs[0::2])
    to, so = half(imgs[1::2])
    pe = Params(**{**ph.__dict__, "triMode": "first"})
    r = align_image_stack(np.stack([se, so]), pe)
    aligned = r["aligned"]
    # (hstack needs halves of equal length; only its first two columns are read)
    trans = np.hstack([te, to]) if len(te) == len(to) else te
    frc, _, _ = tiled_frc(aligned[0], aligned[1], trans, box, overlap)
    return frc, te, to


def align_image_stack(images: np.ndarray, p: Params, mask: np.ndarray | None = None, triIndices=None):
    """``alignImageStack`` for ``xcorrMode='zorro'``.  ``zorro/zorro.py:1569-1716``.
    With ``shapeBinned`` set the stack is Fourier-binned first and the pixel
    size scaled by the mean bin factor (``:1591-1592``, ``:1835-1836``)."""
    images = np.asarray(images, dtype=np.float64)
    if p.shapeBinned is not None and np.any(p.shapeBinned):
        images, scale = bin_stack(images, p.shapeBinned)
        if p.pixelsize:
            p = Params(**{**p.__dict__, "pixelsize": p.pixelsize * float(np.mean(scale))})
    res_hot = None
    if p.filterMode and "hot" in p.filterMode.lower():                    # :1595-1596, after binning
        info = dict(p.hotpixInfo or {})
        if p.hotMuSigma is not None:
            info["mu_sigma"] = p.hotMuSigma
        filt, res_hot = hotpix_filter(images, info)
        images = filt.astype(np.float64)
    n, h, w = images.shape
    padded, fmask = pad_stack(images, p.shapePadded)
    xc = xcorrnm2_tri(padded, fmask if mask is None else mask, p, triIndices=triIndices)
    if p.triMode in ("first", "refine"):                                 # :1626-1647, no solver
        st = xc["shiftsTriMat"]
        t = -(st[0, :] if p.triMode == "first" else np.stack([st[k, k] for k in range(n)])).astype(np.float64)
        ed = dict(shiftsTriMat=st, corrTriMat=xc["corrTriMat"], peaksigTriMat=xc["peaksigTriMat"], translations=t)
    else:
        ed = shifts_solver(xc["shiftsTriMat"], xc["corrTriMat"], xc["peaksigTriMat"], p, padded.shape[1:])
    ed["originTriMat"] = xc["originTriMat"]
    aligned, image_sum, trans = apply_shifts(images, ed["translations"], p)
    res = dict(xcorr=xc, errorDict=ed, translations=trans, imageSum=image_sum, aligned=aligned, hotpix=res_hot)
    if p.doLazyFRC:
        frc, info = lazy_frc(aligned, trans, p.frc_box, p.frc_overlap)
        res.update(FRC=frc, FRC2D=info["FRC2D"], evenSum=info["evenSum"], oddSum=info["oddSum"],
                   frc_tiles=info["n_tiles"])
    filt = None
    for step in filter_steps(p.filterMode):                              # :1692-1711
        if step in ("dose", "dosenorm"):
            filt, info = dose_filter(aligned, image_sum, p, p.shapePadded, normalize=(step == "dosenorm"))
            res.update(dose=info)
        elif step == "background":                                       # :1702-1706, in float32
            base = np.asarray(image_sum if filt is None else filt, np.float32)
            back, _ = background_estimate(base)
            filt = (base.astype(np.float64) - back).astype(np.float32)
        else:                                                            # gausslp
            filt = gauss_lowpass(image_sum if filt is None else filt)
    if filt is not None:
        res.update(filtSum=filt)
    return res
