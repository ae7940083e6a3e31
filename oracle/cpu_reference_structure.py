"""CPU baseline: the reference pipeline's per-stage work, restated with the same
array sizes, dtypes and transform counts, timed on a bounded sample.

TEST / BASELINE INFRASTRUCTURE ONLY (imported by ``bench.py``'s cpu_baseline
leg).  It is labelled a "reference-structure CPU restatement", not the
reference: the reference itself cannot run here (pyFFTW, numexprz and mrcz are
absent).  What it keeps from ``zorro/zorro.py``:

* K1 (``:870-927``): per frame two complex64 forward FFTs at the padded size
  (m*f and (m*f)^2) and two four-quadrant crops;
* K2 (``:1014-1097``): per pair five raw complex64 inverse FFTs at the cropped
  size, the element-wise passes of the single-mask path (numexpr there, NumPy
  here), the full ``ifftshift`` and the window crop; ``locatePeak``
  (``:1099-1231``) with the 'conv' Gaussian, disc argmax and the 16^2 -> 256^2
  Fourier interpolation;
* the solver (``:1260-1567``): ``cdfLogisticCurve`` basinhopping (niter=50) and
  two basinhopping L-BFGS-B solves (niter=100) on the real equation count;
* ``applyShifts`` (``:1734-1808``): np.roll + 3x3 ``ndimage.convolve`` per frame
  in a thread pool, then the stack sum;
* ``tiledFRC`` (``:2246-2310``) and ``doseFilter`` (``:2349-2494``).

FFTs run through ``scipy.fft`` with ``workers=cores`` (FFTW threads in the
reference).  Stage times are measured on a few frames / pairs / hops and scaled
to the whole movie; the sample actually run is reported.
"""
from __future__ import annotations

import math
import os
import time
from multiprocessing.pool import ThreadPool

import numpy as np
import scipy.fft as sfft
import scipy.ndimage
import scipy.optimize

from . import zorro_oracle as O


def _crop4(F, hc, wc, out):
    h2, w2 = hc // 2, wc // 2
    out[:h2, :w2] = F[:h2, :w2]
    out[:h2, -w2:] = F[:h2, -w2:]
    out[-h2:, :w2] = F[-h2:, :w2]
    out[-h2:, -w2:] = F[-h2:, -w2:]


def time_movie(n_frames=40, shape=4096, crop=3072, max_shift=40, brad=150.0, cores=None, seed=0,
               frames_sample=2, pairs_sample=6, hops_sample=4):
    """Return (movies_per_min, details) for one movie of n_frames x shape^2."""
    cores = cores or os.cpu_count() or 1
    rng = np.random.default_rng(seed)
    H = W = shape
    hc = wc = crop
    P = n_frames * (n_frames - 1) // 2
    t = {}
    frames = [rng.poisson(1.0, size=(H, W)).astype(np.float32) for _ in range(frames_sample)]
    mask = np.ones((H, W), np.float32)
    # ---- K1: forward spectra (two FFTs per frame)
    base = np.empty((frames_sample, hc, wc), np.complex64)
    sq = np.empty((frames_sample, hc, wc), np.complex64)
    t0 = time.perf_counter()
    for i, f in enumerate(frames):
        tr = (mask * f).astype(np.complex64)
        _crop4(sfft.fft2(tr, workers=cores), hc, wc, base[i])
        _crop4(sfft.fft2((tr * tr).astype(np.complex64), workers=cores), hc, wc, sq[i])
    t["spectra"] = (time.perf_counter() - t0) / frames_sample * n_frames
    # mask spectrum terms (once per movie; not timed: negligible)
    M = np.empty((hc, wc), np.complex64)
    _crop4(sfft.fft2(mask.astype(np.complex64), workers=cores), hc, wc, M)
    nc2 = np.float32(1.0 / (float(hc) * wc) ** 2)
    mp = (nc2 * np.real(sfft.ifft2(np.conj(M) * M, workers=cores, norm="forward"))).astype(np.float32)
    Mc = np.conj(M)
    # ---- K2: per pair five raw inverse FFTs + element-wise passes + locatePeak
    sigma = np.array([H, W], float) / (math.sqrt(2) * math.pi * brad)
    xm, ym = np.meshgrid(np.arange(-max_shift, max_shift), np.arange(-max_shift, max_shift))
    disc = (xm * xm + ym * ym) < max_shift ** 2
    t0 = time.perf_counter()
    for k in range(pairs_sample):
        tI, tJ = np.conj(base[0]), base[1 % frames_sample]
        tQI, QJ = np.conj(sq[0]), sq[1 % frames_sample]
        ct = np.real(sfft.ifft2(M * tI, workers=cores, norm="forward")) * nc2
        cb = np.real(sfft.ifft2(Mc * tJ, workers=cores, norm="forward")) * nc2
        d1 = np.real(sfft.ifft2(M * tQI, workers=cores, norm="forward"))
        d2 = np.real(sfft.ifft2(Mc * QJ, workers=cores, norm="forward"))
        den = np.sqrt((d2 * nc2 - cb * (cb / mp)) * (d1 * nc2 - ct * (ct / mp)))
        den = np.clip(den, 1, np.inf)
        num = np.real(sfft.ifft2(tJ * tI, workers=cores, norm="forward"))
        C = (num * nc2 - ct * cb / mp) / den
        C[0, 0] = 0.125 * (C[1, 0] + C[0, 1] + 2 * C[-1, 0] + C[1, 1] + 2 * C[-1, 1] + C[-1, -1])
        C = np.fft.ifftshift(C)
        C = C[hc // 2 - max_shift:hc // 2 + max_shift, wc // 2 - max_shift:wc // 2 + max_shift]
        Cf = scipy.ndimage.gaussian_filter(C, sigma)
        Cm = Cf * disc
        c = np.unravel_index(np.argmax(Cm), Cm.shape)
        _ = (Cm[c] - np.mean(Cf[disc])) / np.std(Cf[disc])
        sub = Cf[max_shift - 8:max_shift + 8, max_shift - 8:max_shift + 8].astype(np.complex64)
        S = np.fft.fftshift(sfft.fft2(sub))
        Sp = np.zeros((256, 256), np.complex64)
        Sp[120:136, 120:136] = S
        a = np.abs(sfft.ifft2(np.fft.ifftshift(Sp), norm="forward"))
        _ = np.unravel_index(np.argmax(a), a.shape)
    t["pairs"] = (time.perf_counter() - t0) / pairs_sample * P
    # ---- solver: basinhopping as in the reference, on the real equation count
    N = n_frames - 1
    tri = np.triu(np.ones((N, N), bool))
    A = O.build_equations(tri)
    x_true = rng.normal(0.8, 0.3, N)
    b = A @ x_true + rng.normal(0, 0.1, A.shape[0])
    w = rng.uniform(0.2, 1.0, A.shape[0])
    ps = np.concatenate([rng.normal(4, 1, P // 3), rng.normal(12, 2, P - P // 3)])
    ps = ps[ps > 0]
    pdf, e = np.histogram(ps, bins=int(ps.size / 7.0))
    cdf = np.cumsum(pdf.astype(np.float32))
    cdf /= cdf[-1]
    bound = min(H / 2.0, W / 2.0, N * max_shift)
    t0 = time.perf_counter()
    scipy.optimize.basinhopping(O.logistic_objective, [np.mean(ps), 5.0, 1.0], niter=hops_sample, seed=1,
                                minimizer_kwargs=dict(bounds=((ps.min(), ps.max()), (0.1, 20.0), (0.05, 5.0)),
                                                      tol=1e-6, method="L-BFGS-B", args=(e[:-1], cdf)))
    t_logi = (time.perf_counter() - t0) / hops_sample * 50
    t0 = time.perf_counter()
    scipy.optimize.basinhopping(lambda x: np.sum(w * np.abs(A @ x - b)), np.zeros(N), niter=hops_sample, seed=1,
                                minimizer_kwargs=dict(bounds=[(-bound, bound)] * N, tol=1e-5, method="L-BFGS-B"))
    t_l1 = (time.perf_counter() - t0) / hops_sample * 100 * 2
    t["solve"] = t_logi + t_l1
    # ---- applyShifts (thread pool over frames) + sum
    nthr = min(cores, n_frames)
    pool_frames = [frames[i % frames_sample].copy() for i in range(nthr)]
    kern = O.lanczos_kernel([0.3, 0.7])

    def shift_one(i):
        g = np.roll(np.roll(pool_frames[i], 3, axis=0), -2, axis=1)
        pool_frames[i] = scipy.ndimage.convolve(g, kern, mode="reflect")
    t0 = time.perf_counter()
    with ThreadPool(nthr) as tp:
        tp.map(shift_one, range(nthr))
    t_sh = (time.perf_counter() - t0) / nthr * n_frames
    t0 = time.perf_counter()
    s = np.sum(np.stack(pool_frames[:min(nthr, 4)]), axis=0)
    t_sum = (time.perf_counter() - t0) / min(nthr, 4) * n_frames
    t["shifts"] = t_sh + t_sum
    # ---- FRC on the half sums (whole image)
    even, odd = s, s + np.float32(1.0) * frames[0]
    t0 = time.perf_counter()
    box = 512
    hann = O.hann_apodization([box, box]).astype(np.float32)
    tiles = int(np.floor((H - 10 - box) / box / 0.5))
    frc2 = np.zeros((box, box), np.float32)
    for I in range(tiles):
        for J in range(tiles):
            oy, ox = int(I * box * 0.5 + 5), int(J * box * 0.5 + 5)
            E = sfft.fft2((hann * even[oy:oy + box, ox:ox + box]).astype(np.complex64), workers=cores)
            Od = sfft.fft2((hann * odd[oy:oy + box, ox:ox + box]).astype(np.complex64), workers=cores)
            frc2 += np.real(E * np.conj(Od)) / np.sqrt(np.real(Od * np.conj(Od)) * np.real(E * np.conj(E)))
    O.rotmean(np.fft.fftshift(frc2 / frc2[0, 0]))
    t["frc"] = time.perf_counter() - t0
    # ---- dose filter: per frame FFT + two exp meshes + accumulate; one inverse
    t0 = time.perf_counter()
    xmesh, ymesh = np.meshgrid(np.arange(-W / 2, W / 2, dtype=np.float32), np.arange(-H / 2, H / 2, dtype=np.float32))
    ip = np.float32(1.0 / (W * 1.0))
    q = np.fft.fftshift(np.sqrt(xmesh * xmesh * ip * ip + ymesh * ymesh * ip * ip))
    with np.errstate(divide="ignore"):
        crit = np.float32(0.24499) * q ** np.float32(-1.6649) + np.float32(2.8141)
    acc = np.zeros((H, W), np.complex64)
    t_once = time.perf_counter() - t0
    t0 = time.perf_counter()
    for i, f in enumerate(frames):
        filt = np.exp(np.float32(-0.5 * (i + 1)) / crit)
        thres = np.exp(-(q / np.float32(0.3)) ** np.float32(32))
        acc += sfft.fft2(f.astype(np.complex64), workers=cores) * thres * filt
    t_frame = (time.perf_counter() - t0) / frames_sample
    t0 = time.perf_counter()
    _ = np.abs(sfft.ifft2(acc, workers=cores, norm="forward"))
    t["dose"] = t_once + t_frame * n_frames + (time.perf_counter() - t0)
    total = sum(t.values())
    sample = (f"{frames_sample} frames (K1, dose), {pairs_sample} of {P} pairs, {hops_sample} basin hops per "
              f"optimiser, {nthr} frames shifted, full FRC; stage times scaled to {n_frames}x{H}x{W}")
    return 60.0 / total, dict(stage_s=t, total_s=total, cores=cores, sample=sample)
