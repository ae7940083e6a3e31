"""GPU parity of the pair-correlation path (K1 spectra, K2 windows, peak stage)
against the FP64 oracle on the same seeded movies.

Tolerances (stated per result):
* lag windows: max |gpu - oracle| <= 5e-5 * max |oracle| (complex64 FFTs, FP32 accumulation);
* integer peak positions: identical for every pair;
* sub-pixel shifts: identical on the 1/subPixReg grid except near-ties of the
  interpolated maximum, which may move by exactly one grid step (counted, <= 5%);
* peak significance and correlation values: relative 1e-3.
"""
import sys

import numpy as np
import pytest

from oracle import zorro_oracle as O
from zorro_amd.engine import MoviePlan, make_params
from zorro_amd.synth import make_movie

pytestmark = pytest.mark.gpu

CASES = [
    # (name, n, h, w, fouCrop, maxShift, Brad, seed)
    ("512sq_nocrop", 8, 512, 512, (0, 0), 40, 150, 11),
    ("1024sq_C1", 20, 1024, 1024, (0, 0), 40, 150, 1),
    ("2048sq_crop1024", 6, 2048, 2048, (1024, 1024), 40, 150, 5),
    ("4096sq_crop3072", 4, 4096, 4096, (3072, 3072), 40, 150, 2),
    # column pass with two outputs per residue at its limits: L = 64 (uncropped
    # 4096), L = 32 with maxShift = L, L = 48 with maxShift = L
    ("4096sq_nocrop_L64", 3, 4096, 4096, (0, 0), 40, 150, 7),
    ("2048sq_nocrop_m32", 8, 2048, 2048, (0, 0), 32, 150, 9),
    ("4096sq_crop3072_m48", 3, 4096, 4096, (3072, 3072), 48, 150, 4),
    # largest built window: 128x128 lags with a 51-tap Gaussian (LDS-bound peak stage)
    ("4096sq_crop3072_m64", 3, 4096, 4096, (3072, 3072), 64, 150, 3),
]


def _run(name, n, h, w, crop, m, brad, seed):
    print(f"[gpu xcorr] case={name} n={n} shape={h}x{w} crop={crop} m={m} seed={seed}", flush=True)
    frames, drift = make_movie(n, h, w, seed=seed, dose=1.5)
    kw = dict(triMode="tri", maxShift=m, Bmode="conv", Brad=brad, subPixReg=16, fouCrop=crop,
              shapePadded=(h, w))
    plan = MoviePlan(make_params(**kw), n, h, w)
    plan.set_frames(frames)
    plan.xcorr()
    tm = plan.trimats()
    win = plan.windows()
    plan.close()
    op = O.Params(**kw)
    padded, mask = O.pad_stack(frames.astype(np.float64), (h, w))
    ref = O.xcorrnm2_tri(padded, mask, op, return_windows=True)
    return frames, tm, win, ref


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_xcorr_parity(case):
    name, n, h, w, crop, m, brad, seed = case
    frames, tm, win, ref = _run(*case)
    pairs = [(I, J) for I in range(n) for J in range(I + 1, n)]
    # windows
    # the stored GPU window is the raw one; the origin replacement happens in the
    # peak stage (its input value is checked through originTriMat below)
    worst = 0.0
    keep = np.ones((2 * m, 2 * m), bool)
    keep[m, m] = False
    for p, (I, J) in enumerate(pairs):
        o = ref["windows"][(I, J)]
        err = np.max(np.abs(win[p] - o)[keep]) / np.max(np.abs(o))
        worst = max(worst, err)
    print(f"[gpu xcorr] {name}: worst window rel err {worst:.2e}", flush=True)
    assert worst <= 5e-5
    # integer peak: the sub-pixel shift rounds to the same integer part of the window index
    so = ref["shiftsTriMat"]
    sg = tm["shiftsTriMat"]
    scale = np.array([h / (crop[0] or h), w / (crop[1] or w)], np.float32)
    grid = scale / 16.0
    d = np.abs(sg - so)
    iu = np.triu_indices(n, 1)
    steps = d[iu] / grid
    assert np.all(np.abs(steps - np.round(steps)) < 1e-3), "shift differences off the sub-pixel grid"
    moved = np.sum(np.round(steps) > 0)
    assert np.all(np.round(steps) <= 1), f"shift moved by more than one grid step: {np.max(steps)}"
    frac = moved / steps.size
    print(f"[gpu xcorr] {name}: sub-pixel grid moves {moved}/{steps.size}", flush=True)
    assert frac <= 0.05
    # peaksig / corr / origin
    for key, tol in (("peaksigTriMat", 1e-3), ("corrTriMat", 1e-3), ("originTriMat", 1e-3)):
        a, b = tm[key][iu].astype(np.float64), ref[key][iu].astype(np.float64)
        rel = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12 + 1e-3 * np.max(np.abs(b))))
        print(f"[gpu xcorr] {name}: {key} max rel {rel:.2e}", flush=True)
        assert rel <= tol, key


def test_peak_lds_limit_rejected():
    """A Gaussian too wide for the peak stage's LDS at maxShift 64 is refused at plan time."""
    kw = dict(triMode="tri", maxShift=64, Bmode="conv", Brad=20, subPixReg=16, fouCrop=(3072, 3072),
              shapePadded=(4096, 4096))
    with pytest.raises(RuntimeError, match="LDS"):
        MoviePlan(make_params(**kw), 3, 4096, 4096)
