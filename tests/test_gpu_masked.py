"""GPU parity of the masked (padded-frame) correlation path against the FP64 oracle.

Frames smaller than ``shapePadded`` are zero-padded into the top-left corner and
correlated under the mask of the frame region with the general five-term
``mnxc2`` (``zorro/zorro.py:1014-1097``, ``padStack`` ``:1908-1958``).  The
cases cover the detector shapes of the batched workload (3838x3710 padded to
3840², 5760x4092 padded to 5760²) at a few frames, and small shapes with and
without Fourier cropping.

Tolerances (stated per result):
* lag windows: max |gpu - oracle| <= 2e-4 * max |oracle| (complex64 spectra of
  mean-centred frames, FP32 pair transforms, FP64 five-term combination);
* sub-pixel shifts: identical on the 1/subPixReg grid except near-ties, which
  may move by exactly one grid step (<= 5% of pairs);
* peak significance, correlation and origin values: relative 1e-3.
"""
import numpy as np
import pytest

from oracle import zorro_oracle as O
from zorro_amd.engine import MoviePlan, make_params
from zorro_amd.synth import make_movie

pytestmark = pytest.mark.gpu

CASES = [
    # (name, n, h, w, shapePadded, fouCrop, maxShift, seed)
    ("200x232_pad256", 6, 200, 232, (256, 256), (0, 0), 20, 21),
    ("450x480_pad512_crop256", 5, 450, 480, (512, 512), (256, 256), 20, 22),
    ("1000x900_pad1024", 4, 1000, 900, (1024, 1024), (0, 0), 40, 23),
    ("3838x3710_pad3840_crop3072", 3, 3838, 3710, (3840, 3840), (3072, 3072), 40, 24),
    ("5760x4092_pad5760_crop3072", 2, 5760, 4092, (5760, 5760), (3072, 3072), 40, 25),
]


def _run(name, n, h, w, pad, crop, m, seed):
    print(f"[gpu masked] case={name} n={n} shape={h}x{w} pad={pad} crop={crop} m={m} seed={seed}", flush=True)
    frames, drift = make_movie(n, h, w, seed=seed, dose=1.5)
    kw = dict(triMode="tri", maxShift=m, Bmode="conv", Brad=150, subPixReg=16, fouCrop=crop, shapePadded=pad)
    plan = MoviePlan(make_params(**kw), n, h, w)
    info = plan.info()
    assert info["unmasked"] == 0 and info["Hp"] == pad[0] and info["Wp"] == pad[1]
    plan.set_frames(frames)
    plan.xcorr()
    tm = plan.trimats()
    win = plan.windows()
    plan.close()
    print(f"[gpu masked] case={name} GPU done; running the oracle", flush=True)
    op = O.Params(**kw)
    padded, mask = O.pad_stack(frames.astype(np.float64), pad)
    ref = O.xcorrnm2_tri(padded, mask, op, return_windows=True)
    return tm, win, ref


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_masked_xcorr_parity(case):
    name, n, h, w, pad, crop, m, seed = case
    tm, win, ref = _run(*case)
    pairs = [(I, J) for I in range(n) for J in range(I + 1, n)]
    keep = np.ones((2 * m, 2 * m), bool)
    keep[m, m] = False
    worst = 0.0
    for p, (I, J) in enumerate(pairs):
        o = ref["windows"][(I, J)]
        err = np.max(np.abs(win[p] - o)[keep]) / np.max(np.abs(o))
        worst = max(worst, err)
    print(f"[gpu masked] {name}: worst window rel err {worst:.2e}", flush=True)
    assert worst <= 2e-4
    so, sg = ref["shiftsTriMat"], tm["shiftsTriMat"]
    scale = np.array([pad[0] / (crop[0] or pad[0]), pad[1] / (crop[1] or pad[1])], np.float32)
    grid = scale / 16.0
    iu = np.triu_indices(n, 1)
    steps = np.abs(sg - so)[iu] / grid
    assert np.all(np.abs(steps - np.round(steps)) < 1e-3), "shift differences off the sub-pixel grid"
    assert np.all(np.round(steps) <= 1), f"shift moved by more than one grid step: {np.max(steps)}"
    moved = int(np.sum(np.round(steps) > 0))
    print(f"[gpu masked] {name}: sub-pixel grid moves {moved}/{steps.size}", flush=True)
    assert moved <= max(1, int(0.05 * steps.size))
    for key, tol in (("peaksigTriMat", 1e-3), ("corrTriMat", 1e-3), ("originTriMat", 1e-3)):
        a, b = tm[key][iu].astype(np.float64), ref[key][iu].astype(np.float64)
        rel = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12 + 1e-3 * np.max(np.abs(b))))
        print(f"[gpu masked] {name}: {key} max rel {rel:.2e}", flush=True)
        assert rel <= tol, key
