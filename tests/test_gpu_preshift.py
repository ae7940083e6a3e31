"""GPU parity of preShift (zorro/zorro.py:194, :1148-1163): the maxShift disc of
pair (I, J) is rolled by round(shift of (I, J-1)) on the whole correlation map.
The movies drift 2.5 px per frame, so far pairs lie outside a disc centred on
zero lag and only the following disc finds them.  The oracle restates the
reference on full maps (ifftshifted, Gaussian over the whole map, rolled mask);
the GPU correlates windows of maxShift + 8 + Gaussian radius around each disc,
diagonal by diagonal.

Tolerances (as the other correlation tests): integer peaks identical; sub-pixel
shifts equal on the 1/subPixReg grid except near-ties (one grid step, <= 5% of
the shifts); peak significance, correlation and origin values relative 1e-3.
"""
import numpy as np
import pytest

from oracle import zorro_oracle as O
from zorro_amd.engine import MoviePlan, make_params
from zorro_amd.synth import phantom

pytestmark = pytest.mark.gpu


def _drifting_movie(n, h, w, seed, step=(2.5, -1.5), dose=2.0):
    ph = phantom(h, w, seed, dose)
    P = np.fft.rfft2(ph)
    fy = np.fft.fftfreq(h)[:, None]
    fx = np.fft.rfftfreq(w)[None, :]
    rng = np.random.default_rng(seed + 3)
    out = np.empty((n, h, w), np.float32)
    for k in range(n):
        ramp = np.exp(-2j * np.pi * (fy * step[0] * k + fx * step[1] * k))
        out[k] = rng.poisson(np.clip(np.fft.irfft2(P * ramp, s=(h, w)), 0.0, None))
    return out


CASES = [
    # (name, n, size, triMode, diagWidth, maxShift, fouCrop, drift per frame)
    ("512sq_tri_m10", 9, 512, "tri", 5, 10, (0, 0), (2.5, -1.5)),
    ("2048sq_crop1024_diag3_m8", 8, 2048, "diag", 3, 8, (1024, 1024), (6.5, -4.0)),
    ("2048sq_diag2_m10_r2", 6, 2048, "diag", 2, 10, (0, 0), (6.5, -4.0)),
    ("512sq_first_m10", 8, 512, "first", 5, 10, (0, 0), (2.5, -1.5)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_preshift_matches_oracle(case):
    name, n, size, tri, dw, m, crop, step = case
    print(f"[gpu preshift] case={name}", flush=True)
    frames = _drifting_movie(n, size, size, seed=17, step=step)
    kw = dict(triMode=tri, diagWidth=dw, maxShift=m, Bmode="conv", Brad=150, subPixReg=16, fouCrop=crop,
              shapePadded=(size, size), preShift=True)
    plan = MoviePlan(make_params(**kw), n, size, size)
    plan.set_frames(frames)
    plan.xcorr()
    tm = plan.trimats()
    plan.close()
    p = O.Params(**kw)
    pad, mask = O.pad_stack(frames.astype(np.float64), (size, size))
    ref = O.xcorrnm2_tri(pad, mask, p)
    pairs = [(I, J) for I in range(n) for J in range(I + 1, n) if ref["corrTriMat"][I, J] != 0]
    iu = (np.array([q[0] for q in pairs]), np.array([q[1] for q in pairs]))
    hc = crop[0] or size
    grid = (size / hc) / 16.0
    d = np.abs(tm["shiftsTriMat"][iu] - ref["shiftsTriMat"][iu]) / grid
    assert np.all(np.abs(d - np.round(d)) < 1e-3), "shift differences off the sub-pixel grid"
    assert np.all(np.round(d) <= 1), f"shift moved by more than one grid step: {d.max()}"
    moved = int(np.sum(np.round(d) > 0))
    far = max(abs(ref["shiftsTriMat"][iu][:, 0]).max(), abs(ref["shiftsTriMat"][iu][:, 1]).max())
    print(f"[gpu preshift] {name}: {len(pairs)} pairs, largest shift {far:.2f} px, grid moves {moved}", flush=True)
    assert moved <= max(1, 0.05 * d.size)
    assert far > m * size / hc, "the movie must drift beyond the unrolled disc"
    for key in ("peaksigTriMat", "corrTriMat", "originTriMat"):
        a, b = tm[key][iu].astype(np.float64), ref[key][iu].astype(np.float64)
        rel = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12 + 1e-3 * np.max(np.abs(b))))
        print(f"[gpu preshift] {name}: {key} max rel {rel:.2e}", flush=True)
        assert rel <= 1e-3, key


def test_preshift_refusals():
    with pytest.raises(RuntimeError, match="preShift"):
        MoviePlan(make_params(triMode="tri", maxShift=10, Bmode="opti", fouCrop=(0, 0), shapePadded=(512, 512),
                              preShift=True), 4, 512, 512)
    with pytest.raises(RuntimeError, match="preShift"):
        MoviePlan(make_params(triMode="tri", maxShift=10, Bmode="conv", Brad=150, fouCrop=(0, 0),
                              shapePadded=(512, 512), preShift=True), 4, 500, 512)
