"""GPU parity of the filterMode steps after the registration
(zorro/zorro.py:1692-1711) and of cdfLogisticCurve's histogram
(zorro/zorro.py:1988-1998, errorDict pdfPeaks/cdfPeaks/hSigma) against the
FP64 oracle.

Each case registers a movie on the GPU, then the oracle applies the same steps
to the frames aligned with the GPU's translations (the registration itself is
checked in test_gpu_pipeline.py).  Tolerances:
* filtSum after any chain of dose / dosenorm / gausslp: relative L2 <= 1e-4;
* histogram: pdfPeaks and hSigma bit-identical (float32 counts and numpy's
  float32 bin edges), cdfPeaks to 1 float32 ulp.
"""
import numpy as np
import pytest

from oracle import zorro_oracle as O
from zorro_amd.engine import MoviePlan, make_params
from zorro_amd.synth import make_movie

pytestmark = pytest.mark.gpu

CASES = {
    # square unpadded frames: the spectral dose path
    "512sq_8f_dosenorm": dict(n=8, h=512, w=512, seed=31, filt="dosenorm"),
    "512sq_8f_dose_gausslp": dict(n=8, h=512, w=512, seed=32, filt="dose, gausslp"),
    # non-square frames: the direct dose path
    "512x1024_6f_dosenorm_gausslp": dict(n=6, h=512, w=1024, seed=33, filt="dosenorm,gausslp"),
    # padded frames: masked correlation, direct dose path on the symmetric-padded frames
    "450x480_pad512_6f_gausslp_dosenorm": dict(n=6, h=450, w=480, seed=34, filt="gausslp,dosenorm", pad=(512, 512)),
    # no dose step: the Gaussian starts from a copy of imageSum
    "384x320_pad512_6f_gausslp": dict(n=6, h=384, w=320, seed=35, filt="gausslp", pad=(512, 512)),
}


def _kw(c):
    return dict(triMode="tri", maxShift=20, Bmode="conv", Brad=150, subPixReg=16, shiftMethod="lanczos",
                weightMode="autologistic", originMode="centroid", suppressOrigin=True, fouCrop=(0, 0),
                shapePadded=c.get("pad", (c["h"], c["w"])), filterMode=c["filt"], pixelsize=0.1, voltage=300.0,
                doLazyFRC=False)


@pytest.mark.parametrize("name", list(CASES))
def test_filter_chain_matches_oracle(name):
    c = CASES[name]
    print(f"[gpu filters] case={name} {c}", flush=True)
    frames, _ = make_movie(c["n"], c["h"], c["w"], seed=c["seed"], dose=1.5)
    kw = _kw(c)
    plan = MoviePlan(make_params(**kw), c["n"], c["h"], c["w"])
    try:
        plan.set_frames(frames)
        plan.align()
        sol, sums, tm = plan.solution(), plan.sums(), plan.trimats()
        hist = plan.peaksig_hist()
    finally:
        plan.close()
    p = O.Params(**kw)
    p.originMode = None                      # the GPU translations are already centred
    aligned, isum, _ = O.apply_shifts(frames.astype(np.float64), sol["translations"], p)
    filt = None
    for step in O.filter_steps(kw["filterMode"]):
        if step == "gausslp":
            filt = O.gauss_lowpass(isum if filt is None else filt)
        else:
            filt, _ = O.dose_filter(aligned, isum, p, kw["shapePadded"], normalize=(step == "dosenorm"))
    rel = np.linalg.norm(sums["filtSum"] - filt) / np.linalg.norm(filt)
    print(f"[gpu filters] {name}: filtSum rel L2 {rel:.2e}", flush=True)
    assert rel <= 1e-4
    fit = O.cdf_logistic_curve(tm["peaksigTriMat"])
    assert hist is not None
    assert np.array_equal(hist["pdfPeaks"], fit["pdfPeaks"])
    assert np.array_equal(hist["hSigma"], fit["hSigma"].astype(np.float32))
    assert np.allclose(hist["cdfPeaks"], fit["cdfPeaks"], rtol=1.2e-7, atol=0)


def test_histogram_absent_without_autologistic_weights():
    frames, _ = make_movie(6, 256, 256, seed=36, dose=1.5)
    kw = dict(triMode="tri", maxShift=20, Bmode="conv", Brad=150, weightMode="corr", fouCrop=(0, 0),
              shapePadded=(256, 256), doLazyFRC=False)
    plan = MoviePlan(make_params(**kw), 6, 256, 256)
    try:
        plan.set_frames(frames)
        plan.align()
        assert plan.peaksig_hist() is None
    finally:
        plan.close()
