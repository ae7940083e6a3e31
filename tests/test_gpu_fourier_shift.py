"""GPU parity of shiftMethod='fourier' (the shift theorem on the zero-padded
frames' spectra; zorro/zorro.py:1748-1772 as intended) against the FP64
oracle, given the GPU's translations.

Tolerances: aligned frames, imageSum and filtSum relative L2 <= 1e-4; FRC max
absolute difference <= 1e-3; end-to-end translations within 0.02 px of the
oracle (the registration does not depend on the shift method).
"""
import numpy as np
import pytest

from oracle import zorro_oracle as O
from zorro_amd.engine import MoviePlan, make_params
from zorro_amd.synth import make_movie

pytestmark = pytest.mark.gpu

CASES = {
    # C1 as configured: 20 x 1024^2, no crop, Fourier shifts, dose filter, lazy FRC
    "C1_1024sq_20f": dict(n=20, h=1024, w=1024, seed=1, box=512, filt="dose"),
    "512sq_8f_dosenorm": dict(n=8, h=512, w=512, seed=51, box=256, filt="dosenorm"),
    # padded frames: the masked path's centred spectra plus mean x mask spectrum
    "450x480_pad512_8f": dict(n=8, h=450, w=480, seed=52, box=256, filt="dose", pad=(512, 512)),
    "1000x900_pad1024_crop512_8f_nodose": dict(n=8, h=1000, w=900, seed=53, box=512, filt=None, pad=(1024, 1024),
                                               crop=(512, 512)),
}


def _kw(c):
    return dict(triMode="tri", maxShift=40, Bmode="conv", Brad=150, subPixReg=16, shiftMethod="fourier",
                weightMode="autologistic", originMode="centroid", suppressOrigin=True, fouCrop=c.get("crop", (0, 0)),
                shapePadded=c.get("pad", (c["h"], c["w"])), filterMode=c["filt"], pixelsize=0.1, voltage=300.0,
                doLazyFRC=True, frc_box=c["box"])


@pytest.mark.parametrize("name", list(CASES))
def test_fourier_shifts_match_oracle(name):
    import torch
    c = CASES[name]
    print(f"[gpu fourier] case={name}", flush=True)
    frames, drift = make_movie(c["n"], c["h"], c["w"], seed=c["seed"], dose=1.5)
    kw = _kw(c)
    plan = MoviePlan(make_params(**kw), c["n"], c["h"], c["w"])
    try:
        plan.set_frames(frames)
        plan.align()
        sol, sums, frc = plan.solution(), plan.sums(), plan.frc_curve()
        out = torch.empty((c["n"], c["h"], c["w"]), dtype=torch.float32, device="cuda:0")
        plan.apply_shifts(out)
        plan.synchronize()
        aligned_gpu = out.cpu().numpy()
    finally:
        plan.close()
    p = O.Params(**kw)
    p.originMode = None                       # the GPU translations are already centred
    t = sol["translations"]
    aligned, isum, _ = O.apply_shifts(frames.astype(np.float64), t, p)
    rel_a = np.linalg.norm(aligned_gpu - aligned) / np.linalg.norm(aligned)
    rel = np.linalg.norm(sums["imageSum"] - isum) / np.linalg.norm(isum)
    ofrc, _ = O.lazy_frc(aligned, t, c["box"])
    dfrc = float(np.max(np.abs(frc - ofrc)))
    print(f"[gpu fourier] {name}: aligned rel L2 {rel_a:.2e}, imageSum {rel:.2e}, FRC {dfrc:.2e}", flush=True)
    assert rel_a <= 1e-4 and rel <= 1e-4 and dfrc <= 1e-3
    if c["filt"]:
        fs, _ = O.dose_filter(aligned, isum, p, kw["shapePadded"], normalize=(c["filt"] == "dosenorm"))
        relf = np.linalg.norm(sums["filtSum"] - fs) / np.linalg.norm(fs)
        print(f"[gpu fourier] {name}: filtSum rel L2 {relf:.2e}", flush=True)
        assert relf <= 1e-4
    q = O.Params(**kw)
    q.filterMode, q.doLazyFRC = None, False
    ref = O.align_image_stack(frames, q)
    d = float(np.max(np.abs(sol["translations"] - ref["translations"])))
    print(f"[gpu fourier] {name}: end-to-end |dt|max {d:.4f} px", flush=True)
    assert d <= 0.02
