"""GPU parity of the whole alignImageStack path (xcorr -> FP64 solver -> Lanczos
shifts + sums -> even/odd FRC -> dose filter) against the FP64 oracle.

Tolerances (stated per result, from the path's correctness bar):
* solver on identical trimats: translations equal to 1e-9 px (both take the
  exact optimal vertex), weights relative 1e-6;
* end-to-end per-frame translations: within 0.02 px of the oracle's own run;
* imageSum and filtSum: relative L2 error <= 1e-4;
* FRC curve: max absolute difference <= 1e-3.
"""
import numpy as np
import pytest

from oracle import zorro_oracle as O
from zorro_amd.engine import MoviePlan, make_params
from zorro_amd.synth import make_movie

pytestmark = pytest.mark.gpu

CASES = {
    "512sq_12f": dict(n=12, h=512, w=512, seed=3, dose=1.5, box=256),
    "C1_1024sq_20f": dict(n=20, h=1024, w=1024, seed=1, dose=1.5, box=512),
    "1024sq_10f_nodose": dict(n=10, h=1024, w=1024, seed=9, dose=1.5, box=512, filter=None),
    # non-square frames take the direct dose path (per-frame row FFTs of the aligned frames)
    "512x1024_8f_direct_dose": dict(n=8, h=512, w=1024, seed=13, dose=1.5, box=256),
    # padded frames: masked correlation, crop back, symmetric-padded dose filter
    "450x480_pad512_8f_masked": dict(n=8, h=450, w=480, seed=17, dose=1.5, box=256, pad=(512, 512)),
    "1000x900_pad1024_crop512_10f_masked": dict(n=10, h=1000, w=900, seed=19, dose=1.5, box=512, pad=(1024, 1024),
                                                crop=(512, 512)),
}


def _kw(c):
    return dict(triMode="tri", maxShift=40, Bmode="conv", Brad=150, subPixReg=16, shiftMethod="lanczos",
                weightMode="autologistic", originMode="centroid", suppressOrigin=True, fouCrop=c.get("crop", (0, 0)),
                shapePadded=c.get("pad", (c["h"], c["w"])), filterMode=c.get("filter", "dose"), pixelsize=0.1,
                voltage=300.0, doLazyFRC=True, frc_box=c["box"])


@pytest.fixture(scope="module", params=list(CASES))
def run(request):
    name = request.param
    c = CASES[name]
    print(f"[gpu pipeline] case={name} {c}", flush=True)
    frames, drift = make_movie(c["n"], c["h"], c["w"], seed=c["seed"], dose=c["dose"])
    kw = _kw(c)
    plan = MoviePlan(make_params(**kw), c["n"], c["h"], c["w"])
    plan.set_frames(frames)
    plan.align()
    out = dict(name=name, c=c, frames=frames, drift=drift, kw=kw, tm=plan.trimats(), sol=plan.solution(),
               sums=plan.sums(), frc=plan.frc_curve(), solver=plan.solver_info())
    print(f"[gpu pipeline] case={name} solver {out['solver']}", flush=True)
    plan.close()
    print(f"[gpu pipeline] case={name} GPU done", flush=True)
    return out


def test_solver_matches_oracle_on_same_trimats(run):
    p = O.Params(**run["kw"])
    tm = run["tm"]
    ed = O.shifts_solver(tm["shiftsTriMat"], tm["corrTriMat"], tm["peaksigTriMat"], p,
                         tuple(run["kw"]["shapePadded"]))
    t_or = ed["translations"] - ed["translations"].mean(axis=0)
    t_gpu = run["sol"]["translations"]
    d = np.max(np.abs(t_gpu - t_or))
    assert run["solver"]["status_x"] == 1 and run["solver"]["status_y"] == 1, run["solver"]
    lg = run["sol"]["logistic"]
    print(f"[gpu pipeline] {run['name']}: solver |dt|max={d:.3e} logistic gpu={lg} "
          f"oracle={(ed['peaksigThres'], ed['logisticK'], ed['logisticNu'])}", flush=True)
    assert np.allclose(run["sol"]["weights"], ed["weights"], rtol=1e-6, atol=1e-9)
    assert d <= 1e-9
    assert np.isclose(run["sol"]["mean_errorNorm"], ed["mean_errorNorm"], rtol=1e-6)
    assert np.allclose(run["sol"]["errorX"], ed["errorX"], rtol=1e-6, atol=1e-9)
    assert np.allclose(run["sol"]["errorY"], ed["errorY"], rtol=1e-6, atol=1e-9)


def test_sums_frc_dose_match_oracle_given_translations(run):
    p = O.Params(**run["kw"])
    p.originMode = None                       # GPU translations are already centred
    t = run["sol"]["translations"]
    aligned, isum, _ = O.apply_shifts(run["frames"].astype(np.float64), t, p)
    g = run["sums"]
    rel = np.linalg.norm(g["imageSum"] - isum) / np.linalg.norm(isum)
    print(f"[gpu pipeline] {run['name']}: imageSum rel L2 {rel:.2e}", flush=True)
    assert rel <= 1e-4
    assert np.linalg.norm(g["evenSum"] - aligned[0::2].sum(0)) / np.linalg.norm(aligned[0::2].sum(0)) <= 1e-4
    frc, _ = O.lazy_frc(aligned, t, run["c"]["box"])
    dfrc = np.max(np.abs(run["frc"] - frc))
    print(f"[gpu pipeline] {run['name']}: FRC max abs diff {dfrc:.2e}", flush=True)
    assert run["frc"].shape == frc.shape and dfrc <= 1e-3
    if run["kw"]["filterMode"] != "dose":
        return
    fs, info = O.dose_filter(aligned, isum, p, tuple(run["kw"]["shapePadded"]))
    rel = np.linalg.norm(g["filtSum"] - fs) / np.linalg.norm(fs)
    print(f"[gpu pipeline] {run['name']}: filtSum rel L2 {rel:.2e}", flush=True)
    assert rel <= 1e-4


def test_end_to_end_translations_within_002px(run):
    p = O.Params(**run["kw"])
    p.filterMode = None
    p.doLazyFRC = False
    ref = O.align_image_stack(run["frames"], p)
    d = np.max(np.abs(run["sol"]["translations"] - ref["translations"]))
    truth = -(run["drift"] - run["drift"].mean(axis=0))
    rms_truth = np.sqrt(np.mean((run["sol"]["translations"] - truth) ** 2))
    print(f"[gpu pipeline] {run['name']}: end-to-end |dt|max vs oracle {d:.4f} px; rms vs true drift "
          f"{rms_truth:.3f} px", flush=True)
    assert d <= 0.02
