"""Launch tuning knobs (k1_batch, k2_cols_per_wave, k2_pairs_per_wave) change
how the work is cut into launches and workgroups, never what a frame, pair or
column computes: every output must be bit-identical to the default launch.

The pair transform's radix-64 path (k2_cols_r2) runs for Hc = 64 * {32, 48,
64}; the cases below take it at Hc = 2048 (unmasked and masked) and 3072.
"""
import numpy as np
import pytest

from zorro_amd.engine import MoviePlan, make_params
from zorro_amd.synth import make_movie

pytestmark = pytest.mark.gpu

CASES = {
    "2048sq_10f_crop2048": dict(n=10, h=2048, w=2048, pad=(2048, 2048), crop=(2048, 2048), m=20),
    "1900x2000_pad2048_9f_masked": dict(n=9, h=1900, w=2000, pad=(2048, 2048), crop=(2048, 2048), m=24),
    "3072sq_8f_crop3072": dict(n=8, h=3072, w=3072, pad=(3072, 3072), crop=(3072, 3072), m=40),
}
KNOBS = [dict(k1_batch=1), dict(k1_batch=3), dict(k2_cols_per_wave=3), dict(k2_pairs_per_wave=1),
         dict(k2_pairs_per_wave=3), dict(k2_pairs_per_wave=8)]


def _settings(c):
    return dict(triMode="tri", maxShift=c["m"], Bmode="conv", Brad=150, subPixReg=16, shiftMethod="lanczos",
                weightMode="autologistic", fouCrop=c["crop"], shapePadded=c["pad"], filterMode="dose",
                pixelsize=0.1, voltage=300.0, doLazyFRC=True, frc_box=512)


def _run(frames, settings):
    n, h, w = frames.shape
    plan = MoviePlan(make_params(**settings), n, h, w)
    try:
        plan.set_frames(frames)
        plan.align()
        plan.synchronize()
        tm, sol, sums = plan.trimats(), plan.solution(), plan.sums()
        return dict(tm=tm, t=sol["translations"], sums=sums, frc=plan.frc_curve())
    finally:
        plan.close()


@pytest.mark.parametrize("name", list(CASES))
def test_knobs_do_not_change_results(name):
    c = CASES[name]
    print(f"[gpu knobs] case={name}", flush=True)
    frames = make_movie(c["n"], c["h"], c["w"], seed=41, dose=1.5)[0]
    base = _run(frames, _settings(c))
    for knob in KNOBS:
        print(f"[gpu knobs] {name} {knob}", flush=True)
        got = _run(frames, {**_settings(c), **knob})
        for k in ("shiftsTriMat", "corrTriMat", "peaksigTriMat", "originTriMat"):
            assert np.array_equal(got["tm"][k], base["tm"][k]), (knob, k)
        assert np.array_equal(got["t"], base["t"]), (knob, "translations")
        for k in ("imageSum", "filtSum"):
            assert np.array_equal(got["sums"][k], base["sums"][k]), (knob, k)
        assert np.array_equal(got["frc"], base["frc"]), (knob, "FRC")
