"""GPU parity at full size against the oracle's golden fixtures
(tests/golden/make_golden_full.py): the headline movie (40 x 4096^2, fouCrop
3072, 780 pairs, dose filter, 169-tile FRC), a 5760x4092 movie padded to 5760^2
(masked correlation, direct dose path at Hp = 5760) and the super-resolution
counting movie (uint8 7676x7420 with a gain reference, Fourier-binned to
3838x3710, padded to 3840^2, 0.5 A/px).

The fixtures hold the oracle's outputs only; the frames are rebuilt here from
the seed and must reproduce the stored per-frame integer sums first.
Tolerances (the path's correctness bar):
* per-frame translations within 0.02 px of the oracle;
* imageSum and filtSum: relative L2 <= 1e-4 over 65536 seeded random pixels,
  the 256^2 centre crop, the border rows and columns and the 64^2-block means,
  and the whole-image L2 norms to 1e-4;
* FRC curve: max absolute difference <= 1e-3.
"""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden_full as G  # noqa: E402
from zorro_amd.engine import MoviePlan, make_params  # noqa: E402

pytestmark = pytest.mark.gpu

NAMES = sorted(G.CASES, key=lambda k: G.CASES[k]["n"] * G.CASES[k]["h"] * G.CASES[k]["w"])


def _fixture(name):
    path = os.path.join(HERE, "golden", name + ".npz")
    if not os.path.exists(path):
        pytest.skip(f"fixture {name}.npz not generated")
    return np.load(path)


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _summary_errors(prefix, img, fx):
    h, w = img.shape
    got = G.image_summary(prefix, img)
    return {"sample": _rel(img.ravel()[G.sample_index(h, w)], fx[f"{prefix}_sample"]),
            "centre": _rel(got[f"{prefix}_centre"], fx[f"{prefix}_centre"]),
            "edges": _rel(got[f"{prefix}_edges"], fx[f"{prefix}_edges"]),
            "edgecols": _rel(got[f"{prefix}_edgecols"], fx[f"{prefix}_edgecols"]),
            "blocks": _rel(got[f"{prefix}_blocks"], fx[f"{prefix}_blocks"]),
            "l2": abs(float(got[f"{prefix}_l2"]) / float(fx[f"{prefix}_l2"]) - 1.0)}


@pytest.mark.parametrize("name", NAMES)
def test_fullsize_movie_matches_oracle_fixture(name):
    fx = _fixture(name)
    c = G.CASES[name]
    meta = json.loads(str(fx["meta"]))
    assert meta["seed"] == c["seed"] and meta["n"] == c["n"]
    print(f"[gpu fullsize] case={name}: rebuilding {c['n']}x{c['h']}x{c['w']} frames", flush=True)
    frames, _, gain = G.case_frames(c)
    assert np.array_equal(G.frame_sums(frames), fx["frame_sums"]), "regenerated frames differ from the fixture's"
    params = make_params(**c["params"])
    plan = MoviePlan(params, c["n"], c["h"], c["w"])
    try:
        if gain is not None:
            plan.set_frames_raw(frames, gain=gain)
        else:
            plan.set_frames(frames)
        print(f"[gpu fullsize] case={name}: plan {plan.info()}", flush=True)
        plan.align()
        plan.synchronize()
        tm, sol, sums, frc = plan.trimats(), plan.solution(), plan.sums(), plan.frc_curve()
    finally:
        plan.close()
    dt = float(np.max(np.abs(sol["translations"] - fx["translations"])))
    same = np.mean(np.all(np.abs(tm["shiftsTriMat"] - fx["shiftsTriMat"]) < 1e-3, axis=2)[
        np.triu_indices(c["n"], 1)])
    dfrc = float(np.max(np.abs(frc - fx["FRC"])))
    e_img = _summary_errors("imageSum", sums["imageSum"], fx)
    e_flt = _summary_errors("filtSum", sums["filtSum"], fx)
    print(f"[gpu fullsize] {name}: |dt|max {dt:.2e} px; pair shifts identical {same:.4f}; FRC max diff {dfrc:.2e}",
          flush=True)
    print(f"[gpu fullsize] {name}: imageSum {e_img}", flush=True)
    print(f"[gpu fullsize] {name}: filtSum {e_flt}", flush=True)
    assert dt <= 0.02
    assert dfrc <= 1e-3 and frc.shape == fx["FRC"].shape
    for e in (e_img, e_flt):
        assert max(e.values()) <= 1e-4, e
