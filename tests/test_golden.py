"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU tests pin the seeded generator and the oracle against the stored vectors;
GPU tests run the HIP path on the stored frames and compare with the stored
answers.

Tolerances (stated per result):
* generator and oracle regression: frames bit-identical, trimats and
  translations equal to 1e-9 (FP64 oracle re-run), image summaries relative 1e-9;
* GPU: translations within 0.02 px; sub-pixel pair shifts identical on the
  1/16 grid up to one step at near-ties; imageSum and filtSum relative L2 of the
  8x8 block means and of the 64x64 centre crop <= 1e-4, L2 norms within 1e-4;
  FRC max abs difference <= 1e-3.
"""
import glob
import json
import os

import numpy as np
import pytest

from oracle import zorro_oracle as O
from zorro_amd.synth import make_movie

HERE = os.path.dirname(os.path.abspath(__file__))
# the full-size fixtures (full_*.npz, outputs only) are checked by test_gpu_fullsize.py
FILES = sorted(f for f in glob.glob(os.path.join(HERE, "golden", "*.npz"))
               if not os.path.basename(f).startswith("full_"))
IDS = [os.path.basename(f)[:-4] for f in FILES]


def load(path):
    z = np.load(path)            # no pickles: allow_pickle stays False
    d = {k: z[k] for k in z.files}
    d["meta"] = json.loads(str(d["meta"]))
    return d


def test_fixtures_present():
    assert len(FILES) >= 3


@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_generator_reproduces_fixture_frames(path):
    g = load(path)
    m = g["meta"]
    frames, drift = make_movie(m["n"], m["h"], m["w"], seed=m["seed"], dose=m["dose"])
    assert np.array_equal(frames, g["frames_u8"].astype(np.float32))
    assert np.allclose(drift, g["drift"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_oracle_reproduces_fixture(path):
    g = load(path)
    p = O.Params(**g["meta"]["params"])
    res = O.align_image_stack(g["frames_u8"].astype(np.float64), p)
    xc = res["xcorr"]
    for k in ("shiftsTriMat", "corrTriMat", "peaksigTriMat", "originTriMat"):
        assert np.allclose(xc[k], g[k], rtol=1e-6, atol=1e-9), k
    assert np.allclose(res["translations"], g["translations"], rtol=0, atol=1e-9)
    assert np.allclose(res["FRC"], g["FRC"], rtol=0, atol=1e-9)
    for key in ("imageSum", "filtSum"):
        img = np.asarray(res[key], np.float64)
        assert np.isclose(np.linalg.norm(img), g[key + "_l2"], rtol=1e-9)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_gpu_matches_fixture(path):
    import sys
    from zorro_amd.engine import MoviePlan, make_params
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_golden import block_means, centre_crop

    g = load(path)
    m = g["meta"]
    print(f"[golden] case={m['name']} n={m['n']} {m['h']}x{m['w']} params={m['params']}", flush=True)
    kw = dict(m["params"])
    plan = MoviePlan(make_params(**kw), m["n"], m["h"], m["w"])
    plan.set_frames(g["frames_u8"].astype(np.float32))
    plan.align()
    tm, sol, sums, frc = plan.trimats(), plan.solution(), plan.sums(), plan.frc_curve()
    plan.close()
    n = m["n"]
    iu = np.triu_indices(n, 1)
    pad, crop = kw["shapePadded"], kw["fouCrop"]
    grid = np.array([pad[0] / (crop[0] or pad[0]), pad[1] / (crop[1] or pad[1])]) / 16.0
    steps = np.abs(tm["shiftsTriMat"] - g["shiftsTriMat"])[iu] / grid
    assert np.all(np.abs(steps - np.round(steps)) < 1e-3) and np.all(np.round(steps) <= 1)
    assert int(np.sum(np.round(steps) > 0)) <= max(1, int(0.05 * steps.size))
    dt = float(np.max(np.abs(sol["translations"] - g["translations"])))
    print(f"[golden] {m['name']}: |dt|max {dt:.4f} px", flush=True)
    assert dt <= 0.02
    for key in ("imageSum", "filtSum"):
        img = sums[key]
        rb = _rel(block_means(img.astype(np.float64)), g[key + "_blocks"])
        rc = _rel(centre_crop(img.astype(np.float64)), g[key + "_centre"])
        rn = abs(np.linalg.norm(img.astype(np.float64)) - g[key + "_l2"]) / g[key + "_l2"]
        print(f"[golden] {m['name']}: {key} blocks {rb:.2e} centre {rc:.2e} norm {rn:.2e}", flush=True)
        assert rb <= 1e-4 and rc <= 1e-4 and rn <= 1e-4, key
    dfrc = float(np.max(np.abs(frc - g["FRC"])))
    print(f"[golden] {m['name']}: FRC max abs diff {dfrc:.2e}", flush=True)
    assert dfrc <= 1e-3
