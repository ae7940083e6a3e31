"""Batches of uneven movies (zr_batch): every movie's results bit-identical to
the same movie registered on its own plan.

The batch runs the spectra, pair transforms and peak searches of all movies
as single launches over a ragged pair list and the movies' solver / shift /
filter / FRC stages on streams of their own; each of those steps computes a
frame, a pair or a movie independently of the others, so the bar is exact
equality of trimats, translations, sums and FRC, eager and as a replayed graph.
"""
import numpy as np
import pytest

from zorro_amd.batch import BatchPlan
from zorro_amd.engine import MoviePlan, make_params
from zorro_amd.synth import make_movie

pytestmark = pytest.mark.gpu

CASES = {
    # two detector shapes, uneven frame counts
    "512sq_unmasked_20_33_47": dict(h=512, w=512, ns=(20, 33, 47), pad=(512, 512), box=256, tri="tri"),
    "450x480_pad512_masked_9_14": dict(h=450, w=480, ns=(9, 14), pad=(512, 512), box=256, tri="tri"),
    "512sq_diag_12_17": dict(h=512, w=512, ns=(12, 17), pad=(512, 512), box=256, tri="diag"),
}


def _kw(c):
    return dict(triMode=c["tri"], diagWidth=4, maxShift=20, Bmode="conv", Brad=150, subPixReg=16,
                shiftMethod="lanczos", weightMode="autologistic", fouCrop=(0, 0), shapePadded=c["pad"],
                filterMode="dose", pixelsize=0.1, voltage=300.0, doLazyFRC=True, frc_box=c["box"])


def _results(plan):
    tm, sol, sums = plan.trimats(), plan.solution(), plan.sums()
    return dict(tm=tm, t=sol["translations"], w=sol["weights"], sums=sums, frc=plan.frc_curve())


def _same(a, b, name):
    for k in ("shiftsTriMat", "corrTriMat", "peaksigTriMat", "originTriMat"):
        assert np.array_equal(a["tm"][k], b["tm"][k]), (name, k)
    assert np.array_equal(a["t"], b["t"]), (name, "translations")
    assert np.array_equal(a["w"], b["w"]), (name, "weights")
    for k in ("imageSum", "filtSum", "evenSum", "oddSum"):
        assert np.array_equal(a["sums"][k], b["sums"][k]), (name, k)
    assert np.array_equal(a["frc"], b["frc"]), (name, "FRC")


@pytest.mark.parametrize("name", list(CASES))
def test_batch_movies_equal_their_own_plans(name):
    c = CASES[name]
    print(f"[gpu batch] case={name}", flush=True)
    movies = [make_movie(n, c["h"], c["w"], seed=70 + k, dose=1.5)[0] for k, n in enumerate(c["ns"])]
    kw = _kw(c)
    alone = []
    for k, f in enumerate(movies):
        plan = MoviePlan(make_params(**kw), f.shape[0], c["h"], c["w"])
        try:
            plan.set_frames(f)
            plan.align()
            alone.append(_results(plan))
        finally:
            plan.close()
    batch = BatchPlan(make_params(**kw), c["ns"], c["h"], c["w"])
    try:
        for k, f in enumerate(movies):
            batch.set_frames(k, f)
        batch.align()
        batch.synchronize()
        for k, m in enumerate(batch.movies):
            _same(_results(m), alone[k], f"{name} eager movie {k}")
        batch.use_graph(True)
        for rep in range(2):
            batch.align()
            batch.synchronize()
            for k, m in enumerate(batch.movies):
                _same(_results(m), alone[k], f"{name} graph replay {rep} movie {k}")
        print(f"[gpu batch] {name}: {len(movies)} movies identical to their own plans "
              f"({batch.device_bytes() / 1e6:.0f} MB)", flush=True)
    finally:
        batch.close()


def test_batch_takes_device_frames_written_on_another_stream():
    """Frames produced by torch kernels that may still be running on torch's
    stream when they are handed to the batch (the copy must wait for them)."""
    import torch
    c = CASES["512sq_unmasked_20_33_47"]
    ns = (20, 33)
    movies = [make_movie(n, c["h"], c["w"], seed=80 + k, dose=1.5)[0] for k, n in enumerate(ns)]
    kw = _kw(c)
    batch = BatchPlan(make_params(**kw), ns, c["h"], c["w"])
    ref = BatchPlan(make_params(**kw), ns, c["h"], c["w"])
    try:
        for k, f in enumerate(movies):
            ref.set_frames(k, f)
            t = torch.from_numpy(f).to("cuda:0")
            for _ in range(20):          # queue work on torch's stream ahead of the hand-over
                t = t * 1.0
            batch.set_frames(k, t)
            del t
        for b in (batch, ref):
            b.align()
            b.synchronize()
        for k in range(len(ns)):
            _same(_results(batch.movies[k]), _results(ref.movies[k]), f"device frames movie {k}")
    finally:
        batch.close()
        ref.close()


def test_batch_refuses_unbuilt_settings():
    from zorro_amd import _lib as L
    with pytest.raises(L.ZrError):
        BatchPlan(make_params(triMode="auto", maxShift=20, fouCrop=(0, 0), shapePadded=(256, 256)), [8, 9], 256, 256)
    with pytest.raises(L.ZrError):
        BatchPlan(make_params(triMode="tri", maxShift=20, fouCrop=(0, 0), shapePadded=(256, 256)), [8, 65], 256, 256)
