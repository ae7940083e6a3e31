"""GPU parity of the standalone tiledFRC (zr_tiled_frc; zorro/zorro.py:2246-2310)
against the FP64 oracle, and against the movie's own FRC stage.

Tolerances: the whole rotational-mean curve within 1e-3 (max absolute
difference) of the oracle; the first box/2 values within 1e-6 of the movie
plan's lazy FRC of the same half sums (the same kernels; only the twiddle
tables are computed on the device here and on the host there).
"""
import numpy as np
import pytest

from oracle import zorro_oracle as O
from zorro_amd import _lib as L
from zorro_amd.engine import MoviePlan, make_params, tiled_frc
from zorro_amd.synth import make_movie

pytestmark = pytest.mark.gpu

CASES = [
    # (name, h, w, box, overlap, translations)
    ("1100x1300_box256_trans", 1100, 1300, 256, 0.5, np.array([[-3.2, 7.5], [2.0, -0.4], [0.5, 1.0]])),
    ("1536sq_box512_none", 1536, 1536, 512, 0.5, None),
    ("900x2100_box256_overlap0.25", 900, 2100, 256, 0.25, np.array([[1.0, -6.6], [-2.0, 3.0]])),
    ("2048sq_box1024_overlap0.9", 2048, 2048, 1024, 0.9, None),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_tiled_frc_matches_oracle(case):
    name, h, w, box, ov, trans = case
    print(f"[gpu tiled frc] case={name}", flush=True)
    rng = np.random.default_rng(h + w + box)
    common = rng.standard_normal((h, w))
    a = (3.0 + common + 0.7 * rng.standard_normal((h, w))).astype(np.float32)
    b = (3.0 + common + 0.7 * rng.standard_normal((h, w))).astype(np.float32)
    got = tiled_frc(a, b, trans, box, ov)
    ref, _, nt = O.tiled_frc(a.astype(np.float64), b.astype(np.float64),
                             np.zeros((1, 2)) if trans is None else trans, box, ov)
    assert got.shape == ref.shape
    d = float(np.max(np.abs(got - ref)))
    print(f"[gpu tiled frc] {name}: {nt} tiles, {got.size} rings, max abs diff {d:.2e}", flush=True)
    assert d <= 1e-3


def test_tiled_frc_equals_movie_lazy_frc():
    n, h, w = 8, 1024, 1024
    frames, _ = make_movie(n, h, w, seed=21, dose=1.5)
    kw = dict(triMode="tri", maxShift=20, Bmode="conv", Brad=150, fouCrop=(0, 0), shapePadded=(h, w),
              doLazyFRC=True, frc_box=256)
    plan = MoviePlan(make_params(**kw), n, h, w)
    try:
        plan.set_frames(frames)
        plan.align()
        sums, frc, t = plan.sums(), plan.frc_curve(), plan.solution()["translations"]
    finally:
        plan.close()
    full = tiled_frc(sums["evenSum"], sums["oddSum"], t, 256, 0.5)
    assert np.max(np.abs(full[:128] - frc)) <= 1e-6


def test_tiled_frc_refuses_images_without_a_tile():
    a = np.zeros((300, 300), np.float32)
    with pytest.raises(L.ZrError):
        tiled_frc(a, a, None, 256, 0.5)
