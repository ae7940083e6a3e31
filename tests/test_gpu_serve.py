"""The movie server on the GPU: every movie's result is the same alone and in
the mix.

Movies of four shape classes (unpadded, padded/masked, Fourier-cropped, mixed
frame counts) are registered one at a time, then all submitted at once with
three in flight, then again under an HBM budget that holds one plan at a time
(admission waits and evicts).  Tolerance: bit-identical translations, imageSum
and filtSum (no kernel's result depends on what else runs on the device).
"""
import numpy as np
import pytest

from zorro_amd.serve import MovieServer, latency_summary
from zorro_amd.synth import make_movie

pytestmark = pytest.mark.gpu

BASE = dict(triMode="tri", maxShift=20, Bmode="conv", Brad=150, subPixReg=16, shiftMethod="lanczos",
            weightMode="autologistic", filterMode="dose", pixelsize=0.1, doLazyFRC=True, frc_box=256)

MIX = [
    (10, 512, 512, dict(fouCrop=(0, 0), shapePadded=(512, 512)), 71),
    (8, 450, 480, dict(fouCrop=(0, 0), shapePadded=(512, 512)), 72),
    (12, 1024, 1024, dict(fouCrop=(512, 512), shapePadded=(1024, 1024)), 73),
    (14, 420, 500, dict(fouCrop=(256, 256), shapePadded=(512, 512)), 74),
    (9, 512, 512, dict(fouCrop=(0, 0), shapePadded=(512, 512)), 75),
]


@pytest.fixture(scope="module")
def movies():
    import torch
    out = []
    for n, h, w, kw, seed in MIX:
        f, _ = make_movie(n, h, w, seed=seed, dose=1.5)
        out.append((torch.from_numpy(f).to("cuda:0"), dict(BASE, **kw)))
    return out


def _run(movies, **server_kw):
    with MovieServer(**server_kw) as srv:
        futs = [srv.submit(f, **kw) for f, kw in movies]
        res = [fu.result(timeout=120) for fu in futs]
        stats = dict(srv.stats)
    return res, stats


def _same(a, b):
    return (np.array_equal(a.translations, b.translations) and bool((a.imageSum == b.imageSum).all())
            and bool((a.filtSum == b.filtSum).all()) and np.array_equal(a.FRC, b.FRC))


def test_alone_equals_mixed(movies):
    print(f"[gpu serve] {len(movies)} movies alone", flush=True)
    alone = []
    for m in movies:
        r, _ = _run([m], max_inflight=1)
        alone.append(r[0])
    print("[gpu serve] all at once, 3 in flight", flush=True)
    mixed, st = _run(movies * 2, max_inflight=3)
    print(f"[gpu serve] mixed stats {st} latency {latency_summary(mixed)}", flush=True)
    for i, r in enumerate(mixed):
        assert _same(r, alone[i % len(movies)]), f"movie {i % len(movies)} differs in the mix"


def test_tight_budget_serialises_and_evicts(movies):
    from zorro_amd.serve import estimate_plan_bytes, make_params
    big = max(estimate_plan_bytes(f.shape[0], f.shape[1], f.shape[2], make_params(**kw)) for f, kw in movies)
    print(f"[gpu serve] budget {big / 1e9:.2f} GB (one plan)", flush=True)
    alone = [_run([m], max_inflight=1)[0][0] for m in movies]
    res, st = _run(movies, max_inflight=3, hbm_budget=int(big * 1.2))
    print(f"[gpu serve] tight-budget stats {st}", flush=True)
    assert st["plans_evicted"] >= 1 and st["peak_held"] <= int(big * 1.2)
    for r, a in zip(res, alone):
        assert _same(r, a)
