"""zr_align captured as one HIP graph: every replay equals the eager run.

The graph is replayed several times with eager work and synchronisations in
between (other movies' stages, torch copies into the frames buffer), and each
replay is compared with an eager run on the same frames.  Tolerance:
bit-identical translations, trimats, imageSum, filtSum and FRC.
"""
import numpy as np
import pytest

from zorro_amd.engine import MoviePlan, make_params
from zorro_amd.synth import make_movie

pytestmark = pytest.mark.gpu

KW = dict(triMode="tri", maxShift=20, Bmode="conv", Brad=150, subPixReg=16, shiftMethod="lanczos",
          weightMode="autologistic", fouCrop=(0, 0), filterMode="dose", pixelsize=0.1, doLazyFRC=True, frc_box=256)

CASES = [("512sq_unmasked", 10, 512, 512, (512, 512)), ("450x480_masked", 8, 450, 480, (512, 512))]


def _results(plan):
    return dict(t=plan.solution()["translations"], tm=plan.trimats(), sums=plan.sums(), frc=plan.frc_curve())


def _same(a, b):
    return (np.array_equal(a["t"], b["t"]) and all(np.array_equal(a["tm"][k], b["tm"][k]) for k in a["tm"])
            and all(np.array_equal(a["sums"][k], b["sums"][k]) for k in a["sums"]) and np.array_equal(a["frc"], b["frc"]))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_graph_replays_equal_eager(case):
    import torch
    name, n, h, w, pad = case
    print(f"[gpu graph] case={name}", flush=True)
    movies = [make_movie(n, h, w, seed=100 + k, dose=1.5)[0] for k in range(3)]
    kw = dict(KW, shapePadded=pad)
    eager = []
    for f in movies:
        p = MoviePlan(make_params(**kw), n, h, w)
        p.set_frames(f)
        p.align()
        eager.append(_results(p))
        p.close()
    buf = torch.empty((n, h, w), dtype=torch.float32, device="cuda:0")
    g = MoviePlan(make_params(**kw), n, h, w)
    g.use_graph(True)
    g.set_frames(buf)
    other = MoviePlan(make_params(**kw), n, h, w)      # eager work between replays
    other.set_frames(movies[2])
    for rep in range(6):
        k = rep % 3
        buf.copy_(torch.from_numpy(movies[k]))
        torch.cuda.synchronize()
        g.align()
        g.synchronize()
        assert _same(_results(g), eager[k]), f"replay {rep} (movie {k}) differs from the eager run"
        other.align()
        other.synchronize()
    print(f"[gpu graph] {name}: 6 replays equal the eager runs", flush=True)
    g.close()
    other.close()
