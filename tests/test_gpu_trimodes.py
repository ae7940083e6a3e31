"""GPU parity of the pair-list modes triMode 'first' and 'auto' against the oracle.

* 'first' (zorro/zorro.py:562-565, :1626-1635): frame 0 against every other
  frame, translations = -shiftsTriMat[0, :] with no solver;
* 'auto' (:992-996): each row stops after the first pair with peaksig <=
  peaksigThres or J - I >= autoMax; the solver's equations follow the surviving
  pairs through the reference's rot90 construction, whose rows can differ from
  the pairs they are matched with (:1296-1311).

Tolerances (stated per result): sub-pixel pair shifts identical on the 1/16 grid
up to one step at near-ties; translations within 0.02 px of the oracle end to
end; the solver on the GPU's own trimats equal to the oracle's exact vertex
within 1e-9 px.
"""
import numpy as np
import pytest

from oracle import zorro_oracle as O
from zorro_amd.engine import MoviePlan, make_params
from zorro_amd.synth import make_movie

pytestmark = pytest.mark.gpu


def _kw(mode, size, **extra):
    kw = dict(triMode=mode, maxShift=20, Bmode="conv", Brad=150, subPixReg=16, shiftMethod="lanczos",
              weightMode="logistic", peaksigThres=3.115, logisticK=5.0, logisticNu=0.15, originMode="centroid",
              suppressOrigin=True, fouCrop=(0, 0), shapePadded=(size, size), filterMode=None, doLazyFRC=False)
    kw.update(extra)
    return kw


def _gpu(frames, kw):
    n, h, w = frames.shape
    plan = MoviePlan(make_params(**kw), n, h, w)
    plan.set_frames(frames)
    plan.xcorr()
    plan.solve()
    out = dict(tm=plan.trimats(), sol=plan.solution(), m=plan.equation_count())
    plan.close()
    return out


def test_first_mode_matches_oracle():
    n, size = 8, 512
    print(f"[gpu trimodes] first n={n} {size}^2", flush=True)
    frames, _ = make_movie(n, size, size, seed=61, dose=1.5)
    kw = _kw("first", size)
    g = _gpu(frames, kw)
    ref = O.align_image_stack(frames, O.Params(**dict(kw, filterMode=None, doLazyFRC=False)))
    so, sg = ref["xcorr"]["shiftsTriMat"], g["tm"]["shiftsTriMat"]
    assert np.all(g["tm"]["corrTriMat"][1:] == 0), "only row 0 is correlated"
    steps = np.abs(sg[0, 1:] - so[0, 1:]) / (1 / 16)
    assert np.all(np.abs(steps - np.round(steps)) < 1e-3) and np.all(np.round(steps) <= 1)
    assert g["m"] == 0
    d = np.max(np.abs(g["sol"]["translations"] - ref["translations"]))
    print(f"[gpu trimodes] first: |dt|max {d:.4f} px", flush=True)
    assert d <= 0.02


def test_auto_mode_pruning_and_solver():
    n, size = 8, 256
    print(f"[gpu trimodes] auto n={n} {size}^2", flush=True)
    frames, _ = make_movie(n, size, size, seed=5, dose=4.0)
    kw = _kw("auto", size, autoMax=5)
    g = _gpu(frames, kw)
    tm = g["tm"]
    ps = tm["peaksigTriMat"]
    lengths = []
    for I in range(n - 1):
        row = [J for J in range(I + 1, n) if tm["corrTriMat"][I, J] != 0]
        assert row == list(range(I + 1, I + 1 + len(row))), (I, row)
        last = row[-1]
        assert ps[I, last] <= kw["peaksigThres"] or last - I >= kw["autoMax"] or last == n - 1
        assert all(ps[I, J] > kw["peaksigThres"] and J - I < kw["autoMax"] for J in row[:-1])
        lengths.append(len(row))
    print(f"[gpu trimodes] auto: row lengths {lengths}, equations {g['m']}", flush=True)
    assert g["m"] == sum(lengths)
    # solver on the GPU's trimats: the oracle's exact vertex of the same equations
    p = O.Params(**kw)
    ed = O.shifts_solver(tm["shiftsTriMat"], tm["corrTriMat"], tm["peaksigTriMat"], p, (size, size))
    t_or = ed["translations"] - ed["translations"].mean(axis=0)
    d = np.max(np.abs(g["sol"]["translations"] - t_or))
    print(f"[gpu trimodes] auto: solver |dt|max {d:.2e} px", flush=True)
    assert d <= 1e-9
    assert np.allclose(g["sol"]["weights"], ed["weights"], rtol=1e-6, atol=1e-9)
    # end to end against the oracle when both stopped every row at the same pair
    ref = O.align_image_stack(frames, O.Params(**kw))
    same = np.array_equal(ref["xcorr"]["corrTriMat"] != 0, tm["corrTriMat"] != 0)
    print(f"[gpu trimodes] auto: same pair pattern as the oracle: {same}", flush=True)
    if same:
        assert np.max(np.abs(g["sol"]["translations"] - ref["translations"])) <= 0.02


@pytest.mark.parametrize("masked", [False, True], ids=["unmasked", "masked"])
def test_refine_mode_matches_oracle(masked):
    """triMode 'refine' (zorro/zorro.py:934-956, :1636-1647): every frame but the
    last against the sum of the others; translations = -shiftsTriMat[K, K].
    Windows within 2e-4 of the oracle's (relative to their maximum)."""
    n, size = 8, 256
    h, w = (230, 200) if masked else (size, size)
    print(f"[gpu trimodes] refine masked={masked} n={n} {h}x{w} pad {size}", flush=True)
    frames, _ = make_movie(n, h, w, seed=62, dose=3.0)
    kw = _kw("refine", size)
    plan = MoviePlan(make_params(**kw), n, h, w)
    plan.set_frames(frames)
    plan.xcorr()
    plan.solve()
    tm, sol, win = plan.trimats(), plan.solution(), plan.windows()
    plan.close()
    p = O.Params(**kw)
    padded, mask = O.pad_stack(frames.astype(np.float64), (size, size))
    ref = O.xcorrnm2_tri(padded, mask, p, return_windows=True)
    m = kw["maxShift"]
    keep = np.ones((2 * m, 2 * m), bool)
    keep[m, m] = False
    worst = max(np.max(np.abs(win[k] - ref["windows"][(k, k)])[keep]) / np.max(np.abs(ref["windows"][(k, k)]))
                for k in range(n - 1))
    print(f"[gpu trimodes] refine masked={masked}: worst window rel err {worst:.2e}", flush=True)
    assert worst <= 2e-4
    d = np.array([tm["shiftsTriMat"][k, k] - ref["shiftsTriMat"][k, k] for k in range(n)])
    steps = np.abs(d) * 16
    assert np.all(np.abs(steps - np.round(steps)) < 1e-3) and np.all(np.round(steps) <= 1)
    assert np.all(tm["corrTriMat"][n - 1] == 0)
    t_ref = -np.stack([ref["shiftsTriMat"][k, k] for k in range(n)]).astype(np.float64)
    t_ref -= t_ref.mean(axis=0)
    dt = np.max(np.abs(sol["translations"] - t_ref))
    print(f"[gpu trimodes] refine masked={masked}: |dt|max {dt:.4f} px", flush=True)
    assert dt <= 0.0625 + 1e-6
