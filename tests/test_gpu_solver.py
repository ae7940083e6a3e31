"""GPU weighted-L1 shift solver against the FP64 oracle on injected triangular
matrices (the inputs of the reference's shiftsSolver, zorro/zorro.py:1260).

The measured shifts are quantised the way the peak stage produces them
(multiples of 1/subPixReg, scaled by a float32 fouCrop factor), which makes
degenerate vertices (extra zero-residual equations) common: the solver must
still stop only at an optimal vertex.

Tolerances: translations identical to the oracle's vertex (|dt| <= 1e-9 px)
where the optimum is unique (continuous weights); for unweighted equations,
whose optimum may be a face, the objective must match to 1e-9 relative.
"""
import numpy as np
import pytest

from oracle import zorro_oracle as O
from zorro_amd.engine import MoviePlan, make_params

pytestmark = pytest.mark.gpu


def _trimats(n, seed, outliers=0.15):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 2))
    pos[1:] = np.cumsum(rng.normal(0.8, 0.3, (n - 1, 2)) * rng.choice([-1, 1], 2), axis=0)
    sh = np.zeros((n, n, 2), np.float32)
    ps = np.zeros((n, n), np.float32)
    co = np.zeros((n, n), np.float32)
    scale = np.float32(4.0 / 3.0)
    for i in range(n):
        for j in range(i + 1, n):
            d = pos[j] - pos[i] + rng.normal(0, 0.25, 2)
            bad = rng.random() < outliers
            if bad:
                d = d + rng.normal(0, 6, 2)
            sh[i, j] = (np.round(d * 16) / 16).astype(np.float32) * scale
            ps[i, j] = rng.uniform(2.0, 5.0) if bad else rng.uniform(6.0, 14.0)
            co[i, j] = rng.uniform(0.05, 0.3)
    return sh, co, ps


def _objective(t, sh, w_mode_weights, axis):
    n = sh.shape[0]
    iu = np.triu_indices(n, 1)
    p = -t[:, axis]
    r = p[iu[1]] - p[iu[0]] - sh[iu[0], iu[1], axis].astype(np.float64)
    return float(np.sum(w_mode_weights * np.abs(r)))


CASES = [(8, 1, "autologistic"), (20, 2, "autologistic"), (40, 3, "autologistic"), (40, 4, "corr"),
         (64, 5, "autologistic"), (24, 6, "unweighted"), (40, 7, "unweighted")]


@pytest.mark.parametrize("case", CASES, ids=[f"n{c[0]}_s{c[1]}_{c[2]}" for c in CASES])
def test_solver_on_injected_trimats(case):
    n, seed, wm = case
    print(f"[gpu solver] n={n} seed={seed} weightMode={wm}", flush=True)
    sh, co, ps = _trimats(n, seed)
    kw = dict(triMode="tri", maxShift=40, Bmode="conv", Brad=150, subPixReg=16, fouCrop=(0, 0),
              shapePadded=(256, 256), weightMode=wm, filterMode=None)
    plan = MoviePlan(make_params(**kw), n, 256, 256)
    plan.set_trimats(sh, co, ps)
    plan.solve()
    sol = plan.solution()
    info = plan.solver_info()
    plan.close()
    p = O.Params(**kw)
    ed = O.shifts_solver(sh, co, ps, p, (256, 256))
    t_or = ed["translations"] - ed["translations"].mean(axis=0)
    t_gpu = sol["translations"]
    print(f"[gpu solver] pivots y/x {info['pivots_y']}/{info['pivots_x']} status {info['status_y']}/{info['status_x']}",
          flush=True)
    assert info["status_y"] == 1 and info["status_x"] == 1, info
    w = np.asarray(ed["weights"], np.float64)
    for axis in range(2):
        og = _objective(t_gpu, sh, w, axis)
        oo = _objective(t_or, sh, w, axis)
        print(f"[gpu solver] axis {axis}: objective gpu {og:.12g} oracle {oo:.12g}", flush=True)
        assert og <= oo * (1 + 1e-9) + 1e-12
    if wm != "unweighted":
        d = np.max(np.abs(t_gpu - t_or))
        print(f"[gpu solver] |dt|max {d:.3e}", flush=True)
        assert d <= 1e-9
