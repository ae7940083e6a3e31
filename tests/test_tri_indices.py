"""User-supplied pair masks: xcorrnm2_tri(triIndices=...) (zorro/zorro.py:773,
:552-557, :959-975).

CPU: the C ABI refuses masks outside the built path before touching a device
(entries on or below the diagonal, a triMode other than 'tri', preShift); the
oracle's mask path is pinned by the triMode patterns it must reproduce (a mask
equal to the 'diag' band gives triMode 'diag''s trimats exactly).
GPU: a mask of contiguous rows (the masks whose equations the solver can
hold; with a row with a gap the pairs are correlated and the solve refused)
against the oracle (pair shifts identical on the 1/16 grid up to one step at
near-ties; the solver on those trimats within 1e-9 px of the oracle's exact
vertex; translations within 0.02 px when every pair shift agrees), and
the facade's xcorrnm2_tri + shiftsSolver equal to alignImageStack with the
same mask (bit-identical translations)."""
import numpy as np
import pytest

from oracle import zorro_oracle as O
from zorro_amd._lib import ZrError
from zorro_amd.engine import MoviePlan, make_params


def _kw(mode, size, **extra):
    kw = dict(triMode=mode, maxShift=16, Bmode="conv", Brad=150, subPixReg=16, shiftMethod="lanczos",
              weightMode="logistic", peaksigThres=3.115, logisticK=5.0, logisticNu=0.15, originMode="centroid",
              suppressOrigin=True, fouCrop=(0, 0), shapePadded=(size, size), filterMode=None, doLazyFRC=False)
    kw.update(extra)
    return kw


def test_mask_refusals_before_any_device_call():
    n = 6
    lower = np.zeros((n, n), bool)
    lower[3, 1] = True
    with pytest.raises(ZrError, match="above the diagonal"):
        MoviePlan(make_params(**_kw("tri", 64)), n, 64, 64, triIndices=lower)
    up = np.triu(np.ones((n, n), bool), 1)
    with pytest.raises(ZrError, match="triMode 'tri'"):
        MoviePlan(make_params(**_kw("first", 64)), n, 64, 64, triIndices=up)
    with pytest.raises(ZrError, match="preShift"):
        MoviePlan(make_params(**_kw("tri", 64, preShift=True)), n, 64, 64, triIndices=up)
    with pytest.raises(IndexError):
        MoviePlan(make_params(**_kw("tri", 64)), n, 64, 64, triIndices=up[:5, :5])


def test_oracle_mask_reproduces_trimode_patterns():
    from zorro_amd.synth import make_movie
    n, size = 6, 64
    frames, _ = make_movie(n, size, size, seed=81, dose=4.0)
    padded, mask = O.pad_stack(frames.astype(np.float64), (size, size))
    kw = _kw("diag", size, diagWidth=2, maxShift=8)
    ref = O.xcorrnm2_tri(padded, mask, O.Params(**kw))
    band = O.tri_indices(n, "diag", 2, 0)
    d = np.subtract.outer(np.arange(n), np.arange(n))
    assert np.array_equal(band, (-d >= 1) & (-d <= 2))
    got = O.xcorrnm2_tri(padded, mask, O.Params(**dict(kw, triMode="tri")), triIndices=band)
    for k in ("shiftsTriMat", "corrTriMat", "peaksigTriMat"):
        assert np.array_equal(got[k], ref[k]), k
    assert np.all(got["corrTriMat"][~band] == 0)


def _mask(n, seed):
    """Row I takes J = I+1 .. I+len_I (len_I >= 1): the shape of masks whose
    rot90 equations (zorro/zorro.py:1296-1311) are single frame-to-frame spans;
    at most three pairs per row, so that every pair is well above the noise."""
    rng = np.random.default_rng(seed)
    m = np.zeros((n, n), bool)
    for I in range(n - 1):
        m[I, I + 1:I + 1 + int(rng.integers(1, min(3, n - 1 - I) + 1))] = True
    return m


def test_mask_gives_parallel_equations():
    """The reference's rot90 construction repeats a frame span for uneven rows:
    the GPU test's mask has two identical equation rows, so the solver's
    handling of parallel edges is exercised there."""
    A = O.build_equations(_mask(10, 83)[:-1, 1:])
    rows = [tuple(r) for r in A.astype(int)]
    assert len(rows) == 14 and len(set(rows)) == 13


@pytest.mark.gpu
def test_gpu_mask_matches_oracle():
    from zorro_amd.synth import make_movie
    n, size = 10, 256
    print(f"[gpu triIndices] n={n} {size}^2", flush=True)
    frames, _ = make_movie(n, size, size, seed=82, dose=4.0)
    tri = _mask(n, 83)
    kw = _kw("tri", size)
    plan = MoviePlan(make_params(**kw), n, size, size, triIndices=tri)
    assert plan.npairs == int(tri.sum())
    plan.set_frames(frames)
    plan.xcorr()
    plan.solve()
    tm, sol, m = plan.trimats(), plan.solution(), plan.equation_count()
    plan.close()
    assert np.array_equal(tm["corrTriMat"] != 0, tri) and m == int(tri.sum())
    ref = O.align_image_stack(frames, O.Params(**kw), triIndices=tri)
    so, sg = ref["xcorr"]["shiftsTriMat"], tm["shiftsTriMat"]
    steps = np.abs(sg[tri] - so[tri]) * 16
    assert np.all(np.abs(steps - np.round(steps)) < 1e-3) and np.all(np.round(steps) <= 1)
    # the solver on the GPU's own trimats: the oracle's exact vertex of the same equations
    ed = O.shifts_solver(tm["shiftsTriMat"], tm["corrTriMat"], tm["peaksigTriMat"], O.Params(**kw), (size, size))
    t_or = ed["translations"] - ed["translations"].mean(axis=0)
    ds = np.max(np.abs(sol["translations"] - t_or))
    # weighted-L1 objective of both solutions on the oracle's equations
    A, w = ed["Acoeff"], np.asarray(ed["weights"])
    S = tm["shiftsTriMat"][:-1, 1:, :]
    msk = tm["corrTriMat"][:-1, 1:].astype(bool)
    objs = []
    for ax in (0, 1):
        b = np.ravel(S[msk, ax]).astype(np.float64)
        objs.append((float(np.sum(w * np.abs(A @ sol["relativeEst"][:, ax] - b))),
                     float(np.sum(w * np.abs(A @ ed["relativeEst"][:, ax] - b)))))
    print(f"[gpu triIndices] objectives (gpu, oracle) y {objs[0]}, x {objs[1]}", flush=True)
    for og, oo in objs:
        assert og <= oo * (1 + 1e-9) + 1e-12, "the GPU solution is not optimal"
    # end to end when every pair shift is the oracle's (on a sparse graph one 1/16
    # step can move the weighted-L1 optimum to another vertex)
    same = bool(np.all(np.round(steps) == 0))
    dt = np.max(np.abs(sol["translations"] - ref["translations"]))
    print(f"[gpu triIndices] pairs {int(tri.sum())}, solver |dt|max {ds:.2e} px, end to end |dt|max {dt:.4f} px "
          f"(pair shifts identical: {same})", flush=True)
    # equal objectives: a different vertex of the same optimal face is allowed
    tie = all(abs(og - oo) <= 1e-9 * max(oo, 1e-12) for og, oo in objs)
    assert ds <= 1e-9 or tie
    if same and ds <= 1e-9:
        assert dt <= 0.02
    # a row with a gap gives an equation over non-adjacent frame steps, which the
    # exact solver's graph cannot hold: the pairs are correlated, the solve refused
    gap = tri.copy()
    gap[0, :] = False
    gap[0, [1, 3]] = True
    plan = MoviePlan(make_params(**kw), n, size, size, triIndices=gap)
    try:
        plan.set_frames(frames)
        plan.xcorr()
        tg = plan.trimats()
        assert np.array_equal(tg["corrTriMat"] != 0, gap)
        assert np.array_equal(tg["shiftsTriMat"][gap & tri], tm["shiftsTriMat"][gap & tri])
        with pytest.raises(ZrError, match="cannot represent"):
            plan.solve()
        with pytest.raises(ZrError, match="cannot represent"):
            plan.align()
    finally:
        plan.close()


@pytest.mark.gpu
def test_gpu_facade_xcorr_and_solver():
    from zorro_amd.registrator import ImageRegistrator
    from zorro_amd.synth import make_movie
    n, size = 8, 256
    print(f"[gpu triIndices] facade n={n} {size}^2", flush=True)
    frames, _ = make_movie(n, size, size, seed=84, dose=4.0)
    tri = _mask(n, 85)
    kw = _kw("tri", size)
    reg = ImageRegistrator()
    for k, v in kw.items():
        setattr(reg, k, v)
    reg.images = frames
    tm = reg.xcorrnm2_tri(triIndices=tri)
    ed = reg.shiftsSolver(tm["shiftsTriMat"], tm["corrTriMat"], tm["peaksigTriMat"])
    reg.triIndices = tri
    reg.alignImageStack()
    assert np.array_equal(ed["translations"], reg.errorDictList[-1]["translations"])
    assert ed["M"] == int(tri.sum()) and np.array_equal(ed["corrTriMat"], reg.errorDictList[-1]["corrTriMat"])
    reg.close()


@pytest.mark.gpu
def test_gpu_long_row_mask_solves_to_the_optimum():
    """Regression: rows reaching up to 9 frames apart repeat frame spans in the
    rot90 equations; before parallel edges were summed per span, the simplex
    stopped without proving optimality on this mask.  The GPU objective must
    not exceed the oracle's exact optimum on the same trimats."""
    from zorro_amd.synth import make_movie
    n, size = 10, 256
    print(f"[gpu triIndices] long rows n={n} {size}^2", flush=True)
    frames, _ = make_movie(n, size, size, seed=82, dose=4.0)
    rng = np.random.default_rng(83)
    tri = np.zeros((n, n), bool)
    for I in range(n - 1):
        tri[I, I + 1:I + 1 + int(rng.integers(1, n - I))] = True
    A = O.build_equations(tri[:-1, 1:])
    assert len({tuple(r) for r in A.astype(int)}) < A.shape[0]       # parallel equations present
    kw = _kw("tri", size)
    plan = MoviePlan(make_params(**kw), n, size, size, triIndices=tri)
    try:
        plan.set_frames(frames)
        plan.xcorr()
        plan.solve()
        tm, sol, info = plan.trimats(), plan.solution(), plan.solver_info()
    finally:
        plan.close()
    assert info["status_y"] == 1 and info["status_x"] == 1
    ed = O.shifts_solver(tm["shiftsTriMat"], tm["corrTriMat"], tm["peaksigTriMat"], O.Params(**kw), (size, size))
    w = np.asarray(ed["weights"])
    S = tm["shiftsTriMat"][:-1, 1:, :]
    msk = tm["corrTriMat"][:-1, 1:].astype(bool)
    for ax in (0, 1):
        b = np.ravel(S[msk, ax]).astype(np.float64)
        og = float(np.sum(w * np.abs(A @ sol["relativeEst"][:, ax] - b)))
        oo = float(np.sum(w * np.abs(A @ ed["relativeEst"][:, ax] - b)))
        print(f"[gpu triIndices] long rows axis {ax}: objective gpu {og:.6e} oracle {oo:.6e}", flush=True)
        assert og <= oo * (1 + 1e-9) + 1e-12
