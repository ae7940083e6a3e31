"""The user's single mask (ImageRegistrator.masks of shape [1,H,W] or [H,W],
zorro/zorro.py:584-589, zero-padded with the frames by padStack :1942-1955):
the correlation runs under the frame region times the mask with the general
five-term mnxc2 (:1014-1097).

CPU tests: the facade's checks (a mask per frame is refused with the
reference's own failure, values must be 0 or 1, an all-ones mask is the
default mask).  GPU tests: lag windows, trimats and end-to-end translations
against the FP64 oracle run with the same padded mask.

Tolerances (as for the padded-frame masked path): lag windows max |gpu -
oracle| <= 2e-4 * max |oracle|; sub-pixel shifts identical on the 1/16 grid
except near-ties (<= 5% of pairs, one grid step); translations within 0.02 px.
"""
import numpy as np
import pytest

from oracle import zorro_oracle as O
from zorro_amd.registrator import ImageRegistrator


def _mask(h, w):
    m = np.ones((h, w), np.float32)
    m[h // 5: h // 5 + h // 6, w // 3: w // 3 + w // 4] = 0.0     # a blanked detector patch
    m[:, : w // 25] = 0.0                                          # a dead strip on the left
    yy, xx = np.mgrid[:h, :w]
    m[(yy - 0.7 * h) ** 2 + (xx - 0.7 * w) ** 2 < (0.08 * min(h, w)) ** 2] = 0.0   # and a disc
    return m


def test_mask_per_frame_is_refused_like_the_reference():
    reg = ImageRegistrator()
    reg.images = np.zeros((4, 64, 64), np.float32)
    reg.masks = np.ones((4, 64, 64), np.float32)
    with pytest.raises(NotImplementedError, match="zorro/zorro.py:970"):
        reg._params()


def test_mask_values_and_default():
    reg = ImageRegistrator()
    reg.images = np.zeros((4, 64, 64), np.float32)
    reg.masks = np.full((64, 64), 0.5, np.float32)
    with pytest.raises(ValueError, match="0 and 1"):
        reg._params()
    reg.masks = np.ones((1, 64, 64), bool)
    assert reg._user_mask() is None and reg._params().user_mask == 0
    reg.masks = _mask(64, 64)[None]
    assert reg._params().user_mask == 1 and reg._user_mask().shape == (64, 64)
    reg.masks = np.ones((32, 64), np.float32)
    with pytest.raises(ValueError, match="shape"):
        reg._params()


CASES = [
    # (name, n, h, w, shapePadded, fouCrop, maxShift, seed)
    ("512sq_unpadded", 8, 512, 512, (512, 512), (0, 0), 20, 51),
    ("450x480_pad512", 6, 450, 480, (512, 512), (0, 0), 20, 52),
    ("1000x900_pad1024_crop512", 5, 1000, 900, (1024, 1024), (512, 512), 30, 53),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_user_mask_xcorr_parity(case):
    from zorro_amd.engine import MoviePlan, make_params
    from zorro_amd.synth import make_movie
    name, n, h, w, pad, crop, m, seed = case
    print(f"[gpu user mask] case={name}", flush=True)
    frames, _ = make_movie(n, h, w, seed=seed, dose=1.5)
    um = _mask(h, w)
    kw = dict(triMode="tri", maxShift=m, Bmode="conv", Brad=150, subPixReg=16, fouCrop=crop, shapePadded=pad)
    plan = MoviePlan(make_params(user_mask=True, **kw), n, h, w)
    try:
        assert plan.info()["unmasked"] == 0
        plan.set_frames(frames)
        plan.set_mask(um)
        plan.xcorr()
        tm, win = plan.trimats(), plan.windows()
    finally:
        plan.close()
    padded, region = O.pad_stack(frames.astype(np.float64), pad)
    mask = region.copy()
    mask[0, :h, :w] &= um.astype(bool)
    ref = O.xcorrnm2_tri(padded, mask, O.Params(**kw), return_windows=True)
    pairs = [(I, J) for I in range(n) for J in range(I + 1, n)]
    keep = np.ones((2 * m, 2 * m), bool)
    keep[m, m] = False
    worst = max(float(np.max(np.abs(win[p] - ref["windows"][(I, J)])[keep]) / np.max(np.abs(ref["windows"][(I, J)])))
                for p, (I, J) in enumerate(pairs))
    print(f"[gpu user mask] {name}: worst window rel err {worst:.2e}", flush=True)
    assert worst <= 2e-4
    scale = np.array([pad[0] / (crop[0] or pad[0]), pad[1] / (crop[1] or pad[1])], np.float32)
    iu = np.triu_indices(n, 1)
    steps = np.abs(tm["shiftsTriMat"] - ref["shiftsTriMat"])[iu] / (scale / 16.0)
    assert np.all(np.abs(steps - np.round(steps)) < 1e-3) and np.all(np.round(steps) <= 1)
    assert int(np.sum(np.round(steps) > 0)) <= max(1, int(0.05 * steps.size))


@pytest.mark.gpu
def test_user_mask_end_to_end_translations():
    n, h, w = 10, 512, 512
    frames, _ = make_movie_np(n, h, w, 61)
    um = _mask(h, w)
    reg = ImageRegistrator()
    kw = dict(triMode="tri", maxShift=20, Bmode="conv", Brad=150, subPixReg=16, fouCrop=(0, 0),
              shapePadded=(h, w), weightMode="autologistic", originMode="centroid", doLazyFRC=False)
    for k, v in kw.items():
        setattr(reg, k, v)
    reg.images = frames.copy()
    reg.masks = um[None]
    reg.alignImageStack()
    p = O.Params(**{k: v for k, v in kw.items() if k != "doLazyFRC"}, doLazyFRC=False)
    region = np.ones((1, h, w), bool)
    region[0] &= um.astype(bool)
    ref = O.align_image_stack(frames.astype(np.float64), p, mask=region)
    dt = float(np.max(np.abs(reg.translations - ref["translations"])))
    print(f"[gpu user mask] end to end |dt|max {dt:.4f} px", flush=True)
    assert dt <= 0.02


def make_movie_np(n, h, w, seed):
    from zorro_amd.synth import make_movie
    return make_movie(n, h, w, seed=seed, dose=1.5)
