"""GPU parity of the input stage: gain reference product and Fourier binning
(loadData, zorro/zorro.py:3416-3419, :3455-3482; binStack('fourier'),
:1826-1904), including the super-resolution counting shape 7676x7420 -> 3838x3710
whose DFT lengths are not 2^a 3^b 5^c.

Tolerances (stated per result):
* binned frames: relative L2 error <= 2e-6 per frame against the FP64 oracle
  (float32 transforms: the polyphase form for exact halving, chirp-z
  transforms otherwise);
* gain product without binning: bit-identical (both float32 gainRef * data / gain);
* end to end on a binned, gain-corrected, padded movie: translations within
  0.02 px of the oracle; imageSum relative L2 <= 1e-4.
"""
import numpy as np
import pytest

from oracle import zorro_oracle as O
from zorro_amd.engine import MoviePlan, make_params
from zorro_amd.synth import make_movie

pytestmark = pytest.mark.gpu


def _counts(n, h, w, seed, dose):
    f, d = make_movie(n, h, w, seed=seed, dose=dose)
    assert f.max() <= 255
    return f.astype(np.uint8), d


def _gain(h, w, seed):
    return np.random.default_rng(seed).normal(1.0, 0.05, (h, w)).astype(np.float32)


CASES = [
    # (name, n, raw h, raw w, binned, padded, dtype, gain, gain scalar)
    ("superres_7676x7420_to_3838x3710_u8_gain", 2, 7676, 7420, (3838, 3710), (3840, 3840), "u8", True, None),
    ("300x260_to_150x130_f32", 3, 300, 260, (150, 130), (256, 256), "f32", False, None),
    ("1022x990_to_512x480_u16_gain_scalar", 2, 1022, 990, (512, 480), (512, 512), "u16", True, 1.7),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_binned_frames_match_oracle(case):
    name, n, h, w, sb, pad, dt, use_gain, gs = case
    print(f"[gpu bin] case={name}", flush=True)
    raw, _ = _counts(n, h, w, seed=91, dose=0.25)
    raw = {"u8": raw, "u16": raw.astype(np.uint16), "f32": raw.astype(np.float32)}[dt]
    gain = _gain(h, w, 92) if use_gain else None
    crop = (3072, 3072) if pad[0] > 3072 else (0, 0)
    kw = dict(triMode="tri", maxShift=20, fouCrop=crop, shapePadded=pad, shapeBinned=sb)
    plan = MoviePlan(make_params(**kw), n, h, w)
    assert (plan.hb, plan.wb) == sb
    plan.set_frames_raw(raw, gain=gain, gain_scalar=gs)
    got = plan.frames()
    plan.close()
    ref, scale = O.bin_stack(O.apply_gain(raw, gain, gs), sb)
    for J in range(n):
        rel = np.linalg.norm(got[J] - ref[J]) / np.linalg.norm(ref[J])
        print(f"[gpu bin] {name}: frame {J} rel L2 {rel:.2e}", flush=True)
        assert rel <= 2e-6


def test_gain_without_binning_is_exact():
    n, h, w = 3, 256, 200
    raw, _ = _counts(n, h, w, seed=93, dose=1.0)
    gain = _gain(h, w, 94)
    plan = MoviePlan(make_params(triMode="tri", maxShift=20, fouCrop=(0, 0), shapePadded=(256, 256)), n, h, w)
    plan.set_frames_raw(raw, gain=gain, gain_scalar=2.5)
    got = plan.frames()
    plan.close()
    assert np.array_equal(got, O.apply_gain(raw, gain, 2.5))


def test_binned_gain_movie_end_to_end():
    """C4 in small: uint8 counts at twice the size, a gain reference, 2x Fourier
    binning to a non-square frame that is then padded (masked correlation),
    dose filter at the binned pixel size."""
    n, h, w, sb = 8, 960, 900, (480, 450)
    print(f"[gpu bin] end to end n={n} {h}x{w} -> {sb}", flush=True)
    raw, drift = _counts(n, h, w, seed=95, dose=0.25)
    gain = _gain(h, w, 96)
    kw = dict(triMode="tri", maxShift=20, Bmode="conv", Brad=150, subPixReg=16, shiftMethod="lanczos",
              weightMode="autologistic", fouCrop=(0, 0), shapePadded=(512, 512), shapeBinned=sb, filterMode="dose",
              pixelsize=0.05, voltage=300.0, doLazyFRC=True, frc_box=256)
    plan = MoviePlan(make_params(**kw), n, h, w)
    plan.set_frames_raw(raw, gain=gain)
    plan.align()
    sol, sums = plan.solution(), plan.sums()
    plan.close()
    ref = O.align_image_stack(O.apply_gain(raw, gain), O.Params(**kw))
    dt = np.max(np.abs(sol["translations"] - ref["translations"]))
    rel = np.linalg.norm(sums["imageSum"] - ref["imageSum"]) / np.linalg.norm(ref["imageSum"])
    relf = np.linalg.norm(sums["filtSum"] - ref["filtSum"]) / np.linalg.norm(ref["filtSum"])
    print(f"[gpu bin] end to end: |dt|max {dt:.4f} px, imageSum rel {rel:.2e}, filtSum rel {relf:.2e}", flush=True)
    assert dt <= 0.02 and rel <= 1e-4 and relf <= 1e-4


@pytest.mark.parametrize("binned", [False, True], ids=["unbinned", "binned"])
def test_packed4_and_int16_inputs(binned):
    """MRC mode 101 (4-bit counts, two per byte, first pixel in the low nibble)
    and mode 1 (int16) frames read on the GPU: the gain product is bit-identical
    to the oracle's float32 product of the unpacked samples, and the binned
    frames equal those of the same samples given as uint8 / float32."""
    from zorro_amd import mrc
    n, h, w = 3, 300, 260
    sb = (150, 130) if binned else None
    print(f"[gpu bin] packed4/int16 binned={binned}", flush=True)
    rng = np.random.default_rng(97)
    c4 = rng.integers(0, 16, (n, h, w)).astype(np.uint8)
    s16 = rng.integers(-3000, 3000, (n, h, w)).astype(np.int16)
    gain = _gain(h, w, 98)
    kw = dict(triMode="tri", maxShift=20, fouCrop=(0, 0), shapePadded=(256, 256) if binned else (512, 512))
    if binned:
        kw["shapeBinned"] = sb
    outs = {}
    for name, frames, extra in (("u4", mrc.pack4(c4), {"packed4": True}), ("u8", c4, {}),
                                ("i16", s16, {}), ("f32", s16.astype(np.float32), {})):
        plan = MoviePlan(make_params(**kw), n, h, w)
        plan.set_frames_raw(frames, gain=gain, gain_scalar=1.5, **extra)
        outs[name] = plan.frames()
        plan.close()
    assert np.array_equal(outs["u4"], outs["u8"])
    assert np.array_equal(outs["i16"], outs["f32"])
    if not binned:
        assert np.array_equal(outs["u4"], O.apply_gain(c4, gain, 1.5))
        assert np.array_equal(outs["i16"], O.apply_gain(s16, gain, 1.5))
    plan = MoviePlan(make_params(**kw), n, h, w)
    with pytest.raises(ValueError):
        plan.set_frames_raw(mrc.pack4(c4)[:, :, :-1], packed4=True)
    plan.close()
