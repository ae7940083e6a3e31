"""GPU parity of the outlier-pixel filter ('hot' in filterMode; hotpixFilter,
zorro/zorro.py:2497-2662) against the oracle on the same synthetic stacks with
planted hot, dead and clustered pixels (near the frame edges too, where the
reference's python slices shrink or vanish).

Tolerances (stated per result):
* with the normal-CDF parameters fixed to the oracle's L-BFGS-B fit: the two
  cut-offs equal as float32 up to 2e-6 absolute (float64 least squares in a
  different order), the dead/hot/clustered-pixel counts identical, filtered
  frames within 2e-5 x max|oracle| (float32 exp of a different library, float32
  means of at most 8 values in a different order);
* with explicit cut-offs: counts identical, frames as above;
* the filter's own fit (a deterministic grid + pattern search of the same
  objective): objective no worse than the oracle's L-BFGS-B point + 1e-9, mu and
  sigma within 2% of sigma, cut-offs within 0.05 sigma;
* end to end ('hot' then the registration, fixed fit): translations within
  0.02 px of the oracle's.
"""
import numpy as np
import pytest

from oracle import zorro_oracle as O
from zorro_amd.engine import MoviePlan, make_params
from zorro_amd.synth import make_movie

pytestmark = pytest.mark.gpu


def _stack(n, h, w, lam, seed):
    rng = np.random.default_rng(seed)
    st = rng.poisson(lam, (n, h, w)).astype(np.float32)
    st[:, 20, 30] += 10 * lam            # singleton hot
    st[:, h // 2, w // 3] = 0.0          # singleton dead
    st[:, 40:42, 70:72] += 10 * lam      # 2x2 cluster
    st[:, 0, 5:7] += 10 * lam            # cluster on the first row (empty python slice)
    st[:, h - 1, w - 2:w] += 10 * lam    # cluster on the last row / last column
    st[:, 100, 0] += 10 * lam            # singleton on the first column
    return st


def _run_gpu(st, **hot):
    n, h, w = st.shape
    pad = tuple(max(256, 1 << (int(v) - 1).bit_length()) for v in (h, w))
    kw = dict(triMode="diag", diagWidth=2, maxShift=20, Bmode="conv", Brad=150, fouCrop=(0, 0), shapePadded=pad,
              filterMode="hot", **hot)
    plan = MoviePlan(make_params(**kw), n, h, w)
    plan.set_frames(st)
    plan.xcorr()
    plan.synchronize()
    out, info = plan.frames(), plan.hotpix_info()
    plan.close()
    return out, info


def _compare(name, out, info, ref, rinfo):
    for k in ("guessDeadpix", "guessHotpix", "neighborPix"):
        print(f"[gpu hot] {name}: {k} gpu {info[k]} oracle {rinfo[k]}", flush=True)
        assert info[k] == rinfo[k], k
    err = float(np.max(np.abs(out - ref))) / float(np.max(np.abs(ref)))
    print(f"[gpu hot] {name}: frames max rel err {err:.2e}", flush=True)
    assert err <= 2e-5


CASES = [("1024sq_16f_lam4", 16, 1024, 1024, 4.0, 3), ("768x1000_40f_lam1", 40, 768, 1000, 1.0, 4)]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_filter_with_fixed_fit_matches_oracle(case):
    name, n, h, w, lam, seed = case
    print(f"[gpu hot] case={name}", flush=True)
    st = _stack(n, h, w, lam, seed)
    ref, rinfo = O.hotpix_filter(st)
    out, info = _run_gpu(st, hotMuSigma=(rinfo["mu"], rinfo["sigma"]))
    for k in ("cutoffLower", "cutoffUpper"):
        print(f"[gpu hot] {name}: {k} gpu {info[k]:.7f} oracle {rinfo[k]:.7f}", flush=True)
        assert abs(info[k] - rinfo[k]) <= 2e-6, k
    _compare(name, out, info, ref, rinfo)


def test_filter_with_explicit_cutoffs_matches_oracle():
    st = _stack(12, 256, 256, 4.0, 9)
    print("[gpu hot] case=256sq_explicit_cutoffs", flush=True)
    ref, rinfo = O.hotpix_filter(st, cutoffLower=-4.0, cutoffUpper=3.25)
    out, info = _run_gpu(st, hotMuSigma=(rinfo["mu"], rinfo["sigma"]), hotCutoffs=(-4.0, 3.25))
    assert info["cutoffLower"] == -4.0 and info["cutoffUpper"] == 3.25
    _compare("256sq_explicit_cutoffs", out, info, ref, rinfo)


def test_own_fit_is_no_worse_than_reference_optimiser():
    st = _stack(16, 1024, 1024, 4.0, 5)
    print("[gpu hot] case=own_fit_1024sq", flush=True)
    ref, rinfo = O.hotpix_filter(st)
    _, info = _run_gpu(st)
    sg = rinfo["sigma"]
    print(f"[gpu hot] own fit: mu {info['mu']:.5f} / {rinfo['mu']:.5f}, sigma {info['sigma']:.5f} / {sg:.5f}, "
          f"objective {info['objective']:.6e} / {rinfo['objective']:.6e}, cut-offs "
          f"({info['cutoffLower']:.4f}, {info['cutoffUpper']:.4f}) / ({rinfo['cutoffLower']:.4f}, "
          f"{rinfo['cutoffUpper']:.4f})", flush=True)
    assert info["objective"] <= rinfo["objective"] + 1e-9
    assert abs(info["mu"] - rinfo["mu"]) <= 0.02 * sg and abs(info["sigma"] - sg) <= 0.02 * sg
    assert abs(info["cutoffLower"] - rinfo["cutoffLower"]) <= 0.05
    assert abs(info["cutoffUpper"] - rinfo["cutoffUpper"]) <= 0.05


def test_hot_then_registration_matches_oracle():
    n, h, w = 8, 512, 512
    print("[gpu hot] case=end_to_end_512sq", flush=True)
    frames, _ = make_movie(n, h, w, seed=21, dose=2.0)
    frames = frames.astype(np.float32)
    frames[:, 100, 200] += 60.0
    frames[:, 300:302, 50:52] += 60.0
    _, rinfo = O.hotpix_filter(frames)
    ms = (rinfo["mu"], rinfo["sigma"])
    kw = dict(triMode="tri", maxShift=20, Bmode="conv", Brad=150, subPixReg=16, fouCrop=(0, 0), shapePadded=(h, w),
              filterMode="hot", doLazyFRC=False)
    plan = MoviePlan(make_params(**kw, hotMuSigma=ms), n, h, w)
    plan.set_frames(frames)
    plan.align()
    plan.synchronize()
    t = plan.solution()["translations"]
    plan.close()
    res = O.align_image_stack(frames, O.Params(**kw, hotMuSigma=ms))
    d = float(np.max(np.abs(t - res["translations"])))
    print(f"[gpu hot] end to end: max |dt| {d:.4f} px", flush=True)
    assert d <= 0.02
