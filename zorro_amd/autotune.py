"""Launch-parameter autotuner for the registration path.

The reference plans its FFTs with FFTW_MEASURE and keeps FFTW wisdom in a
pickle (``zorro/zorro_util.py:966-1066``, ``fftw_wisdom.pkl``).  The GPU path
has no FFT planner to call, but a few launch shapes of its own kernels matter
per detector shape: the number of frames per forward-spectra launch pair
(``k1_batch``: the row pass's intermediate of a small batch stays in the
256 MB last-level cache), the lag-window columns each wave of the pair
transform walks (``k2_cols_per_wave``: L2 reuse of a column group against the
number of workgroups) and the pair-list entries each wave walks in the pair
transform's pair walk (``k2_pairs_per_wave``: one column of every frame per
XCD at a time, against the per-workgroup twiddle-table load).  ``tune`` times
whole movies for candidate values (coordinate descent: the batch, then the
columns, then the pairs), and the winners are kept per (frames, shape,
padding, crop, maxShift, masked, dose, device) in a JSON wisdom file,
``tuning.json`` beside this module by default.  Results do not depend on the
knobs; only the time does.

    python -m zorro_amd.autotune --frames 40 --shape 4096 4096 --crop 3072 3072
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from .engine import MoviePlan, make_params

HERE = os.path.dirname(os.path.abspath(__file__))
WISDOM = os.path.join(HERE, "tuning.json")
CANDIDATES = {"k1_batch": [1, 2, 3, 4, 8], "k2_cols_per_wave": [2, 4, 6, 8, 12, 16],
              "k2_pairs_per_wave": [8, 16, 32, 48]}
KNOBS = tuple(CANDIDATES)


def device_name(device=0):
    import torch
    return torch.cuda.get_device_properties(device).name


def tuning_key(n, h, w, settings, device=0):
    """Wisdom key of a movie class: the sizes and settings that change the launches."""
    p = make_params(**settings)
    masked = (p.shape_padded[0] or h) != h or (p.shape_padded[1] or w) != w
    return (f"{n}x{h}x{w}|pad{p.shape_padded[0]}x{p.shape_padded[1]}|crop{p.fou_crop[0]}x{p.fou_crop[1]}|"
            f"m{p.max_shift}|{'masked' if masked else 'unmasked'}|{'dose' if p.do_dose else 'nodose'}|"
            f"tri{p.tri_mode}|{device_name(device)}")


def load_wisdom(path=WISDOM):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save_wisdom(wisdom, path=WISDOM):
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(wisdom, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tuned_settings(n, h, w, settings, path=WISDOM, device=0):
    """settings plus the knobs the wisdom file holds for this movie class (unchanged if none)."""
    entry = load_wisdom(path).get(tuning_key(n, h, w, settings, device))
    if not entry:
        return dict(settings)
    return {**settings, **{k: int(entry["best"][k]) for k in KNOBS if k in entry.get("best", {})}}


def time_movie(frames, settings, reps=5, warmup=2, device=0):
    """ms per movie of zr_align replayed as a graph on a plan for these settings;
    also the spectra / pair-transform / rest split from the plan's HIP events."""
    n, h, w = (int(v) for v in frames.shape)
    plan = MoviePlan(make_params(**settings), n, h, w, device=device)
    try:
        plan.set_frames(frames)
        plan.use_graph(True)
        plan.set_timing(True)
        for _ in range(warmup):
            plan.align()
        plan.synchronize()
        plan.stage_times(reset=True)
        t0 = time.perf_counter()
        for _ in range(reps):
            plan.align()
        plan.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        st = {k: v[0] / v[1] for k, v in plan.stage_times().items() if v[1]}
    finally:
        plan.close()
    return ms, st


def tune(frames, settings, reps=5, candidates=CANDIDATES, device=0, log=print):
    """Coordinate descent over the knobs; returns (best knobs, table of trials)."""
    best = {k: 0 for k in KNOBS}
    best_ms, _ = time_movie(frames, settings, reps, device=device)
    trials = [dict(knobs=dict(best), ms=best_ms)]
    log(f"[autotune] defaults: {best_ms:.3f} ms/movie")
    for knob in KNOBS:
        for v in candidates[knob]:
            trial = {**best, knob: v}
            ms, st = time_movie(frames, {**settings, **trial}, reps, device=device)
            trials.append(dict(knobs=dict(trial), ms=ms, stages=st))
            log(f"[autotune] {trial}: {ms:.3f} ms/movie {', '.join(f'{k} {x:.3f}' for k, x in st.items())}")
            if ms < best_ms * 0.995:
                best, best_ms = trial, ms
    return dict(best=best, ms=best_ms), trials


def main():
    ap = argparse.ArgumentParser(description="tune the launch knobs for one movie class and store the wisdom")
    ap.add_argument("--frames", type=int, default=40)
    ap.add_argument("--shape", type=int, nargs=2, default=(4096, 4096))
    ap.add_argument("--pad", type=int, nargs=2, default=None)
    ap.add_argument("--crop", type=int, nargs=2, default=(3072, 3072))
    ap.add_argument("--max-shift", type=int, default=40)
    ap.add_argument("--no-dose", action="store_true")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--wisdom", default=WISDOM)
    args = ap.parse_args()
    import torch
    from .synth import make_movie_torch
    n, (h, w) = args.frames, args.shape
    settings = dict(triMode="tri", maxShift=args.max_shift, Bmode="conv", Brad=150, subPixReg=16,
                    shiftMethod="lanczos", weightMode="autologistic", fouCrop=tuple(args.crop),
                    shapePadded=tuple(args.pad or (h, w)), filterMode=None if args.no_dose else "dose",
                    pixelsize=0.1, voltage=300.0, doLazyFRC=True, frc_box=512)
    frames, _ = make_movie_torch(n, h, w, seed=5, dose=1.0, device="cuda:0")
    torch.cuda.synchronize()
    res, trials = tune(frames, settings, args.reps)
    key = tuning_key(n, h, w, settings)
    wisdom = load_wisdom(args.wisdom)
    wisdom[key] = dict(best=res["best"], ms=round(res["ms"], 4),
                       trials=[dict(knobs=t["knobs"], ms=round(t["ms"], 4)) for t in trials])
    save_wisdom(wisdom, args.wisdom)
    print(json.dumps({"key": key, **wisdom[key]}))


if __name__ == "__main__":
    main()
