"""Seeded synthetic cryo-EM movies with known drift.

The generator follows the workload definitions the benchmarks are quoted on:

* phantom: white N(0,1) noise low-passed by an ideal disc at 0.1 Nyquist
  (|f| <= 0.05 cycles/px), normalised to unit std and mapped affinely to
  ``dose * (1 + contrast * p)`` (contrast 0.3, negatives clipped to 0);
* drift: a smooth random walk, per-frame step ~ N(0.8, 0.3) px along a heading
  that starts uniformly at random and turns by U(-15, 15) degrees per frame,
  plus an optional exponentially decaying initial jump (``jump`` px in the
  first step, ``jump*e^-k`` added to step k along the initial heading);
* each frame is the phantom shifted by the cumulative drift with the Fourier
  shift theorem (periodic), then Poisson-sampled.

Frame J holds ``phantom(x - d_J)``, so the reference's pair shift (I, J) is
d_J - d_I and the recovered translations are ``-(d_J - mean(d))``.

Two back ends produce the same distribution: NumPy (small movies, tests) and
torch on the GPU (full-size benchmark movies; random streams differ, so the two
do not give identical pixels for one seed).
"""
from __future__ import annotations

import math

import numpy as np


def drift_path(n_frames, seed, step_mean=0.8, step_std=0.3, turn_deg=15.0, jump=0.0):
    """Cumulative drift d [n_frames, 2] (y, x) in pixels, d[0] = 0."""
    rng = np.random.default_rng(seed + 7919)
    heading = rng.uniform(0.0, 2.0 * math.pi)
    h0 = heading
    d = np.zeros((n_frames, 2))
    for k in range(1, n_frames):
        step = rng.normal(step_mean, step_std)
        heading += math.radians(rng.uniform(-turn_deg, turn_deg))
        dy = step * math.sin(heading)
        dx = step * math.cos(heading)
        if jump:
            j = jump * math.exp(-(k - 1))
            dy += j * math.sin(h0)
            dx += j * math.cos(h0)
        d[k] = d[k - 1] + (dy, dx)
    return d


def _lowpass_disc(h, w):
    fy = np.fft.fftfreq(h)[:, None]
    fx = np.fft.rfftfreq(w)[None, :]
    return (fy * fy + fx * fx) <= 0.05 ** 2


def phantom(h, w, seed, dose, contrast=0.3):
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((h, w))
    F = np.fft.rfft2(noise) * _lowpass_disc(h, w)
    p = np.fft.irfft2(F, s=(h, w))
    p = (p - p.mean()) / p.std()
    return np.clip(dose * (1.0 + contrast * p), 0.0, None)


def make_movie(n_frames, h, w, seed=0, dose=1.5, jump=0.0, poisson=True):
    """NumPy generator.  Returns (frames float32 [n,h,w], drift [n,2])."""
    ph = phantom(h, w, seed, dose)
    d = drift_path(n_frames, seed, jump=jump)
    P = np.fft.rfft2(ph)
    fy = np.fft.fftfreq(h)[:, None]
    fx = np.fft.rfftfreq(w)[None, :]
    rng = np.random.default_rng(seed + 104729)
    out = np.empty((n_frames, h, w), dtype=np.float32)
    for k in range(n_frames):
        ramp = np.exp(-2j * np.pi * (fy * d[k, 0] + fx * d[k, 1]))
        fr = np.clip(np.fft.irfft2(P * ramp, s=(h, w)), 0.0, None)
        out[k] = rng.poisson(fr) if poisson else fr
    return out, d


def make_movie_torch(n_frames, h, w, seed=0, dose=1.0, jump=0.0, device="cuda"):
    """torch generator for full-size movies on the GPU (synthetic input only).
    Returns (frames float32 tensor [n,h,w] on ``device``, drift [n,2] numpy)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    noise = torch.randn((h, w), generator=g, device=device, dtype=torch.float32)
    fy = torch.fft.fftfreq(h, device=device)[:, None]
    fx = torch.fft.rfftfreq(w, device=device)[None, :]
    disc = (fy * fy + fx * fx) <= 0.05 ** 2
    p = torch.fft.irfft2(torch.fft.rfft2(noise) * disc, s=(h, w))
    p = (p - p.mean()) / p.std()
    ph = torch.clamp(dose * (1.0 + 0.3 * p), min=0.0)
    P = torch.fft.rfft2(ph)
    d = drift_path(n_frames, seed, jump=jump)
    out = torch.empty((n_frames, h, w), device=device, dtype=torch.float32)
    for k in range(n_frames):
        ramp = torch.exp(-2j * math.pi * (fy * float(d[k, 0]) + fx * float(d[k, 1])))
        fr = torch.clamp(torch.fft.irfft2(P * ramp, s=(h, w)), min=0.0)
        out[k] = torch.poisson(fr, generator=g)
    return out, d
