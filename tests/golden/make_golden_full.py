"""Write the full-size golden fixtures (outputs only) of the registration path:

    python tests/golden/make_golden_full.py [case ...]

The frames of these movies are hundreds of MB, so a fixture holds no frames:
the GPU test regenerates them from the seed with ``zorro_amd.synth.make_movie``
(NumPy, deterministic) and checks the per-frame integer sums stored here before
it compares anything.  Each case runs the FP64 oracle
(``oracle.zorro_oracle.align_image_stack``) once, at the real size, and stores:

* the errorDict trimats, the translations, the logistic fit, the FRC curve;
* for imageSum and filtSum: the L2 norm, 64x64-pixel block means, a 256x256
  centre crop, the first/last rows and columns and 65536 pixels at seeded
  random positions (``sample_index``), so that a relative L2 error over the
  sample estimates the error over the whole image;
* the dose filter's per-frame cut-off radii and dose per frame.

Cases: the headline movie (40 x 4096^2, fouCrop 3072, dose + FRC); a
5760x4092 movie padded to 5760^2 (masked correlation, direct dose path at
Hp = 5760); the super-resolution counting movie (uint8 7676x7420 with a gain
reference, Fourier-binned to 3838x3710, padded to 3840^2, 0.5 A/px after
binning).  The C2 case needs about 30 GB of host memory and 20 minutes.
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import zorro_oracle as O  # noqa: E402
from zorro_amd.synth import make_movie  # noqa: E402

SAMPLE_SEED = 20240611
N_SAMPLE = 65536

BASE = dict(triMode="tri", maxShift=40, Bmode="conv", Brad=150, subPixReg=16, shiftMethod="lanczos",
            weightMode="autologistic", originMode="centroid", suppressOrigin=True, filterMode="dose",
            pixelsize=0.1, voltage=300.0, doLazyFRC=True, frc_box=512, frc_overlap=0.5)

CASES = {
    "full_c2_40x4096sq_crop3072": dict(
        n=40, h=4096, w=4096, seed=2, dose=1.0, jump=3.0,
        params=dict(BASE, fouCrop=(3072, 3072), shapePadded=(4096, 4096))),
    "full_5760x4092_pad5760_crop3072_8f": dict(
        n=8, h=5760, w=4092, seed=57, dose=1.0, jump=0.0,
        params=dict(BASE, fouCrop=(3072, 3072), shapePadded=(5760, 5760))),
    "full_c4_7676x7420_u8_gain_bin3838x3710_8f": dict(
        n=8, h=7676, w=7420, seed=76, dose=0.25, jump=0.0, gain_seed=77,
        params=dict(BASE, fouCrop=(3072, 3072), shapePadded=(3840, 3840), shapeBinned=(3838, 3710),
                    pixelsize=0.025)),
}


def case_frames(c):
    """The case's input stack as the GPU test rebuilds it: uint8 counts for the
    counting case (plus its float32 gain reference), float32 counts otherwise."""
    frames, drift = make_movie(c["n"], c["h"], c["w"], seed=c["seed"], dose=c["dose"], jump=c["jump"])
    gain = None
    if "gain_seed" in c:
        assert frames.max() <= 255
        frames = frames.astype(np.uint8)
        gain = np.random.default_rng(c["gain_seed"]).normal(1.0, 0.05, (c["h"], c["w"])).astype(np.float32)
    return frames, drift, gain


def frame_sums(frames):
    return frames.reshape(frames.shape[0], -1).astype(np.int64).sum(axis=1)


def sample_index(h, w):
    """Seeded pixel positions (flat indices) of the image samples."""
    return np.sort(np.random.default_rng(SAMPLE_SEED).choice(h * w, size=min(N_SAMPLE, h * w), replace=False))


def block_means(img, b=64):
    h, w = (img.shape[0] // b) * b, (img.shape[1] // b) * b
    return img[:h, :w].reshape(h // b, b, w // b, b).mean(axis=(1, 3))


def centre_crop(img, s=256):
    y0, x0 = (img.shape[0] - s) // 2, (img.shape[1] - s) // 2
    return img[y0:y0 + s, x0:x0 + s]


def image_summary(prefix, img):
    img = np.asarray(img, np.float64)
    h, w = img.shape
    return {f"{prefix}_l2": np.float64(np.linalg.norm(img)),
            f"{prefix}_blocks": block_means(img),
            f"{prefix}_centre": centre_crop(img).astype(np.float32),
            f"{prefix}_edges": np.stack([img[0], img[-1]]).astype(np.float32),
            f"{prefix}_edgecols": np.stack([img[:, 0], img[:, -1]]).astype(np.float32),
            f"{prefix}_sample": img.ravel()[sample_index(h, w)].astype(np.float32)}


def make_case(name, c):
    t0 = time.time()
    frames, drift, gain = case_frames(c)
    print(f"{name}: frames {frames.shape} {frames.dtype} in {time.time() - t0:.0f} s", flush=True)
    p = O.Params(**c["params"])
    stack = O.apply_gain(frames, gain) if gain is not None else frames
    res = O.align_image_stack(stack, p)
    xc, ed = res["xcorr"], res["errorDict"]
    out = dict(frame_sums=frame_sums(frames), drift=drift,
               meta=np.array(json.dumps(dict(name=name, **{k: v for k, v in c.items() if k != "params"},
                                             params=c["params"]))),
               shiftsTriMat=xc["shiftsTriMat"], corrTriMat=xc["corrTriMat"], peaksigTriMat=xc["peaksigTriMat"],
               originTriMat=xc["originTriMat"], translations=res["translations"], FRC=res["FRC"],
               frc_tiles=np.int64(res["frc_tiles"]),
               logistic=np.array([ed["peaksigThres"], ed["logisticK"], ed["logisticNu"]]),
               dose_thresq=np.asarray(res["dose"]["thresQ"], np.float64),
               dose_per_frame=np.float64(res["dose"]["dosePerFrame"]))
    out.update(image_summary("imageSum", res["imageSum"]))
    out.update(image_summary("filtSum", res["filtSum"]))
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print(f"{path}: {os.path.getsize(path) / 1024:.0f} KiB, {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    names = sys.argv[1:] or list(CASES)
    for k in names:
        make_case(k, CASES[k])
