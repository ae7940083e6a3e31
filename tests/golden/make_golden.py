"""Write the golden fixtures of the registration path (run from the repository root):

    python tests/golden/make_golden.py

Each case is a small seeded movie from ``zorro_amd.synth.make_movie`` (Poisson
counts, stored exactly as uint8) run through the FP64 oracle
(``oracle.zorro_oracle.align_image_stack``).  A fixture holds the inputs and the
expected outputs: the errorDict trimats, the translations, the FRC curve, and
for the image outputs (imageSum, filtSum) their L2 norms, 8x8-block means and a
full-precision 64x64 centre crop.  The reference's own tests hold no vectors for
this path, so these pin the oracle against regressions and give the GPU tests
inputs and answers that do not depend on re-running the oracle.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import zorro_oracle as O  # noqa: E402
from zorro_amd.synth import make_movie  # noqa: E402

BASE = dict(triMode="tri", maxShift=20, Bmode="conv", Brad=150, subPixReg=16, shiftMethod="lanczos",
            weightMode="autologistic", originMode="centroid", suppressOrigin=True, filterMode="dose",
            pixelsize=0.1, voltage=300.0, doLazyFRC=True, frc_box=256, frc_overlap=0.5)

CASES = {
    "unmasked_512sq_6f": dict(n=6, h=512, w=512, seed=41, dose=1.5,
                              params=dict(BASE, fouCrop=(0, 0), shapePadded=(512, 512))),
    "masked_450x480_pad512_6f": dict(n=6, h=450, w=480, seed=42, dose=1.5,
                                     params=dict(BASE, fouCrop=(0, 0), shapePadded=(512, 512))),
    "masked_450x480_pad512_crop256_6f": dict(n=6, h=450, w=480, seed=43, dose=1.5,
                                             params=dict(BASE, fouCrop=(256, 256), shapePadded=(512, 512))),
}


def block_means(img, b=8):
    h, w = (img.shape[0] // b) * b, (img.shape[1] // b) * b
    return img[:h, :w].reshape(h // b, b, w // b, b).mean(axis=(1, 3))


def centre_crop(img, s=64):
    y0, x0 = (img.shape[0] - s) // 2, (img.shape[1] - s) // 2
    return img[y0:y0 + s, x0:x0 + s]


def image_summary(prefix, img):
    img = np.asarray(img, np.float64)
    return {f"{prefix}_l2": np.float64(np.linalg.norm(img)), f"{prefix}_blocks": block_means(img),
            f"{prefix}_centre": centre_crop(img)}


def make_case(name, c):
    frames, drift = make_movie(c["n"], c["h"], c["w"], seed=c["seed"], dose=c["dose"])
    assert frames.min() >= 0 and frames.max() <= 255 and np.all(frames == np.round(frames))
    p = O.Params(**c["params"])
    res = O.align_image_stack(frames.astype(np.float64), p)
    xc, ed = res["xcorr"], res["errorDict"]
    out = dict(frames_u8=frames.astype(np.uint8), drift=drift,
               meta=np.array(json.dumps(dict(name=name, n=c["n"], h=c["h"], w=c["w"], seed=c["seed"],
                                             dose=c["dose"], params=c["params"]))),
               shiftsTriMat=xc["shiftsTriMat"], corrTriMat=xc["corrTriMat"], peaksigTriMat=xc["peaksigTriMat"],
               originTriMat=xc["originTriMat"], translations=res["translations"], FRC=res["FRC"],
               logistic=np.array([ed["peaksigThres"], ed["logisticK"], ed["logisticNu"]]))
    out.update(image_summary("imageSum", res["imageSum"]))
    out.update(image_summary("filtSum", res["filtSum"]))
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print(f"{path}: {os.path.getsize(path) / 1024:.0f} KiB")


if __name__ == "__main__":
    for k, v in CASES.items():
        make_case(k, v)
