"""GPU parity of the independent even/odd FRC (evenOddFouRingCorr with the
reference's steps made runnable, zorro/zorro.py:2099-2224): even and odd frames
registered on their own under the 64-pixel edge mask, their sums aligned with
triMode 'first', and the tiled FRC of the two aligned sums cropped by the even
half's translations.  The GPU registers each half on the frames cropped to the
mask rectangle (the same masked correlation) and shifts the whole frames; the
oracle uses the edge mask itself.

Tolerances: half translations within 0.02 px of the oracle's; FRC curve within
2e-3 absolute per ring.
"""
import numpy as np
import pytest

from oracle import zorro_oracle as O
from zorro_amd.registrator import ImageRegistrator
from zorro_amd.synth import make_movie

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [16, 17], ids=["16f", "17f_uneven_halves"])
def test_even_odd_frc_matches_oracle(n):
    h = w = 1024
    print(f"[gpu evenodd] case={n} frames {h}x{w}", flush=True)
    frames, _ = make_movie(n, h, w, seed=41 + n, dose=2.0)
    kw = dict(triMode="tri", maxShift=20, Bmode="conv", Brad=150, subPixReg=16, fouCrop=(0, 0), shapePadded=(h, w),
              weightMode="autologistic", originMode="centroid")
    reg = ImageRegistrator()
    for k, v in kw.items():
        setattr(reg, k, v)
    reg.filterMode = None
    reg.doEvenOddFRC = True
    reg.images = frames
    reg.evenOddFouRingCorr(box=(256, 256), overlap=0.5)
    reg.close()
    frc, te, to = O.even_odd_frc(frames, O.Params(**kw, filterMode=None), box=256, overlap=0.5)
    dte = float(np.max(np.abs(reg.transEven - te)))
    dto = float(np.max(np.abs(reg.transOdd - to)))
    g = np.asarray(reg.FRC)[:128]
    dfrc = float(np.max(np.abs(g - frc[:128])))
    print(f"[gpu evenodd] {n} frames: |dt| even {dte:.4f} odd {dto:.4f} px, FRC max |d| {dfrc:.2e}", flush=True)
    assert dte <= 0.02 and dto <= 0.02
    assert dfrc <= 2e-3
