"""Pins of the oracle's outlier-pixel filter (``hotpix_filter``, restating
``zorro/zorro.py:2497-2662``) against independent arithmetic (CPU only).

* the PSF kernel: zero centre, unit sum, corner/edge ratio p1 (K2 value);
* the normal-CDF fit: on an integer-valued normal sample the CDF at the bin
  centres is Phi((c + 1/2 - mu)/sigma), so the fit must return mu - 1/2 and
  sigma (closed form of the binning);
* planted outliers: a singleton hot pixel is replaced by the explicit
  PSF-weighted mean of its 8 neighbours (loops, no scipy), a dead pixel likewise;
  a 2x2 hot cluster gets the mean of its non-outlier 3x3 neighbours times
  (1 - w); untouched pixels keep their value to float32 rounding;
* the cut-off fits: on a large clean Poisson stack the upper cut-off lies
  between 2.5 and 5 sigma and flags well under 1% of the pixels.
"""
import numpy as np

from oracle import zorro_oracle as O


def test_psf_kernel_k2_and_uniform():
    k = O.hotpix_psf_kernel("K2")
    assert k.dtype == np.float32 and k[1, 1] == 0.0
    assert abs(float(np.sum(k, dtype=np.float64)) - 1.0) < 1e-6
    p1 = np.float32(0.173235968)
    assert np.isclose(k[0, 0] / k[0, 1], p1, rtol=1e-6)
    u = O.hotpix_psf_kernel("uniform")
    assert np.allclose(u[np.array([[1, 1, 1], [1, 0, 1], [1, 1, 1]], bool)], 0.125)


def test_normal_cdf_fit_recovers_binned_normal():
    rng = np.random.default_rng(5)
    mu, sg = 400.0, 20.0
    s = np.round(rng.normal(mu, sg, size=400_000)).astype(np.float32)
    st = O.hotpix_stats(s, dict(maxSigma=8.0))
    # bins are [c - 1/2, c + 1/2): the CDF at the centre c is P(X <= c + 1/2)
    assert abs(st["mu"] - (mu - 0.5)) < 0.05 * sg
    assert abs(st["sigma"] - sg) < 0.05 * sg
    # the CDF ends at 1 and the bin centres are integers
    assert st["cdf"][-1] == 1.0 and np.all(st["counts"] == np.round(st["counts"]))


def _psf_mean(um, y, x, k):
    acc = 0.0
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            acc += float(k[dy + 1, dx + 1]) * float(um[y + dy, x + dx])
    return acc


def test_planted_outliers_are_replaced():
    rng = np.random.default_rng(11)
    n, h, w = 16, 96, 96
    st = rng.poisson(4.0, (n, h, w)).astype(np.float32)
    st[:, 20, 30] += 40.0          # singleton hot
    st[:, 60, 40] = 0.0            # singleton dead
    st[:, 40:42, 70:72] += 40.0    # 2x2 hot cluster
    # explicit cut-offs (the reference defaults): on a 96x96 frame the sigma-curve
    # fits are unstable (few pixels in the tails), and this test pins the replacement
    out, info = O.hotpix_filter(st, cutoffLower=-4.0, cutoffUpper=3.25)
    assert info["cutoffLower"] == -4.0 and info["cutoffUpper"] == 3.25
    assert info["guessHotpix"] >= 5 and info["guessDeadpix"] >= 1
    k = O.hotpix_psf_kernel("K2")
    um = st.sum(axis=0) / n
    for (y, x) in ((20, 30), (60, 40)):
        wgt = float(info["logisticMask"][y, x])
        assert wgt < 1e-3
        ref = wgt * st[:, y, x] + (1.0 - wgt) * _psf_mean(um, y, x, k)
        assert np.allclose(out[:, y, x], ref, rtol=1e-4, atol=1e-4)
    # cluster pixel (40, 70): mean of its non-outlier neighbours (the other 3
    # cluster pixels are outliers), times (1 - w); its own weight ~ 0
    y, x = 40, 70
    wgt = float(info["logisticMask"][y, x])
    nb = [(yy, xx) for yy in (39, 40, 41) for xx in (69, 70, 71) if not (40 <= yy <= 41 and 70 <= xx <= 71)]
    ref = (1.0 - wgt) * np.mean([um[p] for p in nb])
    assert np.allclose(out[:, y, x], wgt * st[:, y, x] + ref, rtol=1e-4, atol=1e-4)
    # an ordinary pixel is untouched (weight 1 to float32 rounding)
    assert np.allclose(out[:, 80, 10], st[:, 80, 10], rtol=1e-6)


def test_cutoff_fits_on_large_clean_stack():
    rng = np.random.default_rng(3)
    st = rng.poisson(4.0, (16, 1024, 1024)).astype(np.float32)
    _, info = O.hotpix_filter(st)
    assert 2.5 < info["cutoffUpper"] < 5.0
    assert info["guessHotpix"] < 0.01 * 1024 * 1024
