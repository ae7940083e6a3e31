"""CPU check of the identity the spectral dose path relies on (float64 NumPy
emulation of the HIP kernels' formulas, no GPU):

  FFT2(aligned)(ky,kx) = F(ky,kx) Py(ky) Px(kx) + R0(kx) + RH(kx) e^{-2 pi i ky (H-1)/H}
                         + C0(ky) + CW(ky) e^{-2 pi i kx (W-1)/W}

where aligned is the reference's Lanczos shift (banker's-rounded roll, then the
3x3 kernel with reflect edges; the oracle's apply_shifts), F the frame spectrum,
Py/Px the roll ramps times the separable kernel's circular transfer function,
and R0/RH/C0/CW the 1-D spectra of the border lines of (reflect - circular)."""
import numpy as np
import pytest

from oracle import zorro_oracle as O


def _spectral(f, t):
    H, W = f.shape
    iy, ix = int(np.round(t[0])), int(np.round(t[1]))
    rem = np.remainder(t, 1)
    conv = not (np.isclose(rem[0], 0) and np.isclose(rem[1], 0))
    K = O.lanczos_kernel(rem) if conv else np.zeros((3, 3))
    u = K.sum(axis=1)
    v = K.sum(axis=0)
    g = np.roll(np.roll(f, iy, 0), ix, 1)
    # border difference between reflect and circular 3x3 convolution
    E = np.zeros_like(f)
    if conv:
        def refl(i, n):
            return -1 - i if i < 0 else (2 * n - 1 - i if i >= n else i)
        for y in range(H):
            for x in range(W):
                if 0 < y < H - 1 and 0 < x < W - 1:
                    continue
                s = 0.0
                for a in range(3):
                    for b in range(3):
                        yr, xr = y + 1 - a, x + 1 - b
                        s += K[a, b] * (g[refl(yr, H), refl(xr, W)] - g[yr % H, xr % W])
                E[y, x] = s
    ky = np.fft.fftfreq(H) * H
    kx = np.fft.fftfreq(W) * W
    ey = lambda d: np.exp(-2j * np.pi * ky * d / H)
    ex = lambda d: np.exp(-2j * np.pi * kx * d / W)
    Py = ey(iy) * ((u[0] * ey(-1) + u[1] + u[2] * ey(1)) if conv else 1.0)
    Px = ex(ix) * ((v[0] * ex(-1) + v[1] + v[2] * ex(1)) if conv else 1.0)
    R0 = np.fft.fft(E[0, :])
    RH = np.fft.fft(E[H - 1, :])
    c0 = E[:, 0].copy(); c0[0] = c0[-1] = 0
    cw = E[:, W - 1].copy(); cw[0] = cw[-1] = 0
    C0, CW = np.fft.fft(c0), np.fft.fft(cw)
    F = np.fft.fft2(f)
    return (F * Py[:, None] * Px[None, :] + R0[None, :] + RH[None, :] * ey(H - 1)[:, None]
            + C0[:, None] + CW[:, None] * ex(W - 1)[None, :])


@pytest.mark.parametrize("t", [(2.3, -1.7), (0.0, 0.0), (-3.0, 4.0), (0.5, 1e-7), (-7.49, 2.51)])
def test_spectral_dose_identity(t):
    rng = np.random.default_rng(5)
    f = rng.poisson(3.0, (24, 24)).astype(float)
    p = O.Params(shiftMethod="lanczos", originMode=None, subPixReg=16)
    aligned, _, _ = O.apply_shifts(f[None], np.array([t]), p)
    want = np.fft.fft2(aligned[0])
    got = _spectral(f, np.array(t))
    assert np.max(np.abs(got - want)) <= 1e-10 * np.max(np.abs(want))
