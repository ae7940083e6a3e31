"""CPU check of the pruned pair-transform decomposition used by the HIP K2
kernels (float64 NumPy emulation, no GPU): kx-major cropped half-spectrum
with a side row for the original +Hc/2 frequency, column weights c(kx), the
Nyquist-row combination, DC removal, a column transform to the 2m lags and a
row transform over kx.  It must equal the oracle's full-grid window."""
import numpy as np
import pytest

from oracle import zorro_oracle as O


def _window_by_decomposition(FI, FJ, Hp, Hc, Wc, m):
    Wh = Wc // 2 + 1
    rows = np.array([k if k < Hc // 2 else k + Hp - Hc for k in range(Hc)])
    SI, SJ = FI[rows, :Wh].T, FJ[rows, :Wh].T          # [kx][ky']
    sI, sJ = FI[Hc // 2, :Wh], FJ[Hc // 2, :Wh]          # original row +Hc/2
    dy = np.arange(-m, m)
    E_y = np.exp(2j * np.pi * np.arange(Hc)[:, None] * dy[None, :] / Hc)
    Z = np.zeros((Wh, 2 * m), complex)
    for kx in range(Wh):
        cw = 1.0 if (kx == 0 or 2 * kx == Wc) else 2.0
        x = cw * np.conj(SI[kx]) * SJ[kx]
        if kx == 0:
            x[0] = 0.0
        nm, ns = np.conj(SI[kx, Hc // 2]) * SJ[kx, Hc // 2], np.conj(sI[kx]) * sJ[kx]
        x[Hc // 2] = nm if kx == 0 else (ns if 2 * kx == Wc else nm + ns)
        Z[kx] = x @ E_y
    E_x = np.exp(2j * np.pi * np.arange(Wh)[:, None] * dy[None, :] / Wc)
    return np.real(Z.T @ E_x)


@pytest.mark.parametrize("Hp,Hc,Wc", [(128, 128, 128), (128, 96, 96), (128, 64, 96)])
def test_decomposition_equals_oracle_window(Hp, Hc, Wc):
    rng = np.random.default_rng(Hp + Hc + Wc)
    f = 5 + rng.standard_normal((Hp, Hp))
    g = np.roll(f, (3, -2), (0, 1)) + 0.7 * rng.standard_normal((Hp, Hp))
    F, G = np.fft.fft2(f), np.fft.fft2(g)
    m = 12
    FIc, FJc = O.fourier_crop(F, (Hc, Wc)), O.fourier_crop(G, (Hc, Wc))
    Wo, _ = O.mnxc2_window_unmasked(FIc, FJc, f.sum(), (f * f).sum(), g.sum(), (g * g).sum(), (Hp, Hp),
                                    (Hc, Wc), m, suppressOrigin=False)
    nc2 = 1.0 / (Hc * Wc) ** 2
    den = max(1.0, nc2 * np.sqrt((Hp * Hp * (g * g).sum() - g.sum() ** 2) * (Hp * Hp * (f * f).sum() - f.sum() ** 2)))
    Wd = _window_by_decomposition(F, G, Hp, Hc, Wc, m) * nc2 / den
    assert np.max(np.abs(Wd - Wo)) <= 1e-12 * np.max(np.abs(Wo))
