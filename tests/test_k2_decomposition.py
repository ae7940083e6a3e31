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


def _inv(x, m):
    return next(i for i in range(1, m) if (x * i) % m == 1)


@pytest.mark.parametrize("L", [32, 48, 64])
def test_lane_dft_split_and_residue_map(L):
    """The length-L lane DFT of the pruned column pass (Y[r] = sum_b v[b]
    e^{+2 pi i b r / L}) as R1 = L/16 point DFTs then sixteen-point DFTs.  R1
    odd: prime-factor split, input n = (16 n1 + R1 n2) mod L, no twiddles,
    output k2 of group k1 at r = (16 a k1 + R1 b k2) mod L, a = 16^-1 mod R1,
    b = R1^-1 mod 16.  R1 even: Cooley-Tukey, n = 16 n1 + n2, twiddles
    e^{+2 pi i n2 k1 / L}, r = k1 + R1 k2."""
    rng = np.random.default_rng(L)
    v = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    R1 = L // 16
    direct = np.array([np.sum(v * np.exp(2j * np.pi * np.arange(L) * r / L)) for r in range(L)])
    pfa = R1 % 2 == 1
    got = np.zeros(L, complex)
    for k1 in range(R1):
        u = np.zeros(16, complex)
        for n2 in range(16):
            idx = [(16 * n1 + R1 * n2) % L if pfa else 16 * n1 + n2 for n1 in range(R1)]
            t = np.sum(v[idx] * np.exp(2j * np.pi * np.arange(R1) * k1 / R1))
            u[n2] = t if pfa else t * np.exp(2j * np.pi * n2 * k1 / L)
        y = np.array([np.sum(u * np.exp(2j * np.pi * np.arange(16) * k2 / 16)) for k2 in range(16)])
        for k2 in range(16):
            r = (16 * _inv(16 % R1, R1) * k1 + R1 * _inv(R1, 16) * k2) % L if pfa else k1 + R1 * k2
            got[r] = y[k2]
    np.testing.assert_allclose(got, direct, rtol=0, atol=1e-10 * L)
    if pfa:   # every residue is produced exactly once
        rs = {(16 * _inv(16 % R1, R1) * k1 + R1 * _inv(R1, 16) * k2) % L for k1 in range(R1) for k2 in range(16)}
        assert rs == set(range(L))
