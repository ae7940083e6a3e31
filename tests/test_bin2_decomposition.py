"""The polyphase form of Fourier binning by two (the GPU input stage for
Hr = 2 Hb, Wr = 2 Wb) checked in FP64 NumPy against the oracle's binStack
(zorro/zorro.py:1866-1901: full-size FFT, four-quadrant crop, 1/(Hb Wb), raw
inverse, real part).

Along an axis of length N = 2 Nh the kept frequencies [-Nh/2, Nh/2) are one
period of the length-Nh DFT, so the raw inverse of the crop is Nh x[2m] plus
the odd samples convolved (circularly, length Nh) with
s[d] = sum_k e^{i pi k (2d-1)/Nh}; Im s = (-1)^d leaves, after the real part,
one checkerboard term q (-1)^(my+mx) / (Hb Wb) in 2D.  The half-sample kernel
g = Re s comes from the closed form of the geometric sum, as the library's
table does; it is checked against the direct sum.  Tolerance: 1e-12 relative.
"""
import numpy as np
import pytest

from oracle import zorro_oracle as O


def half_sample_kernel(nh):
    K = nh // 2
    th = np.pi * (2 * np.arange(nh) - 1) / nh
    return (np.exp(-1j * th * K) * (1 - np.exp(2j * th * K)) / (1 - np.exp(1j * th))).real


def circular(o, g):
    return np.real(np.fft.ifft(np.fft.fft(o, axis=-1) * np.fft.fft(g), axis=-1))


def bin2_polyphase(x):
    h, w = x.shape
    hb, wb = h // 2, w // 2
    z = x[:, 0::2] + circular(x[:, 1::2], half_sample_kernel(wb)) / wb
    c = np.sum(x[:, 1::2] * (-1.0) ** np.arange(wb), axis=1)
    q = np.sum(c[1::2] * (-1.0) ** np.arange(hb))
    out = z[0::2] + circular(z[1::2].T, half_sample_kernel(hb)).T / hb
    return out - (-1.0) ** (np.arange(hb)[:, None] + np.arange(wb)[None, :]) * q / (hb * wb)


@pytest.mark.parametrize("nh", [4, 30, 38, 742, 3710])
def test_half_sample_kernel_closed_form(nh):
    K = nh // 2
    d = np.arange(nh)[:, None]
    direct = np.sum(np.exp(1j * np.pi * np.arange(-K, K)[None, :] * (2 * d - 1) / nh), axis=1)
    np.testing.assert_allclose(half_sample_kernel(nh), direct.real, atol=1e-9 * nh)
    np.testing.assert_allclose(direct.imag, (-1.0) ** np.arange(nh) * 1.0, atol=1e-9 * nh)


@pytest.mark.parametrize("shape", [(8, 12), (76, 60), (40, 44), (300, 260), (1916, 1484)])
def test_polyphase_binning_equals_fourier_crop(shape):
    h, w = shape
    rng = np.random.default_rng(h * 1000 + w)
    x = rng.poisson(2.0, shape) * rng.normal(1.0, 0.05, shape)
    ref, scale = O.bin_stack(x[None], (h // 2, w // 2))
    got = bin2_polyphase(x)
    assert np.allclose(scale, 2.0)
    assert np.max(np.abs(got - ref[0])) <= 1e-12 * np.max(np.abs(ref[0]))


def test_checkerboard_term_is_needed():
    # the (Nyquist_y, Nyquist_x) corner: a checkerboard over the odd-odd samples
    x = np.zeros((8, 12))
    x[1::2, 1::2] = (-1.0) ** (np.arange(4)[:, None] + np.arange(6)[None, :])
    ref, _ = O.bin_stack(x[None], (4, 6))
    got = bin2_polyphase(x)
    np.testing.assert_allclose(got, ref[0], atol=1e-12)
    hb, wb = 4, 6
    q = np.sum(x[1::2, 1::2] * (-1.0) ** (np.arange(hb)[:, None] + np.arange(wb)[None, :]))
    assert abs(q) > 1.0          # a non-zero correction on this input
