"""Pins of the FP64 oracle against independent mathematics (CPU only).

Each test states which oracle function it pins and by what: brute force on tiny
inputs, closed forms, invariants, or known-answer cases chosen so that a
plausible mistake (dropped term, sign or index error, transposed operand) fails.
The reference ships no golden vectors for this path (its only registrator test,
``test/test_zorro.py:38-75``, is disabled), so these checks are the pins.
"""
import math

import numpy as np
import pytest
import scipy.ndimage
import scipy.optimize

from oracle import zorro_oracle as O


RNG = np.random.default_rng(1234)


# ----------------------------------------------------------------- transforms
def test_raw_ifft2_closed_forms():
    # delta -> constant 1 (raw sum); constant c -> c*N at DC; Parseval
    X = np.zeros((8, 6), complex)
    X[0, 0] = 1.0
    assert np.allclose(O.raw_ifft2(X), 1.0)
    X = np.full((8, 6), 2.0 + 0j)
    y = O.raw_ifft2(X)
    assert np.isclose(y[0, 0], 96.0) and np.allclose(y.ravel()[1:], 0.0)
    x = RNG.standard_normal((8, 6))
    F = np.fft.fft2(x)
    assert np.isclose(np.sum(np.abs(F) ** 2), 48 * np.sum(x ** 2))
    # raw inverse of the forward transform returns N * x
    assert np.allclose(O.raw_ifft2(F), 48 * x)


def test_fourier_crop_is_decimation_of_bandlimited_image():
    n = 64
    F = np.zeros((n, n), complex)
    k = RNG.standard_normal((15, 15)) + 1j * RNG.standard_normal((15, 15))
    F[:8, :8] = k[:8, :8]
    F[-7:, :8] = k[8:, :8]
    F[:8, -7:] = k[:8, 8:]
    F[-7:, -7:] = k[8:, 8:]
    img = np.real(np.fft.ifft2(F))
    Fc = O.fourier_crop(np.fft.fft2(img), (32, 32))
    dec = np.real(np.fft.ifft2(Fc)) * (32 * 32) / (64 * 64)
    # the band limit is below the crop's Nyquist except for the real-part fold
    assert np.allclose(dec, img[::2, ::2], atol=1e-12)


def test_shape_cropped_rules():
    assert O.shape_cropped((0, 0), (1024, 1024)) == (1024, 1024)
    assert O.shape_cropped((3073, 3071), (4096, 4096)) == (3072, 3070)
    with pytest.raises(ValueError):
        O.shape_cropped((4098, 10), (4096, 4096))


# ------------------------------------------------------------------ mnxc2
def _brute_masked_ncc(fI, fJ, m, mshift):
    """Padfield masked NCC by direct real-space sums over circular lags."""
    n = fI.size
    H, W = fI.shape
    out = np.zeros((H, W))
    origin = None
    for dy in range(H):
        for dx in range(W):
            sJ = np.roll(np.roll(fJ, -dy, 0), -dx, 1)     # fJ(x+d)
            sm = np.roll(np.roll(m, -dy, 0), -dx, 1)      # m(x+d)
            num = np.sum(fI * sJ) / n
            ct = np.sum(fI * sm) / n
            cb = np.sum(m * sJ) / n
            mp = np.sum(m * sm) / n
            t1 = np.sum(fI * fI * sm) / n
            t2 = np.sum(m * sJ * sJ) / n
            den = max(math.sqrt((t2 - cb * cb / mp) * (t1 - ct * ct / mp)), 1.0)
            out[dy, dx] = (num - ct * cb / mp) / den
    return out


def test_mnxc2_matches_bruteforce_masked_ncc():
    H = W = 24
    mask = np.zeros((H, W))
    mask[:19, :17] = 1.0
    fI = mask * (5 + 3 * RNG.standard_normal((H, W)))
    fJ = mask * (5 + 3 * RNG.standard_normal((H, W)))
    brute = _brute_masked_ncc(fI, fJ, mask, mask)
    M, mp = O.mask_spectrum(mask > 0, (H, W))
    FI, QI = np.fft.fft2(fI), np.fft.fft2(fI * fI)
    FJ, QJ = np.fft.fft2(fJ), np.fft.fft2(fJ * fJ)
    m = 6
    W_or, origin = O.mnxc2_window(FI, QI, FJ, QJ, M, mp, (H, W), m, suppressOrigin=False)
    expect = np.fft.ifftshift(brute)[H // 2 - m:H // 2 + m, W // 2 - m:W // 2 + m]
    assert np.allclose(W_or, expect, rtol=1e-9, atol=1e-9)
    assert np.isclose(origin, brute[0, 0])


def test_mnxc2_roll_known_answer():
    # J = I rolled by d: the peak sits at lag d with C = 1 (variance > 1, no clip)
    H = W = 64
    f = 10 + 4 * RNG.standard_normal((H, W))
    d = (5, -3)
    g = np.roll(np.roll(f, d[0], 0), d[1], 1)
    F, G = np.fft.fft2(f), np.fft.fft2(g)
    m = 16
    Wn, _ = O.mnxc2_window_unmasked(F, G, f.sum(), (f * f).sum(), g.sum(), (g * g).sum(), (H, W), (H, W), m)
    iy, ix = np.unravel_index(np.argmax(Wn), Wn.shape)
    assert (iy - m, ix - m) == d
    assert np.isclose(Wn[iy, ix], 1.0, atol=1e-12)


def test_unmasked_closed_form_equals_general_formula_with_crop():
    H = W = 64
    shapeC = (48, 40)
    f = 3 + RNG.standard_normal((H, W))
    g = np.roll(f, 4, 0) + 0.5 * RNG.standard_normal((H, W))
    mask = np.ones((H, W), bool)
    M, mp = O.mask_spectrum(mask, shapeC)
    F = O.fourier_crop(np.fft.fft2(f), shapeC)
    Q = O.fourier_crop(np.fft.fft2(f * f), shapeC)
    G = O.fourier_crop(np.fft.fft2(g), shapeC)
    R = O.fourier_crop(np.fft.fft2(g * g), shapeC)
    a, oa = O.mnxc2_window(F, Q, G, R, M, mp, shapeC, 12)
    b, ob = O.mnxc2_window_unmasked(F, G, f.sum(), (f * f).sum(), g.sum(), (g * g).sum(), (H, W), shapeC, 12)
    assert np.allclose(a, b, rtol=1e-8, atol=1e-10)
    assert np.isclose(oa, ob, rtol=1e-8)


def test_origin_suppression_stencil_quirk():
    C = np.arange(64.0).reshape(8, 8)
    W, origin = O.window_from_full(C, 4, True)
    # lag (0,0) is window index (4,4); the stencil repeats C[-1,0] and C[-1,1]
    expect = 0.125 * (C[1, 0] + C[0, 1] + 2 * C[-1, 0] + C[1, 1] + 2 * C[-1, 1] + C[-1, -1])
    assert W[4, 4] == expect and origin == 0.0
    assert W[5, 4] == C[1, 0] and W[3, 3] == C[-1, -1]


# --------------------------------------------------------------- locatePeak
def test_b_filter_conv_equals_explicit_reflect_gaussian():
    C = RNG.standard_normal((80, 80))
    sig = 1024 / (math.sqrt(2) * math.pi * 150)
    got = O.b_filter(C, "conv", 150, (1024, 1024))
    r = int(4.0 * sig + 0.5)
    k = np.exp(-0.5 * (np.arange(-r, r + 1) / sig) ** 2)
    k /= k.sum()
    P = np.pad(C, r, mode="symmetric")          # scipy 'reflect' == numpy 'symmetric'
    t = np.array([[np.dot(k, P[i:i + 2 * r + 1, j + r]) for j in range(80)] for i in range(80)])
    P2 = np.pad(t, ((0, 0), (r, r)), mode="symmetric")
    expect = np.array([[np.dot(k, P2[i, j:j + 2 * r + 1]) for j in range(80)] for i in range(80)])
    assert np.allclose(got, expect, atol=1e-12)


def test_b_filter_opti_minimises_peak_contrast_objective():
    """Bmode='opti' (zorro/zorro.py:1105-1124): the width Brent's bounded search
    returns is the minimiser of std/(max - mean) over [0, 7] to within a fine
    grid's resolution, and the window comes back filtered with it."""
    yy, xx = np.mgrid[0:80, 0:80].astype(float)
    for seed, (y0, x0, w) in enumerate([(43.2, 37.9, 2.0), (20.0, 55.5, 3.5), (60.7, 21.3, 1.2)]):
        rng = np.random.default_rng(100 + seed)
        C = np.exp(-((yy - y0) ** 2 + (xx - x0) ** 2) / (2 * w * w)) + 0.3 * rng.standard_normal((80, 80))

        def f(s):
            cf = scipy.ndimage.gaussian_filter(C, s)
            return np.std(cf) / (np.max(cf) - np.mean(cf))
        grid = np.linspace(0.0, 7.0, 701)
        g = grid[int(np.argmin([f(s) for s in grid]))]
        got = O.b_filter(C, "opti", 512, (1024, 1024))
        # recover the width from the returned window: it is gaussian_filter(C, s*) for the s* found
        s_star = scipy.optimize.minimize_scalar(f, bounds=[0.0, 7.0], method="bounded",
                                                options={"maxiter": 15, "xatol": 0.01}).x
        assert np.allclose(got, scipy.ndimage.gaussian_filter(C, s_star), atol=1e-12)
        assert abs(s_star - g) <= 0.05, (s_star, g)
        assert f(s_star) <= f(g) + 1e-3 * abs(f(g))


def _brute_subpixel(patch, P=16):
    n = 16 * P
    f = np.arange(-8, 8)
    tau = np.arange(n)[:, None] - 16 * np.arange(16)[None, :]
    K = np.exp(2j * np.pi * f[None, None, :] * tau[:, :, None] / n).sum(axis=2)   # [n,16]
    Y = K @ patch @ K.T
    return np.abs(Y)


def test_subpixel_matches_bruteforce_interpolation():
    m = 40
    yy, xx = np.mgrid[0:2 * m, 0:2 * m]
    y0, x0 = m + 7.3, m - 11.71
    C = np.exp(-((yy - y0) ** 2 + (xx - x0) ** 2) / (2 * 4.0 ** 2)) + 0.01 * RNG.standard_normal((2 * m, 2 * m))
    p = O.Params(maxShift=m, Bmode="none", subPixReg=16)
    shift, corr, ps, cmax = O.locate_peak(C, p, (1024, 1024))
    patch = C[cmax[0] - 8:cmax[0] + 8, cmax[1] - 8:cmax[1] + 8]
    A = _brute_subpixel(patch)
    s = np.unravel_index(np.argmax(A), A.shape)
    rp = np.array(cmax, float) - m
    expect = rp + (np.array(s) - 128 + np.sign(rp)) / 16
    assert np.allclose(shift, expect)
    assert np.isclose(corr, A[s] / 256)


def test_subpixel_known_answer_on_bandlimited_patch():
    # A patch that is exactly a 16-periodic trig polynomial is interpolated
    # exactly: the 1/16 grid point of the true peak wins, plus the Q4 bias.
    m = 40
    yy, xx = np.mgrid[0:2 * m, 0:2 * m].astype(float)
    y0, x0 = m + 7 + 5 / 16, m - 11 - 3 / 16
    C = 3 + np.cos(2 * np.pi * (yy - y0) / 16) + np.cos(2 * np.pi * (xx - x0) / 16)
    C[(np.abs(yy - y0) > 9) | (np.abs(xx - x0) > 9)] = 0.0
    p = O.Params(maxShift=m, Bmode="none", subPixReg=16)
    shift, corr, ps, cmax = O.locate_peak(C, p, (1024, 1024))
    assert np.allclose(shift, [y0 - m + 1 / 16, x0 - m - 1 / 16])
    assert np.isclose(corr, 5.0)


def test_subpixel_edge_fallback_uses_unfiltered_integer_peak():
    m = 40
    C = np.zeros((80, 80))
    C[40, 3] = 5.0                     # inside the disc, 3 px from the window edge
    p = O.Params(maxShift=m, Bmode="none", subPixReg=16)
    shift, corr, ps, cmax = O.locate_peak(C, p, (1024, 1024))
    assert cmax == (40, 3) and np.allclose(shift, [0, -37]) and corr == 5.0


def test_peak_argmax_restricted_to_disc_and_first_occurrence():
    m = 40
    C = np.full((80, 80), -1.0)
    C[0, 0] = 100.0                    # outside the disc: masked to zero
    C[50, 50] = 2.0
    C[60, 30] = 2.0                    # tie later in C order
    p = O.Params(maxShift=m, Bmode="none", subPixReg=1)
    shift, corr, ps, cmax = O.locate_peak(C, p, (1024, 1024))
    assert cmax == (50, 50) and np.allclose(shift, [10, 10])


# ------------------------------------------------------------------- solver
def test_build_equations_tri_and_diag_intervals():
    for tri in (O.tri_indices(6, "tri")[:-1, 1:], O.tri_indices(6, "diag", diagWidth=2)[:-1, 1:]):
        A = O.build_equations(tri)
        rows = [(I, J) for I in range(5) for J in range(5) if tri[I, J]]
        assert A.shape == (len(rows), 5)
        for r, (I, j) in enumerate(rows):
            expect = np.zeros(5)
            expect[I:j + 1] = 1.0             # pair (I, J=j+1) sums x_I..x_{J-1}
            assert np.array_equal(A[r], expect)


def test_build_equations_auto_pattern_known_answer():
    """triMode 'auto' rows of different lengths: the rot90 construction
    (zorro/zorro.py:1296-1311) pairs a measurement with the reversed row of a
    LOWER row of the block, worked out by hand for 5 frames with pairs
    (0,1) (0,2) (0,3) | (1,2) | (2,3) (2,4) | (3,4)."""
    pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (2, 3), (2, 4), (3, 4)]
    tri = np.zeros((4, 4), bool)
    for I, J in pairs:
        tri[I, J - 1] = True
    A = O.build_equations(tri)
    expect = np.array([[1, 0, 0, 0],     # b(0,1): x0
                       [0, 1, 0, 0],     # b(0,2): x1 (row 1 of the block, reversed), not x0 + x1
                       [1, 1, 1, 0],     # b(0,3)
                       [0, 1, 0, 0],     # b(1,2)
                       [0, 0, 1, 0],     # b(2,3)
                       [0, 0, 1, 1],     # b(2,4)
                       [0, 0, 0, 1]])    # b(3,4)
    assert np.array_equal(A, expect)


def test_auto_stops_each_row_after_low_significance_or_automax():
    """xcorrnm2_tri's 'auto' loop (zorro/zorro.py:992-996) on a movie whose frames
    decorrelate: every row is a contiguous run from J = I+1, its last pair is the
    first with peaksig <= peaksigThres or J - I >= autoMax."""
    from zorro_amd.synth import make_movie
    n = 8
    frames, _ = make_movie(n, 256, 256, seed=5, dose=4.0)
    p = O.Params(triMode="auto", autoMax=5, peaksigThres=3.115, maxShift=20, Bmode="conv", Brad=150, subPixReg=1,
                 fouCrop=(0, 0), shapePadded=(256, 256))
    padded, mask = O.pad_stack(frames.astype(np.float64), (256, 256))
    xc = O.xcorrnm2_tri(padded, mask, p)
    ps = xc["peaksigTriMat"]
    lengths = []
    for I in range(n - 1):
        row = [J for J in range(I + 1, n) if xc["corrTriMat"][I, J] != 0]
        assert row == list(range(I + 1, I + 1 + len(row)))
        last = row[-1]
        stop = ps[I, last] <= p.peaksigThres or last - I >= p.autoMax
        assert stop or last == n - 1
        assert all(ps[I, J] > p.peaksigThres and J - I < p.autoMax for J in row[:-1])
        lengths.append(len(row))
    assert min(lengths) == 1 and max(lengths) >= 3        # rows of different lengths occur


def test_refine_known_answer_against_scaled_template():
    """triMode 'refine' (zorro/zorro.py:934-956): with frames [roll(f, d), f, f, f]
    the template of frame 0 is (sum F - F_0)/4 = 3/4 F and its squared term
    3/4 Q, so the peak sits at lag d with
    C = (3/4)(Np SQ - S^2) / sqrt((Np SQ - S^2)(3/4 Np SQ - 9/16 S^2))."""
    H = W = 64
    f = 10 + 4 * RNG.standard_normal((H, W))
    d = (3, -5)
    frames = np.stack([np.roll(np.roll(f, d[0], 0), d[1], 1), f, f, f])
    p = O.Params(triMode="refine", maxShift=16, Bmode="none", subPixReg=1, suppressOrigin=False, fouCrop=(0, 0),
                 shapePadded=(H, W))
    padded, mask = O.pad_stack(frames, (H, W))
    xc = O.xcorrnm2_tri(padded, mask, p, return_windows=True)
    assert np.allclose(xc["shiftsTriMat"][0, 0], d)
    Np, S, SQ = H * W, f.sum(), (f * f).sum()
    expect = 0.75 * (Np * SQ - S * S) / math.sqrt((Np * SQ - S * S) * (0.75 * Np * SQ - 0.5625 * S * S))
    win = xc["windows"][(0, 0)]
    assert np.isclose(win[16 + d[0], 16 + d[1]], expect, rtol=1e-9)
    assert np.all(xc["corrTriMat"][3] == 0)               # the last frame is never the base
    # the general five-term form with an all-ones mask gives the same window
    # (the closed form holds for the refine template, whose squared term is
    # not the square of its image)
    F, Q = O.frame_spectra(frames, np.ones((H, W)), (H, W))
    M, mp = O.mask_spectrum(np.ones((H, W), bool), (H, W))
    TF, TQ = (F.sum(0) - F[0]) / 4, (Q.sum(0) - Q[0]) / 4
    gen, _ = O.mnxc2_window(TF, TQ, F[0], Q[0], M, mp, (H, W), 16, suppressOrigin=False)
    assert np.allclose(gen, win, rtol=1e-7, atol=1e-9)


def test_l1_solver_exact_on_consistent_system_and_optimal_otherwise():
    N = 12
    tri = O.tri_indices(N + 1, "tri")[:-1, 1:]
    A = O.build_equations(tri)
    x_true = np.round(RNG.normal(0.8, 0.5, N) * 16) / 16
    b = A @ x_true
    w = RNG.uniform(0.1, 1, A.shape[0])
    x = O.l1_solve(A, b, w, 1e3)
    xe = O.exact_vertex(A, b, x)
    assert np.array_equal(xe, x_true)
    # noisy: the LP optimum is no worse than random restarts of a local method
    b2 = b + RNG.standard_normal(b.size) * 0.2
    xo = O.l1_solve(A, b2, w, 1e3)
    fo = np.sum(w * np.abs(A @ xo - b2))
    for s in range(5):
        r = scipy.optimize.minimize(lambda z: np.sum(w * np.abs(A @ z - b2)), RNG.standard_normal(N),
                                    method="Powell")
        assert fo <= r.fun + 1e-9


def test_logistic_fit_no_worse_than_reference_basinhopping():
    ps = np.concatenate([RNG.normal(4, 1, 300), RNG.normal(12, 2, 480)])
    ps = ps[ps > 0]
    pdf, edges = np.histogram(ps, bins=int(ps.size / 7.0))
    h = edges[:-1]
    cdf = np.cumsum(pdf.astype(float))
    cdf /= cdf[-1]
    thr, K, nu, f = O.logistic_fit(h, cdf, ps.min(), ps.max())
    assert ps.min() <= thr <= ps.max() and 0.1 <= K <= 20 and 0.05 <= nu <= 5
    assert np.isclose(f, O.logistic_objective([thr, K, nu], h, cdf))
    # the reference's optimiser (zorro/zorro.py:1998-2008), seeded
    bh = scipy.optimize.basinhopping(O.logistic_objective, [np.mean(ps), 5.0, 1.0], niter=50, seed=7,
                                     minimizer_kwargs=dict(bounds=((ps.min(), ps.max()), (0.1, 20), (0.05, 5)),
                                                           tol=1e-6, method="L-BFGS-B", args=(h, cdf)))
    assert f <= bh.fun * (1 + 1e-6) + 1e-9


# -------------------------------------------------------------- applyShifts
def test_lanczos_integer_shift_is_roll_and_probe_values():
    img = RNG.standard_normal((3, 20, 24))
    t = np.array([[2.0, -3.0], [0.0, 0.0], [-5.0, 7.0]])
    p = O.Params(shiftMethod="lanczos", originMode=None, subPixReg=16)
    out, s, _ = O.apply_shifts(img, t, p)
    for j in range(3):
        assert np.array_equal(out[j], np.roll(np.roll(img[j], int(t[j, 0]), 0), int(t[j, 1]), 1))
    # kernel centre of mass (applied sub-pixel shift) at rem=0.7 is +0.0095 px
    k = O.lanczos_kernel([0.7, 0.7])
    kx = k.sum(axis=0)
    assert abs(np.dot(kx, [-1, 0, 1]) - 0.0095) < 5e-5
    # t=0.25 applies -0.0417 px; rem -> 0 gives a half-pixel shift (the Q9 discontinuity)
    assert abs(np.dot(O.lanczos_kernel([0.25, 0.25]).sum(axis=0), [-1, 0, 1]) + 1 / 24) < 1e-12
    assert abs(np.dot(O.lanczos_kernel([1e-7, 1e-7]).sum(axis=0), [-1, 0, 1]) + 0.5) < 1e-6
    assert np.isclose(k.sum(), 1.0)


def test_lanczos_convolution_reflect_boundary():
    img = RNG.standard_normal((1, 9, 11))
    t = np.array([[0.3, 1.6]])
    p = O.Params(shiftMethod="lanczos", originMode=None, subPixReg=16)
    out, _, _ = O.apply_shifts(img, t, p)
    g = np.roll(img[0], 2, axis=1)      # round(1.6) = 2, round(0.3) = 0
    K = O.lanczos_kernel(np.remainder(t[0], 1))
    P = np.pad(g, 1, mode="symmetric")
    expect = np.zeros_like(g)
    for i in range(9):
        for j in range(11):
            expect[i, j] = sum(K[a, b] * P[i + 2 - a, j + 2 - b] for a in range(3) for b in range(3))
    assert np.allclose(out[0], expect)


def test_bankers_rounding_of_half_shifts():
    img = RNG.standard_normal((2, 8, 8))
    t = np.array([[0.5, 2.5], [-0.5, 1.5]])
    p = O.Params(shiftMethod=None, originMode=None, subPixReg=16)
    out, _, _ = O.apply_shifts(img, t, p)
    assert np.array_equal(out[0], np.roll(img[0], 2, axis=1))              # round(0.5)=0, round(2.5)=2
    assert np.array_equal(out[1], np.roll(np.roll(img[1], 0, 0), 2, 1))    # round(-0.5)=0, round(1.5)=2


# ---------------------------------------------------------------------- FRC
def test_rotmean_constant_and_linear_profile():
    r, ax = O.rotmean(np.full((64, 64), 3.0))
    assert np.allclose(r[:45], 3.0)
    yy, xx = np.mgrid[-32:32, -32:32]
    r, _ = O.rotmean(np.sqrt(yy ** 2 + xx ** 2))
    assert np.allclose(r[5:30], np.arange(5, 30), atol=0.05)


def test_frc_identical_halves_is_one_and_noise_is_small():
    a = RNG.standard_normal((700, 700))
    trans = np.zeros((4, 2))
    r, frc2, nt = O.tiled_frc(a, a.copy(), trans, box=128)
    assert nt > 0 and np.allclose(r[:64], 1.0)
    # positive means (like counts) make the DC term of every tile agree, so the
    # [0,0] normalisation is the tile count
    b = RNG.standard_normal((700, 700))
    # (a small mean: the Hann window's own spectrum correlates the low rings)
    r, _, nt = O.tiled_frc(a + 3, b + 3, trans, box=128)
    assert np.all(np.abs(r[10:64]) < 6.0 / math.sqrt(nt * 2 * math.pi * 10))


def test_frc_crop_limit_quirk():
    t = np.array([[-3.2, 7.5], [2.0, -0.4]])
    c = O.frc_crop_limits(t)
    assert list(c) == [5, 8, 4, 1]        # end crops follow min(floor(.),5), not max


# --------------------------------------------------------------- dose filter
def test_dose_filter_constant_image_keeps_dc_raw_scaled():
    n, h, w = 4, 32, 32
    imgs = np.full((n, h, w), 2.5)
    p = O.Params(pixelsize=0.1, filterMode="dose")
    fs, info = O.dose_filter(imgs, imgs.sum(0), p, (h, w))
    assert np.allclose(fs, n * 2.5 * h * w)
    assert np.isclose(info["dosePerFrame"], 2.5)


def test_dose_filter_single_cosine_follows_published_exposure_curve():
    # one frame, cosine at (ky,kx) = (0,5): filtSum = |cos| * H*W * exp(-Ne/(2 Nc(q))) * cutoff
    h = w = 64
    yy, xx = np.mgrid[0:h, 0:w]
    img = 10.0 + np.cos(2 * np.pi * 5 * xx / w)
    p = O.Params(pixelsize=0.1, filterMode="dose")
    fs, info = O.dose_filter(img[None], img, p, (h, w))
    q = 5 / (w * 1.0)                                   # 1/Angstrom at 1 A/px
    crit = 0.24499 * q ** -1.6649 + 2.8141              # Grant & Grigorieff 2015
    dose = info["accumDose"][1]
    tq = info["thresQ"][0]
    g = math.exp(-0.5 * dose / crit) * math.exp(-(q / tq) ** 32)
    expect = np.abs(10.0 + g * np.cos(2 * np.pi * 5 * xx / w)) * h * w
    assert np.allclose(fs, expect, rtol=1e-10)


# ------------------------------------------------------------- gain + binning
def test_bin_stack_band_limited_image_is_decimated_and_constant_scaled():
    """binStack('fourier') (zorro/zorro.py:1866-1901) by exactly 2 on non-smooth
    sizes: an image whose spectrum lies inside the kept box comes back as 4 x its
    even samples; a constant c comes back as 4c (the raw inverse at the binned size)."""
    H, W = 2 * 19, 2 * 21
    F = np.zeros((H, W), complex)
    k = RNG.standard_normal((9, 9)) + 1j * RNG.standard_normal((9, 9))
    F[:5, :5] = k[:5, :5]
    F[-4:, :5] = k[5:, :5]
    F[:5, -4:] = k[:5, 5:]
    F[-4:, -4:] = k[5:, 5:]
    img = np.real(np.fft.ifft2(F))
    out, scale = O.bin_stack(img[None], (H // 2, W // 2))
    assert np.allclose(out[0], 4 * img[::2, ::2], atol=1e-12) and np.allclose(scale, [2, 2])
    out, _ = O.bin_stack(np.full((1, H, W), 2.5), (H // 2, W // 2))
    assert np.allclose(out, 10.0)


def test_bin_stack_matches_bruteforce_dft_crop():
    H, W, hb, wb = 14, 10, 6, 4
    img = RNG.standard_normal((H, W))
    out, _ = O.bin_stack(img[None], (hb, wb))
    ky = np.r_[np.arange(hb // 2), np.arange(H - hb // 2, H)]
    kx = np.r_[np.arange(wb // 2), np.arange(W - wb // 2, W)]
    y, x = np.arange(H), np.arange(W)
    Fk = np.exp(-2j * np.pi * np.outer(ky, y) / H) @ img @ np.exp(-2j * np.pi * np.outer(x, kx) / W)
    yb, xb = np.arange(hb), np.arange(wb)
    # the cropped grid: index i of the binned spectrum carries frequency ky[i]; inverse at the binned size
    fy = np.r_[np.arange(hb // 2), np.arange(-(hb // 2), 0)]
    fx = np.r_[np.arange(wb // 2), np.arange(-(wb // 2), 0)]
    brute = np.real(np.exp(2j * np.pi * np.outer(yb, fy) / hb) @ Fk @ np.exp(2j * np.pi * np.outer(fx, xb) / wb))
    assert np.allclose(out[0], brute / (hb * wb), atol=1e-12)


def test_gain_orientation_and_product():
    g = np.arange(12, dtype=float).reshape(3, 4)
    assert np.array_equal(O.orient_gain(g, {"Diagonal": True}), np.flipud(g.T))
    assert np.array_equal(O.orient_gain(g, {"Horizontal": True, "Vertical": True}), g[::-1, ::-1])
    assert np.array_equal(O.orient_gain(g, {"Horizontal": True}), g[:, ::-1])
    assert np.array_equal(O.orient_gain(g, {"Vertical": True}), g[::-1, :])
    assert np.array_equal(O.orient_gain(g, {"Diagonal": True, "Vertical": True}), np.flipud(g.T)[::-1, :])
    raw = np.array([[[0, 1, 2, 3]] * 3], np.uint8)
    gr = np.full((3, 4), 1.5, np.float32)
    assert np.array_equal(O.apply_gain(raw, gr, 2.0), raw.astype(np.float32) * 1.5 / 2)


def test_preshift_rolled_disc_known_answer():
    # whole 64x64 map with a band-limited peak at lag (+14, -9) and a higher one
    # at lag (-2, +1): the unrolled disc (radius 8) finds the near peak, a disc
    # rolled by (+13, -10) finds the far one at its exact lag
    n, m = 64, 8
    yy, xx = np.meshgrid(np.arange(n) - n // 2, np.arange(n) - n // 2, indexing="ij")
    C = 0.6 * np.exp(-((yy - 14) ** 2 + (xx + 9) ** 2) / 4.0) + 0.9 * np.exp(-((yy + 2) ** 2 + (xx - 1) ** 2) / 4.0)
    p = O.Params(maxShift=m, Bmode=None, Brad=0, subPixReg=1)
    s0, _, _, c0 = O.locate_peak(C, p, (n, n), roll=(0, 0))
    assert tuple(s0) == (-2.0, 1.0)
    s1, _, ps1, c1 = O.locate_peak(C, p, (n, n), roll=(13, -10))
    assert tuple(s1) == (14.0, -9.0)
    # the statistics run over the rolled disc: brute force
    disc = ((yy - 13) ** 2 + (xx + 10) ** 2) < m * m
    vals = C[disc]
    assert np.isclose(ps1, (C[c1] - vals.mean()) / vals.std())


# ------------------------------------------------------- solver error analysis
def _planted_solver_case(N, I0, J0, r, weights=None):
    """Consistent integer positions, all tri pairs measured exactly except pair
    (I0, J0) whose x shift is off by r; 'unweighted' or given weights."""
    pos = np.cumsum(np.concatenate([[0.0], np.arange(1, N) % 3 + 1.0]))     # x positions, exact in float32
    st = np.zeros((N, N, 2), np.float32)
    for I in range(N):
        for J in range(I + 1, N):
            st[I, J, 1] = pos[J] - pos[I]
            st[I, J, 0] = -(pos[J] - pos[I]) / 2
    st[I0, J0, 1] += r
    corr = np.triu(np.ones((N, N), np.float32), 1)
    ps = np.triu(np.full((N, N), 10.0, np.float32), 1) if weights is None else weights
    return pos, st, corr, ps


def test_solver_error_analysis_consistent_system_is_error_free():
    # zorro/zorro.py:1483-1547 on a consistent system: every error is zero
    N = 6
    pos, st, corr, ps = _planted_solver_case(N, 1, 3, 0.0)
    ed = O.shifts_solver(st, corr, ps, O.Params(weightMode="unweighted", maxShift=40), (256, 256))
    assert np.allclose(ed["relativeEst"][:, 1], np.diff(pos))
    assert np.all(ed["errorXY"] == 0) and ed["mean_errorNorm"] == 0 and ed["std_errorNorm"] == 0
    assert np.all(ed["errorX"] == 0) and np.all(ed["errorY"] == 0) and np.all(ed["errorUnraveled"] == 0)


@pytest.mark.parametrize("I0,J0", [(1, 3), (0, 5), (2, 4)])
def test_solver_error_analysis_planted_equation_known_answer(I0, J0):
    """One bad equation (I0, J0), residual r, uniform weights w = 1: the L1
    optimum keeps the true positions, so errorXY has the single row -r (A x - b),
    errorNorm = |r| there, mean = r/M, std = sqrt(r^2/M - mean^2), and the
    per-frame errorX[J] is the weighted mean of |errorX| over the pairs (J, c)
    and the pairs (c, J-1) (the reference's column J-1, which wraps to the last
    column for J = 0), counted here by walking the pair list."""
    N, r = 6, 0.75
    pos, st, corr, ps = _planted_solver_case(N, I0, J0, r)
    ed = O.shifts_solver(st, corr, ps, O.Params(weightMode="unweighted", maxShift=40), (256, 256))
    assert np.allclose(ed["relativeEst"][:, 1], np.diff(pos))
    pairs = [(I, J) for I in range(N - 1) for J in range(I + 1, N)]          # row-major trimat order
    M = len(pairs)
    e = pairs.index((I0, J0))
    exy = np.zeros((M, 2))
    exy[e, 1] = -r
    assert np.allclose(ed["errorXY"], exy, atol=1e-12)
    mean = r / M
    assert np.isclose(ed["mean_errorNorm"], mean)
    assert np.isclose(ed["std_errorNorm"], math.sqrt(r * r / M - mean * mean))
    # errorUnraveled[I, J-1] holds pair (I, J)
    eun = np.zeros((N - 1, N - 1))
    eun[I0, J0 - 1] = r
    assert np.allclose(ed["errorUnraveled"], eun)
    for J in range(N):
        row = [(I, K) for (I, K) in pairs if I == J]                         # padded row J
        col_frame = (J - 1) % N                                              # padded column J-1 (wraps)
        col = [(I, K) for (I, K) in pairs if K == col_frame]
        den = len(row) + len(col)
        num = r * (((I0, J0) in row) + ((I0, J0) in col))
        assert np.isclose(ed["errorX"][J], num / den), (J, ed["errorX"][J], num / den)
        assert ed["errorY"][J] == 0


# ------------------------------------------------------------- Hann window
def test_hann_apodization_closed_form_values():
    # zorro/zorro_util.py:62-104 'hann': 0.5 (1 + cos(pi r / R)) inside r <= R, 0 outside,
    # with the mesh arange(-box/2, box/2) (the centre at index box/2)
    box = 16
    w = O.hann_apodization([box, box])
    c = box // 2
    assert w.shape == (box, box)
    assert w[c, c] == 1.0                                             # r = 0
    assert np.isclose(w[c, c + 2], (2 + math.sqrt(2)) / 4)            # r = R/4: cos(pi/4)
    assert np.isclose(w[c + 4, c], 0.5)                               # r = R/2
    assert np.isclose(w[c - 4, c - 4], 0.5 * (1 + math.cos(math.pi * math.sqrt(32) / 8)))
    assert w[0, c] == 0.0                                             # r = R (on the circle)
    assert w[0, 0] == 0.0 and w[1, 1] == 0.0                          # outside the circle
    assert np.allclose(w, w.T)                                        # radial
    assert np.all(w[c, c:] >= 0) and np.all(np.diff(w[c, c:]) < 0)   # decreasing from the centre


# ----------------------------------------------------- logistic histogram / CDF
def test_cdf_logistic_curve_hand_counted_histogram():
    # zorro/zorro.py:1988-1998: positive peak significances only, int(n/7) bins,
    # float32 counts, normalised cumulative sum, left bin edges
    ps = np.zeros((8, 8), np.float32)
    vals = [1.0] * 10 + [2.0] * 5 + [4.0] * 6                 # 21 positive values -> 3 bins over [1, 4]
    iu = np.triu_indices(8, 1)
    ps[iu[0][:21], iu[1][:21]] = vals
    ps[iu[0][21], iu[1][21]] = -3.0                           # negatives and zeros are left out
    fit = O.cdf_logistic_curve(ps)
    assert fit["pdfPeaks"].dtype == np.float32
    assert list(fit["pdfPeaks"]) == [10.0, 5.0, 6.0]          # [1,2) [2,3) [3,4]
    assert np.allclose(fit["cdfPeaks"], [10 / 21, 15 / 21, 1.0])
    assert list(fit["hSigma"]) == [1.0, 2.0, 3.0]
    assert 1.0 <= fit["peaksigThres"] <= 4.0


# ------------------------------------------------------- dosenorm / gausslp
def test_dose_filter_normalize_constant_image_divides_dc_by_n_plus_1():
    # at q = 0 every frame's filter is 1 (infinite critical dose, no cut-off), so the
    # normalisation 1/(sum + 1) is 1/(N + 1) there (zorro/zorro.py:2472-2486)
    n, h, w = 4, 32, 32
    imgs = np.full((n, h, w), 2.5)
    p = O.Params(pixelsize=0.1, filterMode="dosenorm")
    fs0, _ = O.dose_filter(imgs, imgs.sum(0), p, (h, w))
    fs1, _ = O.dose_filter(imgs, imgs.sum(0), p, (h, w), normalize=True)
    assert np.allclose(fs0, n * 2.5 * h * w)
    assert np.allclose(fs1, n * 2.5 * h * w / (n + 1))


def test_dose_filter_normalize_cosine_gain_is_g_over_g_plus_1():
    # one frame c + a cos(k x): the plain filter passes the cosine with gain g
    # (read off the plain result), the normalised one with g / (g + 1) and the
    # DC term with 1/2: a relation between the two modes, not a restated formula
    h = w = 64
    yy, xx = np.mgrid[0:h, 0:w]
    c, a = 10.0, 1.0
    img = c + a * np.cos(2 * np.pi * 5 * xx / w)
    p = O.Params(pixelsize=0.1, filterMode="dose")
    fs0, _ = O.dose_filter(img[None], img, p, (h, w))
    fs1, _ = O.dose_filter(img[None], img, p, (h, w), normalize=True)
    g = (fs0.max() - fs0.min()) / (2 * a * h * w)
    assert 0.01 < g < 1.0
    assert np.isclose(fs1.mean(), c * h * w / 2)
    assert np.isclose((fs1.max() - fs1.min()) / (2 * a * h * w), g / (g + 1))


def test_gauss_lowpass_equals_explicit_reflect_separable_sum():
    # scipy.ndimage.gaussian_filter(img, 3.0): truncate 4 -> radius 12, 'reflect'
    # = numpy 'symmetric' padding, normalised taps exp(-x^2/18)
    img = RNG.standard_normal((40, 33))
    x = np.arange(-12, 13)
    k = np.exp(-x * x / 18.0)
    k /= k.sum()
    pad = np.pad(img, 12, mode="symmetric")
    tmp = sum(k[i] * pad[i:i + 40, :] for i in range(25))
    ref = sum(k[i] * tmp[:, i:i + 33] for i in range(25))
    assert np.allclose(O.gauss_lowpass(img), ref, atol=1e-12)
    assert np.allclose(O.gauss_lowpass(np.full((30, 30), 7.0)), 7.0)


def test_filter_steps_follow_token_order():
    assert O.filter_steps("Dose, gaussLP") == ["dose", "gausslp"]
    assert O.filter_steps("hot,dosenorm") == ["dosenorm"]
    assert O.filter_steps(None) == []
    assert O.filter_steps("dose,background, gausslp") == ["dose", "background", "gausslp"]


# ------------------------------------------------------------- Fourier shifts
def test_fourier_shift_integer_is_roll_and_padded_roll():
    img = RNG.standard_normal((24, 20))
    got = O.fourier_shift(img, (3.0, -5.0), (24, 20))
    assert np.allclose(got, np.roll(np.roll(img, 3, axis=0), -5, axis=1), atol=1e-12)
    # padded: a circular shift of the zero-padded frame, cropped back
    pad = np.zeros((32, 32))
    pad[:24, :20] = img
    got = O.fourier_shift(img, (2.0, 4.0), (32, 32))
    assert np.allclose(got, np.roll(np.roll(pad, 2, axis=0), 4, axis=1)[:24, :20], atol=1e-12)


def test_fourier_shift_round_trip_and_cosine_closed_form():
    h = w = 32
    yy, xx = np.mgrid[0:h, 0:w]
    # band-limited (no Nyquist content): shift then unshift is the identity
    F = np.zeros((h, w), complex)
    F[1:6, 1:6] = RNG.standard_normal((5, 5)) + 1j * RNG.standard_normal((5, 5))
    img = np.real(np.fft.ifft2(F))
    t = np.array([0.37, -1.62])
    back = O.fourier_shift(O.fourier_shift(img, t, (h, w)), -t, (h, w))
    assert np.allclose(back, img, atol=1e-12)
    # a fractional shift of a cosine: cos(2 pi k (x - t) / w)
    img = np.cos(2 * np.pi * 3 * xx / w) + 2.0
    got = O.fourier_shift(img, (0.0, 0.4), (h, w))
    assert np.allclose(got, np.cos(2 * np.pi * 3 * (xx - 0.4) / w) + 2.0, atol=1e-12)


def test_fourier_shift_nyquist_row_takes_cos_pi_t():
    # the fftshifted mesh puts the Nyquist frequency at -1/2: (-1)^y shifted by ty
    # keeps the real part cos(pi ty) (-1)^y
    h = w = 16
    yy, xx = np.mgrid[0:h, 0:w]
    img = (-1.0) ** yy
    got = O.fourier_shift(img, (0.3, 0.0), (h, w))
    assert np.allclose(got, np.cos(np.pi * 0.3) * img, atol=1e-12)


def _bg_model(q, x, y):
    xc, yc = x - q[1], y - q[2]
    return q[0] + q[3] * np.exp(-xc * xc / q[4] ** 2 - yc * yc / q[5] ** 2)


def test_background_estimate_keeps_a_drawn_surface():
    """backgroundEstimate (zorro/zorro_util.py:497-549) on a square frame whose
    content is a 2D Gaussian in the fit's grid units, (x - W/2)/128, started at
    that surface: the L1 error there is only the sigma-4 low-pass's (< 1e-4 of the
    amplitude), Powell keeps it, and the full-grid surface reproduces the frame.
    A wrong sample crop, grid centre or evaluation grid moves the fit away."""
    h = w = 2048
    q = (40.0, 1.3, -0.7, 4.0, 4.0, 6.0)
    y, x = np.mgrid[:h, :w]
    img = _bg_model(q, (x - w / 2) / 128.0, (y - h / 2) / 128.0).astype(np.float32)
    back, fit = O.background_estimate(img, p_start=q)
    assert fit.success
    assert np.max(np.abs(back - img)) < 2e-3 * q[3], np.max(np.abs(back - img))


def test_background_estimate_y_centring_by_x_count():
    """On a W > H frame the fit centres y by the x sample count (ycrop.shape[1],
    zorro/zorro_util.py:512) but the surface is evaluated at (y - H/2)/128: a
    surface drawn in the fit's coordinates comes back shifted in y by
    (nx - ny)/2 sample steps."""
    h, w = 1024, 2048
    nx, ny = w // 128, h // 128
    q = (40.0, 0.5, -1.0, 4.0, 4.0, 5.0)
    y, x = np.mgrid[:h, :w]
    img = _bg_model(q, x / 128.0 - nx / 2.0, y / 128.0 - nx / 2.0).astype(np.float32)
    back, fit = O.background_estimate(img, p_start=q)
    assert fit.success
    shifted = _bg_model(q, x / 128.0 - nx / 2.0, y / 128.0 - ny / 2.0)
    assert np.max(np.abs(back - shifted)) < 2e-3 * q[3]
    assert np.max(np.abs(back - img)) > 0.05 * q[3]      # the quirk is visible


def test_background_estimate_of_a_constant_fails_the_fit():
    """A constant frame has std 0, so the start's widths are 0 and the error is
    NaN: the fit reports failure and the background is zero (filtSum unchanged)."""
    img = np.full((1024, 1024), 3.25, np.float32)
    back, fit = O.background_estimate(img)
    assert not fit.success and np.all(back == 0.0)
