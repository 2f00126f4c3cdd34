"""CPU tests that pin the oracle (the checker) against numpy/pocketfft golden
vectors, closed forms, brute force on tiny inputs and invariants -- each
chosen so that a dropped term, a wrong sign or index, or a transposed
operand fails one of them.  No GPU needed.
"""
import os

import numpy as np
import pytest

from _oracle import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def o():
    return Oracle("f64")


@pytest.fixture(scope="module")
def o32():
    return Oracle("f32")


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(HERE, "golden", "fft_golden.npz"))


# ------------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 12, 16, 26, 53, 100, 103, 512, 3710])
def test_fft_c2c_matches_numpy(o, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = o.fft_c2c(x, -1)
    assert np.abs(f - np.fft.fft(x)).max() <= 1e-13 * max(1, np.abs(f).max())
    b = o.fft_c2c(x, +1)
    assert np.abs(b - np.fft.ifft(x) * n).max() <= 1e-13 * max(1, np.abs(b).max())


@pytest.mark.parametrize("n", [16, 26, 32, 64])
def test_r2c2d_golden(o, golden, n):
    s = o.r2c2d(golden["img_%d" % n])
    ref = golden["rfft2_%d" % n]
    assert np.abs(s - ref).max() <= 1e-12 * np.abs(ref).max()


def test_reference_ramp_roundtrip(o, o32, golden):
    """The 26 x 26 ramp of the reference's FFT self-check: R2C matches
    pocketfft and C2R/n^2 gives the ramp back (FP64 and FP32 builds)."""
    ramp = golden["ramp_26"]
    s = o.r2c2d(ramp)
    assert np.abs(s - golden["rfft2_ramp_26"]).max() < 1e-11
    assert np.abs(o.c2r2d(s) / 26 ** 2 - ramp).max() < 1e-12
    s32 = o32.r2c2d(ramp.astype(np.float32))
    assert np.abs(o32.c2r2d(s32) / 26 ** 2 - ramp).max() < 1e-4


def test_fft_closed_forms(o):
    n = 32
    d = np.zeros((n, n))
    d[0, 0] = 1.0
    assert np.allclose(o.r2c2d(d), 1.0)                         # delta -> constant
    c = np.full((n, n), 2.5)
    s = o.r2c2d(c)
    assert abs(s[0, 0] - 2.5 * n * n) < 1e-9                    # constant -> DC delta
    assert np.abs(s.ravel()[1:]).max() < 1e-9
    y, x = np.mgrid[0:n, 0:n]
    e = np.cos(2 * np.pi * (3 * x + 5 * y) / n)                 # one cosine -> two bins
    s = o.r2c2d(e)
    assert abs(s[5, 3] - n * n / 2) < 1e-9 and abs(s[n - 5, 3]) < 1e-9
    mask = np.ones_like(s, bool)
    mask[5, 3] = False
    assert np.abs(s[mask]).max() < 1e-9


def test_parseval(o):
    n = 64
    img = np.random.default_rng(3).standard_normal((n, n))
    s = o.r2c2d(img)
    w = np.full(s.shape[1], 2.0)
    w[0] = w[-1] = 1.0
    assert abs((np.abs(s) ** 2 * w).sum() / n ** 2 - (img ** 2).sum()) < 1e-8 * (img ** 2).sum()


def test_c2r_semantics_match_numpy_irfft2(o):
    """C2R of arbitrary (non-Hermitian) half spectra equals numpy's irfft2 * n^2:
    complex column transforms, then the row C2R ignoring the imaginary parts
    of the DC and Nyquist bins."""
    n = 32
    rng = np.random.default_rng(11)
    X = rng.standard_normal((n, n // 2 + 1)) + 1j * rng.standard_normal((n, n // 2 + 1))
    assert np.abs(o.c2r2d(X) - np.fft.irfft2(X, s=(n, n)) * n * n).max() < 1e-10


# -------------------------------------------------------------- geometry
def test_verify_crop_size(o):
    v = o.lib.or_verify_crop_size
    assert v(3838, 3710, 0, 0, 0, 1.0) == 3710
    assert v(3838, 3710, 0, 0, 0, 2.0) == 3708          # multiple of 4 when binning 2x
    assert v(3838, 3710, 0, 0, 0, 1.5) == 3710
    assert v(4096, 4096, 0, 0, 1000, 1.0) == 1000
    assert v(4096, 4096, 100, 10, 0, 1.0) == 3996
    assert v(4096, 4096, 4096, 0, 0, 1.0) == 0           # offset out of range
    assert v(4096, 4096, 0, 0, 0, 0.0) == 0              # bin <= 0
    assert v(513, 600, 0, 0, 0, 1.0) == 512              # rounded to even


def test_binned_size(o):
    assert o.lib.or_binned_size(4096, 2.0) == 2048
    assert o.lib.or_binned_size(3710, 1.5) == 2472
    assert o.lib.or_binned_size(512, 1.0) == 512


def test_bin_crop_is_slicing_and_scales_means(o):
    nu, n = 64, 32
    img = np.random.default_rng(4).standard_normal((nu, nu)) + 3.0
    S = o.r2c2d(img)
    B = np.zeros((n, n + 2))
    src = o.padded(S)
    o.lib.or_bin_crop(B.ctypes.data, src.ctypes.data, n, nu)
    Bc = B.view(np.complex128)
    ref = np.concatenate([S[: n // 2, : n // 2 + 1], S[nu - n // 2:, : n // 2 + 1]])
    assert np.array_equal(Bc, ref)
    # Fourier crop keeps the DC sum: the binned image mean is bin^2 x the mean
    small = o.c2r2d(Bc) / (n * n)
    assert abs(small.mean() - img.mean() * (nu / n) ** 2) < 1e-9


def test_pad_fill_is_region_mean(o):
    rng = np.random.default_rng(9)
    frame = rng.random((40, 50)).astype(np.float32)
    out = np.zeros((64, 66))
    o.lib.or_pad2fft(frame.ctypes.data, 50, 40, out.ctypes.data, 5, 3, 64)
    region = frame[3:, 5:].astype(np.float64)
    assert np.array_equal(out[:37, :45], region)
    assert np.allclose(out[37:, :64], region.mean()) and np.allclose(out[:, 45:64], region.mean())


# ------------------------------------------------------ spectral kernels
def test_weight_closed_form(o):
    n, B = 64, 150.0
    for kx, ky in [(0, 0), (1, 0), (0, 1), (3, 5), (32, 40), (10, 63)]:
        yy = ky if ky < n // 2 else ky - n
        z = (-1) ** ((kx + yy) % 2) * np.exp(-B * (kx * kx + yy * yy) / (2 * n * n))
        assert abs(o.weight(n, B, kx, ky) - z) < 1e-15


def _brute_xcorr(a, b):
    """c(d) = sum_u a(u + d) b(u), circular, by direct summation."""
    n = a.shape[0]
    c = np.zeros((n, n))
    for dy in range(n):
        for dx in range(n):
            c[dy, dx] = (np.roll(a, (-dy, -dx), axis=(0, 1)) * b).sum()
    return c


def test_cross_power_map_is_centred_cross_correlation(o):
    """B = 0: C2R(F_a conj(F_b) z / sizec) = n^2/sizec * xcorr(a, b) moved by
    n/2 (the (-1)^{x+y'} factor), checked against direct summation."""
    n = 16
    rng = np.random.default_rng(5)
    a, b = rng.standard_normal((n, n)), rng.standard_normal((n, n))
    X = o.cross_power(o.r2c2d(a), o.r2c2d(b), n, 0.0)
    m = o.c2r2d(X)
    c = _brute_xcorr(a, b)
    sizec = n * (n // 2 + 1)
    ref = np.roll(c, (n // 2, n // 2), axis=(0, 1)) * n * n / sizec
    assert np.abs(m - ref).max() < 1e-10 * np.abs(ref).max()


def test_integer_roll_gives_measured_shift(o):
    """frame_j = frame_i moved by t: measured b_ij = t_i - t_j = -t (exact)."""
    n, box = 128, 64
    rng = np.random.default_rng(6)
    a = rng.standard_normal((n, n))
    for t in [(3, -5), (-7, 2), (0, 0), (11, 9)]:
        b = np.roll(a, (t[1], t[0]), axis=(0, 1))       # content moved by (+tx, +ty)
        X = o.cross_power(o.r2c2d(a), o.r2c2d(b), n, 150.0)
        sx, sy, cc, _ = o.find_shift(o.c2r2d(X), box)
        assert (sx, sy) == (-t[0], -t[1])


def test_pruned_box_equals_full_inverse(o):
    n = 128
    rng = np.random.default_rng(8)
    X = o.cross_power(o.r2c2d(rng.standard_normal((n, n))),
                      o.r2c2d(rng.standard_normal((n, n))), n, 150.0)
    full = o.c2r2d(X)
    for box in (32, 96):
        ori = (n - box) // 2
        pr = o.pruned_box(X, box)
        assert np.abs(pr - full[ori:ori + box, ori:ori + box]).max() < 1e-10 * np.abs(full).max()


def test_phase_shift_integer_is_roll(o):
    n = 64
    img = np.random.default_rng(10).standard_normal((n, n))
    s = o.phase_shift(o.r2c2d(img), 5.0, -3.0)
    back = o.c2r2d(s) / (n * n)
    # multiplying by e^{+i 2 pi k s/n} moves the content by -s
    assert np.abs(back - np.roll(img, (3, -5), axis=(0, 1))).max() < 1e-10


# ---------------------------------------------------------- peak search
def test_find_shift_delta(o):
    box, n = 96, 512
    m = np.zeros((box, box))
    m[40, 57] = 1.0
    sx, sy, cc = o.find_shift_box(m, n)
    assert (sx, sy) == (57 - 48, 40 - 48)


def test_find_shift_on_grid_and_subpixel(o):
    box, n = 96, 512
    y, x = np.mgrid[0:box, 0:box]
    for x0, y0 in [(50.3, 47.72), (44.1, 51.55), (48.0, 48.0), (30.9, 60.4)]:
        m = np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * 1.6 ** 2))
        sx, sy, cc = o.find_shift_box(m, n)
        assert abs(sx * 32 - round(sx * 32)) < 1e-9 and abs(sy * 32 - round(sy * 32)) < 1e-9
        assert abs(sx - (x0 - 48)) <= 1 / 64 + 0.01 and abs(sy - (y0 - 48)) <= 1 / 64 + 0.01


def test_find_shift_noise_exclusion(o):
    """wNoise excludes the central (2w+1)^2 window: a stronger centre peak is
    skipped and the off-centre one found (cufunc.cu:416,466)."""
    box, n = 96, 512
    y, x = np.mgrid[0:box, 0:box]
    m = 2.0 * np.exp(-((x - 48) ** 2 + (y - 48) ** 2) / 4.0) + np.exp(-((x - 60) ** 2 + (y - 41) ** 2) / 4.0)
    sx, sy, _ = o.find_shift_box(m, n, wnoise=-1)
    assert (sx, sy) == (0, 0)
    sx, sy, _ = o.find_shift_box(m, n, wnoise=2)
    assert abs(sx - 12) < 0.1 and abs(sy + 7) < 0.1


def _np_find_shift(m, n, wnoise=-1):
    """FindShift (cufunc.cu:396-497) restated in numpy, FP64, independently of
    the oracle's FFT: the 512 x 512 map is evaluated as the matrix product
    Re(Ey . (mask * W * rfft2(window)) . Ex).  Returns (sx, sy, cc, U, (ox, oy))."""
    box = m.shape[0]
    # integer peak: first maximum in row-major order, noise window skipped (:409-424)
    jj, ii = np.mgrid[0:box, 0:box]
    skip = (np.abs(ii - box // 2) <= wnoise) & (np.abs(jj - box // 2) <= wnoise)
    v = np.where(skip, -np.inf, m).ravel()
    sx = sy = 0
    if v.max() > -1e9:
        k = int(np.argmax(v))  # first occurrence of the maximum
        sy, sx = divmod(k, box)
    # 16 x 16 window (:428-437); clamps applied as the reference orders them
    offx, offy = sx - 8, sy - 8
    offx = box - 16 if offx + 16 > box else max(offx, 0)
    offy = box - 16 if offy + 16 > box else max(offy, 0)
    G = np.fft.rfft2(m[offy:offy + 16, offx:offx + 16])  # (16, 9)
    # cosmask2d (func.cpp:348-370): r = int(sqrt(kx^2 + y'^2)), y' = ky or ky - 16
    ky = np.arange(16)[:, None]
    yp = np.where(ky < 8, ky, ky - 16)
    kx = np.arange(9)[None, :]
    r = np.floor(np.sqrt(kx * kx + yp * yp))
    G = G * np.where(r < 8, np.cos(r * np.pi / 8) / 2 + 0.5, 0.0)
    # zero-pad to 512 (:446-453): rows 0..7 -> 0..7, rows 8..15 -> 504..511,
    # then the unnormalised C2R; column weights 1 (DC) and 2 (kx = 1..8)
    q = np.arange(512)
    Ey = np.exp(2j * np.pi * np.outer(q, yp[:, 0]) / 512)  # (512, 16)
    Ex = np.exp(2j * np.pi * np.outer(np.arange(9), q) / 512)  # (9, 512)
    a = np.where(np.arange(9) == 0, 1.0, 2.0)
    U = (Ey @ (G * a[None, :]) @ Ex).real
    # first strict maximum from U[0], scaled noise window skipped (:457-476)
    ox, oy = (box // 2 - offx) * 32, (box // 2 - offy) * 32
    qq, pp = np.mgrid[0:512, 0:512]
    skip = (np.abs(pp - ox) <= 32 * wnoise) & (np.abs(qq - oy) <= 32 * wnoise)
    u = np.where(skip, -np.inf, U).ravel()
    best, sxp, syp = U[0, 0], 0, 0
    k = int(np.argmax(u))
    if u[k] > best:
        best = u[k]
        syp, sxp = divmod(k, 512)
    return (offx + sxp / 32 - box // 2, offy + syp / 32 - box // 2,
            best / n / n / 16 / 16, U, (sxp, syp))


def _top2_margin(U, pick):
    """Relative gap between the picked maximum and the best other point."""
    u = U.copy()
    u[pick[1], pick[0]] = -np.inf
    return (U[pick[1], pick[0]] - u.max()) / abs(U[pick[1], pick[0]])


def _check_same(o, m, n, wnoise=-1):
    sx, sy, cc, U, pick = _np_find_shift(m, n, wnoise)
    osx, osy, occ = o.find_shift_box(m, n, wnoise)
    assert (osx, osy) == (sx, sy), ((osx, osy), (sx, sy))
    assert abs(occ - cc) <= 1e-12 * abs(cc) + 1e-300, (occ, cc)
    return U, pick


def test_find_shift_matches_numpy_restatement_random_maps(o):
    """Random box maps (noise only, and Gaussian peaks anywhere, including the
    border clamps of the 16 x 16 window) at boxes 32-128: identical shifts
    and cc within 1e-12."""
    rng = np.random.default_rng(2024)
    worst = 1.0
    for t in range(60):
        box = int(rng.choice([32, 48, 64, 96, 128]))
        n = int(rng.choice([256, 512, 4096]))
        y, x = np.mgrid[0:box, 0:box]
        m = rng.normal(0, 1.0, (box, box))
        if t % 3:
            x0, y0 = rng.uniform(-2, box + 1, 2)  # peaks past the edges clamp the window
            m += rng.uniform(3, 50) * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / rng.uniform(1, 12))
        U, pick = _check_same(o, m, n)
        worst = min(worst, _top2_margin(U, pick))
    print("smallest top-2 margin of the random maps: %.3g" % worst)


@pytest.mark.parametrize("wnoise", [0, 1, 3])
def test_find_shift_matches_numpy_restatement_noise_window(o, wnoise):
    """-nps > 0: the integer and the 1/32 searches both skip the (scaled)
    central window, also when the window origin is clamped."""
    rng = np.random.default_rng(77 + wnoise)
    for t in range(12):
        box = int(rng.choice([48, 96]))
        y, x = np.mgrid[0:box, 0:box]
        c = box // 2
        m = 4.0 * np.exp(-((x - c) ** 2 + (y - c) ** 2) / 3.0) + rng.normal(0, 0.05, (box, box))
        x0, y0 = c + rng.uniform(-6, 6), c + rng.uniform(-6, 6)
        m += np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / 3.0)
        _check_same(o, m, 512, wnoise)


def test_find_shift_matches_numpy_restatement_near_ties(o):
    """Maps built so that two neighbouring 1/32-grid points nearly tie: a
    Gaussian's sub-pixel position is bisected to where the pick switches
    between them, down to a relative top-2 gap of about 1e-9.  Oracle and
    restatement must still pick the same point."""
    rng = np.random.default_rng(31337)
    box, n = 64, 1024
    y, x = np.mgrid[0:box, 0:box]
    done = 0
    for t in range(8):
        noise = rng.normal(0, 0.02, (box, box))
        w = rng.uniform(1.5, 6)
        y0 = 32 + rng.uniform(-3, 3)
        g = 30 + rng.integers(-3, 3) + rng.integers(0, 32) / 32

        def pick_at(x0):
            m = np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / w) + noise
            return m, _np_find_shift(m, n)

        lo, hi = g, g + 1 / 32
        plo, phi = pick_at(lo)[1][4], pick_at(hi)[1][4]
        if plo == phi:
            continue
        for _ in range(40):
            mid = 0.5 * (lo + hi)
            if pick_at(mid)[1][4] == plo:
                lo = mid
            else:
                hi = mid
            m, r = pick_at(mid)
            gap = _top2_margin(r[3], r[4])
            if gap < 1e-9:
                break
        if gap < 1e-12:  # below FP64 evaluation noise: either pick is right
            continue
        _check_same(o, m, n)
        done += 1
        print("near tie %d: relative top-2 gap %.2g" % (t, gap))
    assert done >= 4


def test_find_shift_comparison_catches_planted_mistakes(o):
    """The comparison has power: a copy of the restatement with one planted
    mistake (no cosine mask, window origin off by one, C2R column weight
    dropped, inverse with the wrong sign, cc without the 16^2 factor)
    disagrees with the oracle on at least one of these maps."""
    rng = np.random.default_rng(5)
    box, n = 96, 512
    y, x = np.mgrid[0:box, 0:box]
    maps = [8 * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / 5.0) + rng.normal(0, 0.3, (box, box))
            for x0, y0 in rng.uniform(20, 76, (6, 2))]
    ref = [o.find_shift_box(m, n) for m in maps]

    def variant(m, mistake):
        jj = np.argmax(m)
        sy, sx = divmod(int(jj), box)
        offx = min(max(sx - (7 if mistake == "origin" else 8), 0), box - 16)
        offy = min(max(sy - 8, 0), box - 16)
        G = np.fft.rfft2(m[offy:offy + 16, offx:offx + 16])
        ky = np.arange(16)[:, None]
        yp = np.where(ky < 8, ky, ky - 16)
        r = np.floor(np.sqrt(np.arange(9)[None, :] ** 2 + yp * yp))
        if mistake != "nomask":
            G = G * np.where(r < 8, np.cos(r * np.pi / 8) / 2 + 0.5, 0.0)
        sg = -1 if mistake == "sign" else 1
        q = np.arange(512)
        Ey = np.exp(sg * 2j * np.pi * np.outer(q, yp[:, 0]) / 512)
        Ex = np.exp(sg * 2j * np.pi * np.outer(np.arange(9), q) / 512)
        a = np.ones(9) if mistake == "weight" else np.where(np.arange(9) == 0, 1.0, 2.0)
        U = (Ey @ (G * a[None, :]) @ Ex).real
        k = int(np.argmax(U))
        syp, sxp = divmod(k, 512)
        cc = U.ravel()[k] / n / n / (1 if mistake == "ccscale" else 256)
        return offx + sxp / 32 - box // 2, offy + syp / 32 - box // 2, cc

    for m, (rx, ry, rc) in zip(maps, ref):  # the faithful copy agrees
        vx, vy, vc = variant(m, None)
        assert (vx, vy) == (rx, ry) and abs(vc - rc) <= 1e-12 * abs(rc)
    for mistake in ("nomask", "origin", "weight", "sign", "ccscale"):
        caught = False
        for m, (rx, ry, rc) in zip(maps, ref):
            vx, vy, vc = variant(m, mistake)
            caught |= (vx, vy) != (rx, ry) or abs(vc - rc) > 1e-9 * abs(rc)
        assert caught, mistake


# ---------------------------------------------------------- drift model
@pytest.mark.parametrize("F,fod,P", [(10, 4, 45), (24, 11, 276), (40, 2, 253), (40, 19, 780),
                                     (2, 0, 1), (3, 0, 3), (4, 0, 3)])
def test_pair_counts(o, F, fod, P):
    p = o.pair_list(F, fod)
    assert len(p) == P
    d = max(1, min(F // 2 - fod, F // 2))
    assert np.all(p[:, 1] - p[:, 0] >= d)


def _design(pairs, F):
    A = np.zeros((len(pairs), F - 1))
    for q, (i, j) in enumerate(pairs):
        A[q, i:j] = 1
    return A


def test_lsq_matches_numpy(o):
    rng = np.random.default_rng(12)
    F = 12
    pairs = o.pair_list(F, 2)
    b = rng.standard_normal((len(pairs), 2))
    s = o.lsq(pairs, F, b)
    ref = np.linalg.lstsq(_design(pairs, F), b, rcond=None)[0]
    assert np.abs(s - ref).max() < 1e-12


def test_lsq_exact_on_consistent_data(o):
    F = 9
    t = np.random.default_rng(13).standard_normal((F, 2))
    pairs = o.pair_list(F, 4)
    b = np.array([t[i] - t[j] for i, j in pairs])
    s = o.lsq(pairs, F, b)
    assert np.abs(s - (t[:-1] - t[1:])).max() < 1e-12


def test_threshold_keeps_full_rank_and_drops_outlier(o):
    F = 10
    pairs = o.pair_list(F, 4)
    t = np.random.default_rng(14).standard_normal((F, 2))
    b = np.array([t[i] - t[j] for i, j in pairs])
    b[7] += (5.0, -4.0)  # one corrupted pair
    s = o.lsq(pairs, F, b)
    ki = o.residual(pairs, s, b)
    good = o.threshold(pairs, F, ki, 1.0)
    A = _design(pairs[good], F)
    assert np.linalg.matrix_rank(A) == F - 1
    assert 7 not in good
    assert np.all(np.diff(good) > 0)
    # kept rows: all rows under threshold, and at least F-1 rows
    assert len(good) >= F - 1
    assert set(np.where(ki < 1.0)[0]) <= set(good) or len(good) == F - 1


def test_threshold_adds_rows_until_connected(o):
    """Only nframe-1 rows pass the threshold but they leave a frame isolated:
    the next-smallest residual rows are added until the rank is full."""
    F = 4
    pairs = o.pair_list(F, 1)  # d = 1: all 6 pairs
    ki = np.array([0.1, 0.2, 5.0, 0.3, 5.0, 6.0])  # rows (0,1),(0,2),(0,3)... see below
    # pairs: (0,1) (0,2) (0,3) (1,2) (1,3) (2,3); smallest three: (0,1),(0,2),(1,2)
    # -> frame 3 disconnected; the 4th smallest (0,3) or (1,3) must be added
    good = o.threshold(pairs, F, ki, 1.0)
    A = _design(pairs[good], F)
    assert np.linalg.matrix_rank(A) == F - 1
    assert list(good) == [0, 1, 2, 3]


# ------------------------------------------------------------ whole movie
def _movie(F, n, drift, dose=20.0, seed=3):
    from motioncorr1exp_amd.synth import make_movie
    return make_movie(n, n, F, dose=dose, seed=seed, drift=drift, poisson=False)


def _rolled_movie(F, n, drift, seed):
    """Circularly shifted copies of one smooth periodic image (exact integer
    drift, so every correlation map is an exactly shifted autocorrelation)."""
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(n)
    w = 1.0 / (1.0 + (k[:, None] ** 2 + k[None, :] ** 2) / 0.05 ** 2)
    img = np.fft.ifft2(np.fft.fft2(rng.standard_normal((n, n))) * np.sqrt(w)).real
    img = 10.0 + img / img.std()
    return np.stack([np.roll(img, (int(t[1]), int(t[0])), axis=(0, 1)) for t in drift]).astype(np.float32)


def test_noise_free_integer_drift_is_exact(o):
    F, n = 8, 256
    rng = np.random.default_rng(15)
    drift = np.cumsum(rng.integers(-3, 4, size=(F, 2)), axis=0).astype(float)
    drift -= drift[0]
    frames = _rolled_movie(F, n, drift, 15)
    r = o.align_movie(frames, dict(fod=2, atm=0, iter=2, box=64, nthreads=4))
    i, j = r["pairs"][:, 0], r["pairs"][:, 1]
    assert np.abs(r["pair_b"] - (drift[i] - drift[j])).max() < 1e-9
    assert np.abs(r["frame_shift"][1:] - (drift[1:] - drift[0])).max() < 1e-9


def test_subpixel_drift_recovered(o):
    F, n = 8, 256
    rng = np.random.default_rng(16)
    drift = np.cumsum(rng.normal(0, 1.2, size=(F, 2)), axis=0)
    drift -= drift[0]
    frames, _ = _movie(F, n, drift)
    r = o.align_movie(frames, dict(fod=2, atm=1, box=64, nthreads=4), want_boxes=False)
    ref = F // 2 + 1  # RefFrame = nz/2+1
    assert np.abs(r["frame_shift"] - (drift - drift[ref])).max() < 0.05


def test_rmsd_recursion_accumulates(o):
    """rmsd is not reset between rounds: rmsd_k = (rmsd_{k-1} + sum|d|^2)/F."""
    from motioncorr1exp_amd.synth import make_movie
    F = 8
    frames, _ = make_movie(256, 256, F, dose=2.0, step=1.0, seed=17)
    r = o.align_movie(frames, dict(fod=2, iter=3, box=64, nthreads=4), want_boxes=False)
    s2 = r["shift2"]
    pos = np.vstack([[0, 0], np.cumsum(s2, axis=0)])
    rm = 0.0
    for k in range(r["niter"]):
        meas = r["iter_meas"][k]
        rm = (rm + ((pos - meas) ** 2).sum()) / F
        assert abs(rm - r["rmsd"][k]) < 1e-12
        pos = meas - meas[0]


def test_first_summed_frame_unshifted_without_atm(o):
    from motioncorr1exp_amd.synth import make_movie
    frames, _ = make_movie(128, 128, 6, dose=3.0, step=1.0, seed=18)
    r = o.align_movie(frames, dict(fod=1, atm=0, box=32, nStartSum=1, nthreads=4), want_boxes=False)
    assert np.isnan(r["frame_shift"][0]).all()
    assert (r["frame_shift"][1] == 0).all()


def test_sum_of_integer_shifts_is_roll(o):
    """Integer shifts: the aligned sum equals the sum of rolled frames."""
    F, n = 6, 128
    rng = np.random.default_rng(19)
    drift = np.cumsum(rng.integers(-2, 3, size=(F, 2)), axis=0).astype(float)
    drift -= drift[0]
    frames = _rolled_movie(F, n, drift, 19)
    r = o.align_movie(frames, dict(fod=1, atm=0, box=32, nthreads=4), want_boxes=False)
    s = np.zeros((n, n))
    for k in range(F):
        sh = r["frame_shift"][k]
        s += np.roll(frames[k].astype(np.float64), (-int(np.rint(sh[1])), -int(np.rint(sh[0]))), axis=(0, 1))
    assert np.abs(r["sum_image"] - s).max() < 1e-8 * np.abs(s).max()


def test_sharpness_gradient_central_differences(o):
    F, n = 4, 64
    rng = np.random.default_rng(20)
    spectra = np.stack([o.padded(o.r2c2d(rng.standard_normal((n, n)))) for _ in range(F)])
    s0 = rng.normal(0, 0.7, size=(F, 2))
    J, g = o.sharpness(spectra, n, 150.0, s0)
    h = 1e-3
    for f in range(F):
        for ax in range(2):
            sp, sm = s0.copy(), s0.copy()
            sp[f, ax] += h
            sm[f, ax] -= h
            fd = (o.sharpness(spectra, n, 150.0, sp, False)[0] - o.sharpness(spectra, n, 150.0, sm, False)[0]) / (2 * h)
            assert abs(fd - g[f, ax]) <= 1e-6 * max(1.0, abs(J))


# ------------------------------------------------------------ FSC table
@pytest.mark.parametrize("n,nbox,step", [(4096, 205, 10), (8192, 205, 20), (1024, 257, 2),
                                         (768, 386, 1), (512, 258, 1), (256, 130, 1)])
def test_fsc_bins_closed_form(o, n, nbox, step):
    """DFAlign.cu:214-218: step = n/400 (clamped to 1), nbox = (n/2+1)/step + 1."""
    assert o.fsc_bins(n) == (nbox, step)


def _rand_half(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, n // 2 + 1)) + 1j * rng.standard_normal((n, n // 2 + 1))


def test_fsc_closed_forms(o):
    """b = c a (c > 0) gives 1, b = -a gives -1, b = i a gives 0, in every shell
    that holds elements; shells beyond n/2+1 stay 0."""
    n = 512
    a = _rand_half(n, 40)
    nb, _ = o.fsc_bins(n)
    d = (0.3, -1.7)
    one = o.fsc(a, 2.5 * a, d)
    filled = o.fsc(a, a, d) != 0
    # step 1: r >= 1 leaves shell 0 empty; r < 257 leaves shell 257 empty
    assert filled[1: nb - 1].all() and not filled[0].any() and not filled[nb - 1].any()
    assert np.abs(one[filled] - 1).max() < 1e-12
    assert np.abs(o.fsc(a, -a, d)[filled] + 1).max() < 1e-12
    assert np.abs(o.fsc(a, 1j * a, d)).max() < 1e-12


def _delta(n, kx, y):
    a = np.zeros((n, n // 2 + 1), complex)
    a[y % n, kx] = 3.0 - 2.0j
    return a


@pytest.mark.parametrize("kx,y,d,shell,col", [
    (100, -30, (1.0, 0.0), 52, 0),     # r = 104.40, |k.d|/r = 0.958
    (100, -30, (0.0, 5.0), 52, 1),     # 0.287
    (100, -30, (0.05, 0.0), 52, 0),    # |d| <= 0.1: direction (1, 0)
    (100, -30, (-3.0, 4.0), 52, 0),    # |-60 - 24| / 104.4 = 0.805
    (100, 30, (-3.0, 4.0), 52, 1),     # |-60 + 24| / 104.4 = 0.345: sign of y counts
    (0, 1, (1.0, 0.0), 0, 1),          # r = 1: first shell
    (512, 0, (1.0, 0.0), 256, 0),      # r = 512 < n/2+1: last shell
    (0, -512, (0.0, 1.0), 256, 0),     # row n/2 is y = -n/2
    (3, 3, (1.0, 0.0), 2, 0),          # float angle 0.7071068 > cos(pi/4)
    (7, 7, (1.0, 0.0), 4, 1),          # float angle rounds to cos(pi/4): not greater
])
def test_fsc_single_element_shell_and_side(o, kx, y, d, shell, col):
    """One nonzero element lands in shell int(sqrtf(kx^2+y^2)/step) on the side
    fixed by the float angle test of DFAlign.cu:243-249 (n = 1024, step 2)."""
    n = 1024
    a = _delta(n, kx, y)
    f = o.fsc(a, a, d)
    expect = np.zeros_like(f)
    expect[shell, col] = 1.0
    assert np.array_equal(f, expect)


@pytest.mark.parametrize("kx,y", [(0, 0), (512, -512), (400, 400)])
def test_fsc_skips_dc_and_corners(o, kx, y):
    n = 1024
    a = _delta(n, kx, y)
    assert not o.fsc(a, a, (1.0, 0.0)).any()


def test_fsc_table_halves_of_a_movie(o):
    """Raw halves are the unshifted frames at alternating positions of
    [nStartSum, nEndSum]; aligned halves are the first summed frame plus every
    second frame after it versus the rest, each with its final shift; the
    direction is the total shift of the last summed frame (DFAlign.cu:808-822,
    1143-1266, 1314-1318).  Frames alternate between two different images so a
    wrong partition changes the table."""
    F, n = 7, 128
    rng = np.random.default_rng(41)
    base = _rolled_movie(2, n, np.zeros((2, 2)), 41)
    P = base[0]
    Q = np.roll(P, 1, axis=0) + 0.3 * rng.standard_normal((n, n)).astype(np.float32)
    drift = np.cumsum(rng.integers(-2, 3, size=(F, 2)), axis=0).astype(float)
    frames = np.stack([np.roll(P if k % 2 == 0 else Q, (int(drift[k, 1]), int(drift[k, 0])), axis=(0, 1))
                       for k in range(F)]).astype(np.float32)
    prm = dict(fod=1, atm=0, box=32, nStartSum=1, nEndSum=5, nthreads=4)
    r = o.align_movie(frames, prm, want_spectra=True, want_boxes=False)
    spec = r["spectra"].view(np.complex128)  # (F, n, n/2+1)
    fs = r["frame_shift"]
    tot = -fs[5]  # atm 0: frame_shift = -total shift
    raw = o.fsc(spec[1] + spec[3] + spec[5], spec[2] + spec[4], tot)

    def sh(k):
        return o.phase_shift(spec[k], fs[k, 0], fs[k, 1])
    corr = o.fsc(sh(1) + sh(3) + sh(5), sh(2) + sh(4), tot)
    assert np.abs(r["fsc"][:, :2] - raw).max() < 1e-10
    assert np.abs(r["fsc"][:, 2:] - corr).max() < 1e-10
    # the partition matters: a first-half / second-half split differs
    other = o.fsc(spec[1] + spec[2], spec[3] + spec[4] + spec[5], tot)
    assert np.abs(other - raw).max() > 1e-3


# ------------------------------------------------------- quick-look images
def test_logmod_image_positions(o):
    """func.cpp:459-489: the right half holds kx = i - n/2 of row ky = (j + n/2)
    mod n, the left half its point mirror, and row 0 copies row 1; each
    pixel is log(1 + |X| scale) (cufunc.cu:178-188)."""
    n, hn, sc = 64, 32, 1.0 / 64  # scale exact in float
    cases = [
        (5, n - 3, [(hn - 3, hn + 5), (hn + 3, hn - 5)]),
        (0, 0, [(hn, hn)]),
        (hn, 4, [(hn - 4, 0)]),        # Nyquist column: left edge only
        (7, hn - 1, [(n - 1, hn + 7), (1, hn - 7), (0, hn - 7)]),
    ]
    for kx, ky, pix in cases:
        X = np.zeros((n, hn + 1), complex)
        X[ky, kx] = 300.0 + 400.0j
        img = o.logmod_image(X, sc)
        expect = np.zeros((n, n))
        for j, i in pix:
            expect[j, i] = np.log(1 + 500.0 * sc)
        assert np.abs(img - expect).max() < 1e-12, (kx, ky)


def test_sum_display_is_cosine_on_the_display_grid(o):
    """A cosine below the crop frequency: crop to 400 and the unnormalised
    C2R give n^2 cos(2 pi (3X -/+ 5Y)/400) (DFAlign.cu:1281-1284, 73-90)."""
    n, d = 512, 400
    y, x = np.mgrid[0:n, 0:n]
    Y, X = np.mgrid[0:d, 0:d]
    for sy in (5, -5):
        img = np.cos(2 * np.pi * (3 * x + sy * y) / n)
        got = o.sum_display(np.fft.rfft2(img))
        assert np.abs(got - n * n * np.cos(2 * np.pi * (3 * X + sy * Y) / d)).max() < 1e-6 * n * n


def test_fft_display_matches_numpy(o):
    """R2C of the log-modulus image, Fourier crop, C2R and / d^2 equal numpy's
    normalised irfft2 of the cropped rfft2 (func.cpp:524-555)."""
    n, d = 512, 400
    rng = np.random.default_rng(50)
    spec = np.fft.rfft2(rng.poisson(3.0, (n, n)).astype(float))
    img = o.logmod_image(spec, 1e-4)
    F = np.fft.rfft2(img)
    crop = np.concatenate([F[: d // 2, : d // 2 + 1], F[n - d // 2:, : d // 2 + 1]])
    ref = np.fft.irfft2(crop, s=(d, d))
    got = o.fft_display(spec, 1e-4, d)
    assert np.abs(got - ref).max() < 1e-9 * np.abs(ref).max()


def test_disp_short_two_levels_and_constant(o):
    b = np.zeros(1000, np.float32)
    b[500:] = 1.0
    out = o.disp_short(b)
    assert (out[:500] == 0).all() and (out[500:] == 255).all()
    assert o.disp_short(np.full(1000, 3.0, np.float32)) is None


def test_disp_short_ramp_clips_half_a_percent(o):
    """Uniform ramp of 160000 values: 400 values per histogram bin and a 0.5%
    threshold of 800 values put the clip points two bins from either end; in
    between the mapping is linear (func.cpp:372-443)."""
    N = 160000
    v = np.arange(N, dtype=np.float32)
    out = o.disp_short(v)
    step = (N - 1) / 400
    lo, hi = 2 * step, (N - 1) - 2 * step
    ref = np.clip(np.floor((v - lo) * 255.0 / (hi - lo)), 0, 255)
    ref[v >= hi] = 255
    assert np.abs(out - ref).max() <= 1
    assert (np.diff(out.astype(int)) >= 0).all()
    # a permutation of the pixels permutes the display
    perm = np.random.default_rng(51).permutation(N)
    assert np.array_equal(o.disp_short(v[perm]), out[perm])


def test_movie_quicklook_inputs(o):
    """_RawFFT is the display of the unaligned sum of [nStartSum, nEndSum],
    _CorrSum and _CorrFFT those of the aligned sum (DFAlign.cu:842-848,
    1270-1285)."""
    from motioncorr1exp_amd.synth import make_movie
    F, n = 6, 512
    frames, _ = make_movie(n, n, F, dose=2.0, step=0.8, seed=52)
    r = o.align_movie(frames, dict(fod=1, box=32, nStartSum=1, nEndSum=4, nthreads=8),
                      want_spectra=True, want_boxes=False, want_disp=True)
    assert r["disp_written"] == 7
    spec = r["spectra"].view(np.complex128)
    fs = r["frame_shift"]
    raw = spec[1:5].sum(axis=0)
    corr = sum(o.phase_shift(spec[k], fs[k, 0], fs[k, 1]) for k in range(1, 5))

    def close(a, b):
        return np.abs(a.astype(int) - b.astype(int)).max() <= 1 and (a == b).mean() > 0.999
    assert close(r["disp_raw_fft"], o.disp_short(o.fft_display(raw).astype(np.float32)))
    assert close(r["disp_corr_fft"], o.disp_short(o.fft_display(corr).astype(np.float32)))
    assert close(r["disp_corr_sum"], o.disp_short(o.sum_display(corr).astype(np.float32)))
    # the unaligned and aligned displays differ
    assert (r["disp_raw_fft"] != r["disp_corr_fft"]).mean() > 0.01
