"""Closed-form and invariant pins of the oracle stages that the reference holds
no vectors for: trajectory smoothing, patch placement, the feature map and the
whole global/patch path on movies whose answer is known exactly."""
import numpy as np
import pytest

import oracle as O


# ---- CStackShift::Smooth (Align/CStackShift.cpp:180-198) ----
def test_smooth_keeps_linear_and_ends():
    t = np.arange(9.0)
    x, y = O.smooth(3 * t - 2, -0.5 * t + 7, 0.5)
    assert np.allclose(x, 3 * t - 2, atol=1e-12)
    assert np.allclose(y, -0.5 * t + 7, atol=1e-12)


@pytest.mark.parametrize("w", [0.5, 0.7])
def test_smooth_impulse_response(w):
    n, k = 9, 4
    sx = np.zeros(n)
    sx[k] = 1.0
    sy = np.zeros(n)
    sy[0] = 5.0  # an end point is never changed ...
    x, y = O.smooth(sx, sy, w)
    want = np.zeros(n)
    want[k] = w
    want[k - 1] = want[k + 1] = 0.5 * (1 - w)
    assert np.allclose(x, want, atol=1e-12)
    # ... but its neighbour sees it
    assert y[0] == 5.0 and abs(y[1] - 0.5 * (1 - w) * 5.0) < 1e-12


def test_smooth_weight_one_is_identity():
    rng = np.random.default_rng(0)
    a, b = rng.normal(size=7), rng.normal(size=7)
    x, y = O.smooth(a, b, 1.0)
    assert np.array_equal(x, a) and np.array_equal(y, b)


# ---- CDetectFeatures::mFindCenters + CPatchCenters (Align/CDetectFeatures.cpp:289-396,
#      Align/CPatchCenters.cpp:42-105) ----
def _grid_starts(px, py, xw, xh):
    """Every grid centre is a feature: centres stay at (i + 0.5) * 128 / p in
    map pixels, scaled by x/128, start = (int)(centre - size/2), clamped."""
    psx, psy = xw // px // 2 * 2, xh // py // 2 * 2
    out = []
    for j in range(py):
        for i in range(px):
            cx = (i + 0.5) * np.float32(128.0 / px) * np.float32(xw / 128.0)
            cy = (j + 0.5) * np.float32(128.0 / py) * np.float32(xh / 128.0)
            sx = min(max(int(cx - psx // 2), 0), xw - psx)
            sy = min(max(int(cy - psy // 2), 0), xh - psy)
            out.append((sx, sy))
    return np.array(out)


@pytest.mark.parametrize("px,py,xw,xh", [(5, 5, 2882, 2048), (7, 5, 2882, 2048),
                                         (5, 3, 1024, 1536)])
def test_patch_starts_all_features(px, py, xw, xh):
    st = O.patch_starts(np.ones((128, 128), np.uint8), px, py, xw, xh)
    assert np.array_equal(st, _grid_starts(px, py, xw, xh))


def test_patch_starts_single_feature_taken_by_first_patch():
    """One feature pixel at map (x=70, y=60).  Patch (2,2) (grid centre (64,64)
    is no feature) searches the 42x42 window from (43,43) and takes the pixel
    centre (70.5, 60.5); that marks a 20x20 area used, so patch (3,2), whose
    window also holds the pixel, keeps its grid centre.  Every other patch
    keeps its grid centre."""
    px = py = 5
    xw, xh = 2560, 1280
    f = np.zeros((128, 128), np.uint8)
    f[60, 70] = 1
    st = O.patch_starts(f, px, py, xw, xh)
    want = _grid_starts(px, py, xw, xh)
    psx, psy = xw // px // 2 * 2, xh // py // 2 * 2
    want[12] = (int(70.5 * (xw / 128.0) - psx // 2), int(60.5 * (xh / 128.0) - psy // 2))
    assert np.array_equal(st, want)


def test_patch_starts_tie_goes_to_the_first_feature_in_scan_order():
    """3 x 3 patches: patch 1's grid centre is x = 1.5 * (128 / 3) = 64.0 in
    FP32 (64.0000019 in FP64), y = 21.33, and its cell holds no feature.
    Two features in row 22 at x = 63 and x = 64 (centres 63.5 and 64.5) are
    at the same FP32 distance from it: the strict < of mFindCenter keeps the
    first one scanned, x = 63.5 (Align/CDetectFeatures.cpp:337-361, all
    FP32).  An FP64 evaluation would take 64.5 (start 346)."""
    px = py = 3
    xw = xh = 1024
    f = np.zeros((128, 128), np.uint8)
    f[22, 63] = f[22, 64] = 1
    st = O.patch_starts(f, px, py, xw, xh)
    want = _grid_starts(px, py, xw, xh)
    want[1] = (338, 10)     # (63.5 * 8 - 170, 22.5 * 8 - 170)
    assert np.array_equal(st, want)


# ---- CDetectFeatures feature map (Align/CDetectFeatures.cpp:58-186) ----
def _xcf_sum(seed, W=512, H=384):
    rng = np.random.default_rng(seed)
    img = rng.normal(size=(H, W))
    k = np.fft.rfft2(img)
    ky = np.fft.fftfreq(H)[:, None]
    kx = np.fft.rfftfreq(W)[None, :]
    k *= np.exp(-(kx ** 2 + ky ** 2) * 400.0)  # structure on the 128-map scale
    return k / (W * H), W, H


def test_feature_map_ignores_dc():
    s, W, H = _xcf_sum(1)
    f0 = O.feature_map(s, W, H)
    s2 = s.copy()
    s2[0, 0] += 37.0  # adds a constant to the image: the +(1-min) offset removes it
    assert 0 < f0.sum() < 128 * 128
    assert np.array_equal(O.feature_map(s2, W, H), f0)


def test_feature_map_follows_image_translation():
    """A circular shift of the image by (+16 map px in x, -8 in y) moves the
    map by the same amount away from the (non-periodic) 5x5 border and the
    clipped-statistics region; transposed axes or a sign slip would not."""
    s, W, H = _xcf_sum(2)
    dx, dy = 16 * W // 128, -8 * H // 128
    ky = np.fft.fftfreq(H)[:, None]
    kx = np.fft.rfftfreq(W)[None, :]
    s2 = s * np.exp(-2j * np.pi * (kx * dx + ky * dy))
    f0 = O.feature_map(s, W, H).astype(int)
    f1 = O.feature_map(s2, W, H).astype(int)
    moved = np.roll(np.roll(f0, 16, axis=1), -8, axis=0)
    inner = (slice(12, 116), slice(20, 108))
    agree = np.mean(moved[inner] == f1[inner])
    assert agree > 0.97, agree


# ---- whole path on movies with a known answer ----
def _drift_movie(W, H, N, dxy, seed=5):
    rng = np.random.default_rng(seed)
    img = rng.normal(size=(H, W))
    k = np.fft.rfft2(img)
    ky = np.fft.fftfreq(H)[:, None]
    kx = np.fft.rfftfreq(W)[None, :]
    img = np.fft.irfft2(k * np.exp(-(kx ** 2 + ky ** 2) * 60.0), s=(H, W))
    img = 50.0 + 10.0 * img / img.std()
    return np.stack([np.roll(img, (int(dxy[i, 1]), int(dxy[i, 0])), axis=(0, 1))
                     for i in range(N)]).astype(np.float32)


def test_global_path_recovers_integer_linear_drift():
    """Frames are exact circular integer shifts along a line: correlation
    peaks are exact, the spline in sqrt(t) (no knots at N=8) holds c*t
    exactly, the smoothing keeps linear data, so the logged shifts equal
    minus the drift up to a constant (full-frame shifts are not made
    relative to a frame) and the 0.01 px truncation."""
    N = 8
    t = np.arange(N)
    d = np.stack([2 * t - 5, -3 * t + 4], axis=1).astype(float)
    mv = _drift_movie(256, 256, N, d)
    r = O.align_movie(mv, O.default_params(bft=(500, 100), iter=7))
    g = r["global_shift"]
    e = (g - g.mean(0)) + (d - d.mean(0))
    assert np.max(np.abs(e)) < 0.02, e
    # the sum is every frame moved by its logged (truncated) shift: content at
    # x goes to x + s, F(k) * exp(-2 pi i k.s / n)
    H, W = mv.shape[1:]
    ky = np.fft.fftfreq(H)[:, None]
    kx = np.fft.rfftfreq(W)[None, :]
    want = sum(np.fft.irfft2(np.fft.rfft2(mv[i].astype(np.float64)) *
                             np.exp(-2j * np.pi * (kx * g[i, 0] + ky * g[i, 1])), s=(H, W))
               for i in range(N))
    assert np.max(np.abs(r["sums"][0] - want)) < 1e-4 * np.max(np.abs(want))


def test_patch_path_rigid_drift_has_no_local_motion():
    N = 8
    t = np.arange(N)
    d = np.stack([t - 3, 2 * t - 7], axis=1).astype(float)
    mv = _drift_movie(512, 512, N, d, seed=7)
    p = O.default_params(patch=(2, 2), bft=(500, 100), iter=7)
    r = O.align_movie(mv, p)
    g = r["global_shift"]
    e = (g - g.mean(0)) + (d - d.mean(0))
    assert np.max(np.abs(e)) < 0.02
    assert np.max(np.abs(r["patch_shift"])) < 0.05
    assert not r["patch_bad"].any()


def test_sums_are_linear_and_clamped():
    """With replayed zero shifts the plain sum is the clamped sum of the
    frames; doubling the movie doubles every sum (dose weights normalised)."""
    rng = np.random.default_rng(11)
    N, H, W = 6, 96, 128
    mv = rng.normal(3.0, 1.0, size=(N, H, W)).astype(np.float32)
    p = O.default_params(fm_dose=1.5, pix_size=1.0, kv=300)
    z = np.zeros((N, 2))
    r1 = O.align_movie(mv, p, replay_global=z)
    r2 = O.align_movie(2 * mv, p, replay_global=z)
    assert r1["sums"].shape[0] == 3
    assert np.allclose(r1["sums"][0], np.maximum(mv.astype(np.float64).sum(0), 0), atol=1e-6)
    assert np.allclose(r2["sums"], 2 * r1["sums"], rtol=1e-9, atol=1e-9)
    # dose weighting leaves the DC (the frame mean) alone
    assert abs(r1["sums"][1].mean() - r1["sums"][0].mean()) < 1e-9 * abs(r1["sums"][0].mean())


# ---- GWeightFrame (Correct/GWeightFrame.cu:19-30, 101-109) ----
# Literal weights exp(-0.5 D / Dc(k)) with Dc = 0.24499 k^-1.6649 + 2.8141 at
# 300 kV (the Grant & Grigorieff critical-exposure fit), k in 1/A.  The
# oracle's normaliser with one frame is that frame's weight itself.
@pytest.mark.parametrize("kx,y,pix,dose,want", [
    (16, 0, 1.0, 10.0, 0.38773409603823816),    # k = 0.25
    (32, 0, 1.0, 10.0, 0.24848023943515027),    # k = 0.5, x Nyquist: fx = kx / (2 (cx - 1))
    (0, 32, 1.0, 10.0, 0.24848023943515027),    # y Nyquist: fy = 0.5 -> -0.5
    (16, 40, 1.0, 10.0, 0.2624264636553408),    # fy = 40/64 - 1 = -0.375
    (16, 24, 1.0, 37.5, 0.006626414602836957),  # |fy| = 0.375, larger dose
    (16, 0, 0.83, 10.0, 0.33886409539204293),   # k = 0.25 / 0.83
    (3, 59, 1.0, 37.5, 0.3106594568913417),     # fx = 3/64, fy = -5/64
])
def test_dose_weight_values(kx, y, pix, dose, want):
    w = O.dose_weight_sum(33, 64, pix, 300, [dose])
    assert w[y, kx] == pytest.approx(want, rel=1e-12)
    assert w[0, 0] == 0.0   # DC is never weighted


@pytest.mark.parametrize("kv,factor", [(300, 1.0), (400, 1.0), (250, 0.9), (200, 0.8),
                                       (160, 0.625), (120, 0.45), (80, 0.45)])
def test_dose_weight_kv_factor(kv, factor):
    """Dc scales by the piecewise-linear kV factor (1 at >= 300 kV, 0.8 at
    200, 0.45 at <= 120), so -ln w(kV) = -ln w(300) / factor everywhere."""
    w300 = O.dose_weight_sum(33, 64, 1.0, 300, [12.0]).ravel()[1:]
    wkv = O.dose_weight_sum(33, 64, 1.0, kv, [12.0]).ravel()[1:]
    assert np.allclose(np.log(wkv) * factor, np.log(w300), rtol=1e-12, atol=0)


def test_dose_weight_normaliser_over_frames():
    """sqrt(sum_i w_i^2) over accumulated doses FmDose * (i + 1), against the
    literal single-frame weights at k = 0.25."""
    ws = O.dose_weight_sum(33, 64, 1.0, 300, [10.0, 37.5])
    assert ws[0, 16] == pytest.approx(np.hypot(0.38773409603823816, 0.028641942865727247),
                                      rel=1e-12)


# ---- CInterpolateShift (Align/CInterpolateShift.cpp:41-142) ----
def test_interpolate_linear_data_exact_with_uneven_groups():
    """50 frames in groups of 4 form 12 x 4 + 1 x 2 groups (centres 1.5,
    5.5, ..., 45.5 and 48.5).  Linear data at the group centres comes back
    exactly at every frame centre, including the extrapolated frames 0, 1
    (before the first centre) and 49 (after the last)."""
    n = 50
    cent = np.array([4 * g + 1.5 for g in range(12)] + [48.5])
    ox, oy, ng = O.interpolate(0.3 * cent - 2.0, -1.7 * cent + 5.0, 4, n)
    assert ng == 13
    t = np.arange(n, dtype=float)
    assert np.allclose(ox, 0.3 * t - 2.0, atol=1e-12)
    assert np.allclose(oy, -1.7 * t + 5.0, atol=1e-12)


def test_interpolate_piecewise_brackets_hand_values():
    """Groups of 4 over 10 frames: centres 1.5, 5.5, 8.5 (the last group has
    two frames).  With x = (0, 4, 1) the frames take the line through their
    bracketing centres, the end lines extended outwards:
    frame 0: 0 + 1.0 * (0 - 1.5) = -1.5;  frame 3: 1.5;  frame 5: 3.5;
    frame 6: 4 - 1.0 * 0.5 = 3.5;  frame 8: 4 - 2.5 = 1.5;
    frame 9: 4 - 3.5 = 0.5."""
    ox, oy, ng = O.interpolate([0.0, 4.0, 1.0], [2.0, 2.0, 2.0], 4, 10)
    assert ng == 3
    want = [-1.5, -0.5, 0.5, 1.5, 2.5, 3.5, 3.5, 2.5, 1.5, 0.5]
    assert np.allclose(ox, want, atol=1e-12)
    assert np.allclose(oy, 2.0, atol=1e-12)


def test_interpolate_nearest_and_one_to_one():
    ox, oy, ng = O.interpolate([1.0, 2.0, 3.0], [7.0, 8.0, 9.0], 4, 10, nearest=True)
    assert np.array_equal(ox, [1, 1, 1, 1, 2, 2, 2, 2, 3, 3])
    assert np.array_equal(oy, [7, 7, 7, 7, 8, 8, 8, 8, 9, 9])
    g = np.arange(6.0) ** 2
    ox, oy, ng = O.interpolate(g, -g, 1, 6)
    assert ng == 6 and np.array_equal(ox, g) and np.array_equal(oy, -g)


def test_interpolate_raw_frame_sizes():
    """Integrated frames of 1, 1, 2, 2, 1, 3 raw frames (centres 0, 1, 2.5,
    4.5, 6, 8).  With -Group 2 a group holds frames until it has 2 x 1 raw
    frames (DataUtil/CFmGroupParam.cpp:59-78): {0, 1}, {2}, {3}, {4, 5},
    centres 0.5, 2.5, 4.5, 7.  x = centre (linear data) comes back as each
    frame's centre."""
    sizes = [1, 1, 2, 2, 1, 3]
    gc = np.array([0.5, 2.5, 4.5, 7.0])
    ox, oy, ng = O.interpolate(gc, 2 * gc, 2, 6, fmsize=sizes)
    assert ng == 4
    fc = np.array([0.0, 1.0, 2.5, 4.5, 6.0, 8.0])
    assert np.allclose(ox, fc, atol=1e-12)
    assert np.allclose(oy, 2 * fc, atol=1e-12)


# ---- GNormByStd2D + CDetectFeatures thresholds ----
def test_norm_by_var_brute_force():
    """Each pixel over the population variance of its in-bounds 5 x 5
    neighbourhood (fewer pixels at the borders)."""
    rng = np.random.default_rng(4)
    img = rng.uniform(1.0, 3.0, (9, 11))
    got = O.norm_by_var(img)
    for y in range(9):
        for x in range(11):
            nb = img[max(0, y - 2):y + 3, max(0, x - 2):x + 3]
            assert got[y, x] == pytest.approx(img[y, x] / (nb.var() + 1e-20), rel=1e-12)


def test_feature_thresholds_on_a_built_map():
    """A 128 x 128 map: 16000 background values, 6 outliers far below and 6
    far above (outside mu +- 4 sigma, dropped from the clipped statistics),
    and probes placed at fractions of the clipped sigma around both
    thresholds.  A feature is mu' - 4 sigma' <= v <= mu' - 0.08 sigma'; the
    clipped mean and sigma are computed here by brute force.  A sign slip
    in either threshold, or statistics taken over the outliers, moves a
    probe across."""
    rng = np.random.default_rng(12)
    n = 128 * 128
    base = rng.normal(10.0, 1.0, n - 12 - 8)
    base = np.clip(base, 7.0, 13.0)
    outl = np.r_[np.full(6, -200.0), np.full(6, 260.0)]
    # the probes sit at known offsets from the clipped statistics, which do
    # include them; solve for that fixed point by iteration
    offs = np.array([-4.05, -3.95, -0.10, -0.06, 0.06, 0.10, -2.0, 1.0])
    probes = 10.0 + offs
    for _ in range(50):
        allv = np.r_[base, probes]
        mu, sd = allv.mean(), allv.std()
        probes = mu + offs * sd
    img = np.r_[base, outl, probes]
    f, (lo, hi) = O.feature_threshold(img)
    # brute force: first-pass statistics over everything, then the clipped set
    m1, s1 = img.mean(), img.std()
    keep = img[(img >= m1 - 4 * s1) & (img <= m1 + 4 * s1)]
    assert keep.size == n - 12
    mu, sd = keep.mean(), keep.std()
    assert lo == pytest.approx(mu - 4 * sd, rel=1e-12)
    assert hi == pytest.approx(mu - 0.08 * sd, rel=1e-12)
    want = ((img >= mu - 4 * sd) & (img <= mu - 0.08 * sd)).astype(np.uint8)
    assert np.array_equal(f, want)
    pf = f[-8:]
    assert list(pf) == [0, 1, 1, 0, 0, 0, 1, 0]   # -4.05 out, -3.95 in, -0.10 in, -0.06 out
    assert f[n - 20:n - 8].sum() == 0             # outliers never features
