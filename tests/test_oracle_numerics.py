"""Pins of the oracle's path stages: reference KAT, the reference's own
compiled Util_LinEqs, closed forms and brute force on tiny inputs."""
import numpy as np
import pytest

import oracle as O

# LibSrc/Util/Util_LinEqs.cpp:160-180 (Util_LinEqs::Test)
KAT_A = np.array([[0.2368, 0.2471, 0.2568, 1.2671],
                  [0.1968, 0.2071, 1.2168, 0.2271],
                  [0.1581, 1.1675, 0.1768, 0.1871],
                  [1.1161, 0.1254, 0.1397, 0.1490]])
KAT_B = np.array([1.8471, 1.7471, 1.6471, 1.5471])
KAT_X = np.array([1.0405, 0.9870, 0.9350, 0.8812])


def test_lin_eqs_kat():
    ok, x = O.lin_eqs(KAT_A, KAT_B)
    assert ok
    # the reference lists the solution truncated to 4 decimals
    assert np.array_equal(np.floor(x * 1e4) / 1e4, KAT_X)
    assert np.allclose(x, np.linalg.solve(KAT_A, KAT_B), atol=1e-12)


def test_lin_eqs_matches_reference_build():
    rng = np.random.default_rng(3)
    r = O.ref_lin_eqs(KAT_A, KAT_B)
    if r is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    assert r[0] and np.allclose(r[1], O.lin_eqs(KAT_A, KAT_B)[1], atol=1e-5)
    for n in (3, 4, 7, 9):
        a = rng.standard_normal((n, n)) + n * np.eye(n)
        b = rng.standard_normal(n)
        ok_r, xr = O.ref_lin_eqs(a, b)
        ok_o, xo = O.lin_eqs(a, b)
        assert ok_r and ok_o
        assert np.allclose(xr, xo, rtol=1e-4, atol=1e-4)


def test_lin_eqs_singular():
    ok, _ = O.lin_eqs(np.zeros((3, 3)), np.ones(3))
    assert not ok


def test_peak_parabola_vertex_exact():
    yy, xx = np.mgrid[0:64, 0:64].astype(float)
    x0, y0 = 30.3, 35.7
    img = 100 - (xx - x0) ** 2 - 2 * (yy - y0) ** 2
    sh = O.peak2d(img)
    assert np.allclose(sh, [x0 - 32, y0 - 32], atol=1e-12)


def test_peak_tie_first_and_border():
    img = np.zeros((64, 64))
    img[10, 20] = 5.0
    img[40, 50] = 5.0
    sh = O.peak2d(img)
    assert np.allclose(sh, [20 - 32, 10 - 32])
    img = np.zeros((64, 64))
    img[0, 0] = 1.0     # moved to (1,1); flat neighbourhood -> no sub-pixel term
    assert np.allclose(O.peak2d(img), [1 - 32, 1 - 32])
    img = np.zeros((64, 64))
    img[63, 63] = 1.0
    assert np.allclose(O.peak2d(img), [62 - 32, 62 - 32])


def test_peak_rejects_large_delta():
    img = np.zeros((16, 16))
    img[5, 5] = 1.0
    img[5, 6] = 1.0 - 1e-9   # nearly flat -> |delta| ~ 0.5, kept
    sh = O.peak2d(img)
    assert abs(sh[0] - (5 + 0.5 - 8)) < 1e-6


def test_phase_shift_integer_delta():
    W, H = 64, 48
    img = np.zeros((H, W))
    img[10, 20] = 1.0
    f = O.rfft2(img)
    g = O.phase_shift(f, 3.0, -2.0)
    out = O.irfft2(g, W) / (W * H)
    # e^{+i k s} moves content by -s (Util/GPhaseShift2D.cu:25-37)
    assert np.unravel_index(np.argmax(out), out.shape) == (12, 17)
    assert np.max(np.abs(out)) == pytest.approx(1.0, abs=1e-9)


def test_phase_shift_subpixel_gaussian():
    W, H = 96, 80
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    s = 4.0
    img = np.exp(-((xx - 40) ** 2 + (yy - 30) ** 2) / (2 * s * s))
    out = O.irfft2(O.phase_shift(O.rfft2(img), 0.37, -1.21), W) / (W * H)
    ref = np.exp(-((xx - 40 + 0.37) ** 2 + (yy - 30 - 1.21) ** 2) / (2 * s * s))
    assert np.max(np.abs(out - ref)) < 1e-9


def test_fourier_resize_index_map():
    # Util/GFourierResize2D.cu:15-35 on an 8 x 6 (cmp 5 x 6) -> 3 x 4 crop
    cmp = (np.arange(30) + 1j * np.arange(30)).reshape(6, 5)
    out = O.fourier_resize(cmp, 3, 4)
    # out rows 0,1,2 <- in rows 0,1,2 ; out row 3 (> 4/2) <- in row 3-4+6 = 5
    assert np.array_equal(out[:3], cmp[:3, :3])
    assert np.array_equal(out[3], cmp[5, :3])


def test_correlate_integer_shift_peak():
    rng = np.random.default_rng(5)
    W, H = 256, 192
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    img = np.zeros((H, W))
    for _ in range(60):
        cx, cy, s = rng.uniform(0, W), rng.uniform(0, H), rng.uniform(2, 5)
        img += np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * s * s))
    g = np.roll(img, (5, -7), axis=(0, 1))
    F, G = O.rfft2(img), O.rfft2(g)
    win = O.correlate_window(F + G, G, 100.0)
    sh = O.peak2d(win)
    # conj(G) * F peaks at lag -d for g = f shifted by d
    assert np.allclose(sh, [7, -5], atol=1e-6)


def test_spline_exact_on_smooth_trajectory():
    n = 50
    t = np.sqrt(np.arange(n) / n)
    sx = 1.0 + 2.0 * t - 3.0 * t ** 2
    sy = -0.5 + 0.7 * t + 0.2 * t ** 2
    x, y = O.remove_spikes(sx, sy)
    assert np.allclose(x, sx, atol=1e-12) and np.allclose(y, sy, atol=1e-12)


def test_spike_removed():
    n = 40
    sx = np.linspace(0, 4, n)
    sy = np.zeros(n)
    sx2 = sx.copy()
    sx2[17] += 25.0
    x, _ = O.remove_spikes(sx2, sy, single=True)
    assert x[17] == pytest.approx(0.5 * (sx[16] + sx[18]))
    assert np.allclose(np.delete(x, 17), np.delete(sx, 17))


def test_dose_weights_normalised_and_monotone():
    cx, cy, N = 33, 64, 12
    acc = 1.5 * (np.arange(N) + 1)
    ws = O.dose_weight_sum(cx, cy, 1.0, 300, acc)
    assert ws[0, 0] == 0.0
    # sqrt(sum w^2) <= sqrt(N) and > 0 away from DC; decreasing with frequency
    assert np.all(ws.ravel()[1:] > 0) and np.all(ws <= np.sqrt(N) + 1e-12)
    assert ws[0, 1] > ws[0, 20] and ws[1, 0] > ws[30, 0]


def _make_relative_emulated(s, quirks=True):
    """Literal flat-array emulation of Align/CPatchShifts.cpp:263-286."""
    nf, npat = s.shape[:2]
    flat = list(s.reshape(-1)) + [0.0] * 8
    means = []
    for i in range(1, nf):
        p = s[i]
        m = p.mean(0)
        sd = np.sqrt(((p - m) ** 2).sum() / npat)
        means.append((np.hypot(*m) + sd, i))
    fmin, ref = 1e20, 0
    for v, i in means:
        if v > fmin:
            continue
        fmin, ref = v, i
    rb = ref * npat * 2
    for i in range(nf):
        pb = i * npat * 2
        for j in range(npat):
            k = 2 * j
            flat[pb + k] -= flat[rb + k]
            flat[pb + k + 1] -= flat[rb + k + 2] if quirks else flat[rb + k + 1]
    return np.array(flat[: nf * npat * 2]).reshape(nf, npat, 2), ref


def test_make_relative_quirk_emulation():
    rng = np.random.default_rng(7)
    s = rng.normal(0, 2, (6, 4, 2))
    out, ref = O.make_relative(s, quirks=True)
    exp, eref = _make_relative_emulated(s, quirks=True)
    assert ref == eref
    assert np.allclose(out, exp, atol=1e-12)


def test_make_relative_plain():
    rng = np.random.default_rng(8)
    s = rng.normal(0, 2, (6, 4, 2))
    out, ref = O.make_relative(s, quirks=False)
    assert np.allclose(out, s - s[ref][None], atol=1e-12)


def test_detect_bads_brute():
    N, P = 3, 9
    cen = np.array([[(i % 3 + 0.5) * 300, (i // 3 + 0.5) * 300] for i in range(P)])
    s = np.full((N, P, 2), 0.5)
    s[1, 4] = [9.0, 0.0]            # an outlier in the middle patch
    bad = O.detect_bads(s, cen, 900, 900)
    assert bad[1, 4] == 1 and bad.sum() == 1


def test_local_warp_identity_and_integer_shift():
    rng = np.random.default_rng(9)
    img = rng.standard_normal((40, 48))
    cen = np.array([[12.0, 10.0], [36.0, 10.0], [12.0, 30.0], [36.0, 30.0]])
    assert np.array_equal(O.local_warp(img, cen, np.zeros((4, 2))), img)
    # every pixel is within r = 0.6 of a patch centre, so all move by 3.25;
    # (int) truncates toward zero: x - 3.25 in (-1, 0) reads column 0
    out = O.local_warp(img, cen, np.tile([3.25, 0.0], (4, 1)))
    assert np.array_equal(out[:, 4:], img[:, :-4])
    assert np.array_equal(out[:, 3], img[:, 0])
    # out of range: both coordinates go through |.| and 811*k mod N
    for x in range(3):
        k = abs(int(x - 3.25))
        assert np.array_equal(out[:, x], img[(811 * np.arange(40)) % 40, (811 * k) % 48])


def test_geometry_of_reference_sizes():
    g = O.geometry(5760, 4092, 5, 5)
    assert g["xcf"] == (2882, 2048) and g["patch"] == (576, 408)
    g = O.geometry(4096, 4096, 7, 5)
    assert g["xcf"] == (2048, 2048) and g["patch"] == (292, 408)
    g = O.geometry(1024, 1024)
    assert g["xcf"] == (1024, 1024) and g["xcf_bin"] == (1.0, 1.0)
