"""Stage kernels through the C ABI against the FP64 oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _blobs(W, H, n, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    img = np.ones((H, W))
    for _ in range(n):
        cx, cy, s = rng.uniform(0, W), rng.uniform(0, H), rng.uniform(2, 6)
        img += rng.uniform(0.5, 1.5) * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * s * s))
    return img


def _col_major(cmp_ref_layout):
    """[H][cx] complex -> complex64 [cx][H] (the library's spectrum layout)."""
    return np.ascontiguousarray(cmp_ref_layout.T).astype(np.complex64)


def test_phase_shift_accumulate(ctx):
    W, H, N = 96, 80, 5
    rng = np.random.default_rng(2)
    specs = [O.rfft2(rng.standard_normal((H, W))) / (W * H) for _ in range(N)]
    shifts = rng.uniform(-6, 6, (N, 2))
    ref = sum(O.phase_shift(specs[i], *shifts[i]) for i in range(N))
    d = ctx.upload(np.stack([_col_major(s) for s in specs]))
    out = ctx.phase_shift_accumulate(d, shifts, W).numpy().T
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 2e-6


@pytest.mark.parametrize("W,H,bft,po", [(256, 192, 500.0, 0), (576, 408, 100.0, 0),
                                        (2882, 2048, 500.0, 0), (340, 340, 100.0, 1)])
def test_correlate_window(ctx, W, H, bft, po):
    img = _blobs(W, H, 80, W)
    g = np.roll(img, (3, -5), axis=(0, 1)) + 0.05 * np.random.default_rng(1).standard_normal((H, W))
    F, G = O.rfft2(img) / (W * H), O.rfft2(g) / (W * H)
    ref = O.correlate_window(F + G, G, bft, bool(po))
    out = ctx.correlate_window(ctx.upload(_col_major(F + G)), ctx.upload(_col_major(G)), W, bft,
                              bool(po))
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(out - ref)) < 2e-5 * scale
    assert np.allclose(O.peak2d(out), O.peak2d(ref), atol=1e-3)


def test_peak_fit(ctx):
    rng = np.random.default_rng(4)
    wins = rng.standard_normal((40, 64, 64)).astype(np.float32)
    wins[3, 10, 20] = wins[3, 40, 50] = 9.0   # tie -> first
    wins[5, 0, 0] = 9.0                       # border clamp
    out = ctx.peak_fit(ctx.upload(wins))
    ref = np.array([O.peak2d(w.astype(np.float64)) for w in wins])
    assert np.allclose(out, ref, atol=1e-5)


def _spiky(rng, n):
    t = np.sqrt(np.arange(n) / n)
    sx = 8 * t - 3 * t ** 2 + rng.normal(0, 0.3, n)
    sy = -5 * t + rng.normal(0, 0.3, n)
    sx[n // 3] += 12.0
    sy[n // 2] -= 7.0
    # a two-frame excursion survives the single-spike pass and is left to
    # the spline refit
    sy[n // 4:n // 4 + 2] += 4.5
    return sx, sy


@pytest.mark.parametrize("quirks", [True, False])
def test_despike_fp64(ctx, quirks):
    """One wave per trajectory: lengths below, at and above the 64-lane wave,
    with 0 to 6 spline knots, against CRemoveSpikes1D restated in the oracle."""
    rng = np.random.default_rng(5)
    for n in (20, 50, 64, 73, 130, 211):
        sx, sy = _spiky(rng, n)
        gx, gy = ctx.despike(sx, sy, quirks)
        ox, oy = O.remove_spikes(sx, sy, quirks=quirks)
        assert np.allclose(gx, ox, atol=1e-9) and np.allclose(gy, oy, atol=1e-9), n


def test_despike_batch_matches_single(ctx):
    rng = np.random.default_rng(11)
    n, count = 97, 9
    tr = [_spiky(rng, n) for _ in range(count)]
    bx, by = ctx.despike(np.stack([a for a, _ in tr]), np.stack([b for _, b in tr]))
    for k, (sx, sy) in enumerate(tr):
        ox, oy = O.remove_spikes(sx, sy)
        assert np.allclose(bx[k], ox, atol=1e-9) and np.allclose(by[k], oy, atol=1e-9), k


@pytest.mark.parametrize("quirks", [True, False])
def test_patch_stats_fp64(ctx, quirks):
    rng = np.random.default_rng(6)
    N, P = 50, 25
    s = rng.normal(0, 1.5, (N, P, 2))
    s[7, 12] = [15, -9]
    cen = np.array([[(i % 5 + 0.5) * 1152, (i // 5 + 0.5) * 818] for i in range(P)])
    gs, gbad, gref = ctx.patch_stats(s, cen, 5760, 4092, quirks)
    os_, oref = O.make_relative(s, quirks)
    obad = O.detect_bads(os_, cen, 5760, 4092)
    assert gref == oref
    assert np.allclose(gs, os_, atol=1e-9)
    assert np.array_equal(gbad, obad)


def _warp_source(W, H, cen, sh, bad):
    """FP64 source coordinates x - sx, y - sy of mGCorrect3D's weighted shift
    (Correct/GCorrectPatchShift.cu:17-66), brute force over all patches."""
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    sx, sy, w = np.zeros((H, W)), np.zeros((H, W)), np.zeros((H, W))
    for p in range(len(cen)):
        if bad[p]:
            continue
        r = np.hypot((xx - cen[p, 0]) / W, (yy - cen[p, 1]) / H)
        e = np.where(r > 0.6, 0.0, np.exp(-100.0 * r))
        w += e
        sx += sh[p, 0] * e
        sy += sh[p, 1] * e
    ok = w > 0
    sx[ok] /= w[ok]
    sy[ok] /= w[ok]
    return xx - sx, yy - sy


def test_local_warp(ctx):
    """Nearest-pixel gather: the GPU forms the weighted shift in FP32 like the
    reference, the oracle in FP64 before the FP32 subtraction.  A pixel may
    read a different source only where its FP64 source coordinate lies
    within 1e-4 px of a truncation boundary; everywhere else the gathered
    values are identical."""
    rng = np.random.default_rng(8)
    N, H, W, P = 3, 300, 420, 9
    frames = rng.standard_normal((N, H, W)).astype(np.float32)
    cen = np.array([[(i % 3 + 0.5) * W / 3, (i // 3 + 0.5) * H / 3] for i in range(P)])
    sh = rng.uniform(-4, 4, (N, P, 2))
    bad = (rng.uniform(0, 1, (N, P)) < 0.15).astype(np.uint8)
    out = ctx.local_warp(ctx.upload(frames), cen.astype(np.float32),
                         sh.astype(np.float32), bad).numpy()
    c32 = cen.astype(np.float32).astype(np.float64)
    for i in range(N):
        s32 = sh[i].astype(np.float32).astype(np.float64)
        ref = O.local_warp(frames[i].astype(np.float64), c32, s32, bad[i])
        mism = out[i] != ref
        xs, ys = _warp_source(W, H, c32, s32, bad[i])
        edge = (np.abs(xs - np.round(xs)) < 1e-4) | (np.abs(ys - np.round(ys)) < 1e-4)
        assert not np.any(mism & ~edge), (i, int(np.sum(mism & ~edge)))
        assert np.sum(mism) <= 20, (i, int(np.sum(mism)))


def test_dose_weight_sum(ctx):
    W, H, N = 512, 384, 30
    out = ctx.dose_weight_sum(W, H, N, 0.83, 300, 1.2).numpy().T
    ref = O.dose_weight_sum(W // 2 + 1, H, 0.83, 300, 1.2 * (np.arange(N) + 1))
    assert np.allclose(out, ref, rtol=2e-6, atol=1e-7)


# ---- stage entry points of the boundary: crop, feature centres, patch
#      extraction, output finalisation ----
@pytest.mark.parametrize("inw,inh,outw,outh,batch", [(5760, 4092, 2882, 2048, 2), (8, 6, 4, 4, 3),
                                                     (14, 10, 20, 16, 2), (1152, 816, 576, 408, 1)])
def test_fourier_crop_bit_exact(ctx, inw, inh, outw, outh, batch):
    """GFourierResize2D is a copy: every kept value is bit-identical to the
    oracle's, padded rows and columns stay zero."""
    rng = np.random.default_rng(inw + outh)
    spec = (rng.standard_normal((batch, inh, inw // 2 + 1)) +
            1j * rng.standard_normal((batch, inh, inw // 2 + 1))).astype(np.complex64)
    out = ctx.fourier_crop(ctx.upload(np.stack([_col_major(x) for x in spec])), inw, outw,
                           outh).numpy()
    for b in range(batch):
        ref = O.fourier_resize(spec[b].astype(np.complex128), outw // 2 + 1, outh)
        assert np.array_equal(out[b].T.astype(np.complex128), ref), b


def _xcf_sum_spec(xw, xh, seed):
    """An aligned-sum-like xcf image (blobs on a background of 50 counts with
    shot noise) as the normalised FP32 spectrum the path forms."""
    rng = np.random.default_rng(seed)
    img = 50.0 * _blobs(xw, xh, 300, seed) / 2.0 + rng.normal(0, 2.0, (xh, xw))
    return (np.fft.rfft2(img) / (xw * xh)).astype(np.complex64)


@pytest.mark.parametrize("xw,xh,px,py", [(2882, 2048, 5, 5), (2048, 2048, 7, 5), (1024, 1024, 3, 3)])
def test_feature_centres_match_oracle(ctx, xw, xh, px, py):
    """The feature map against the oracle's FP64 map on the same spectrum:
    equal except where a normalised value lies within 1e-4 of the threshold
    span from a threshold; the patch starts equal the oracle's placement
    from the GPU map (Align/CDetectFeatures.cpp:289-396,
    Align/CPatchCenters.cpp:42-105), and the oracle's from its own map when
    the maps agree."""
    spec = _xcf_sum_spec(xw, xh, xw + px)
    fm, st = ctx.feature_centres(ctx.upload(_col_major(spec)), xw, px, py)
    s64 = spec.astype(np.complex128)
    ref = O.feature_map(s64, xw, xh)
    img = O.irfft2(O.fourier_resize(s64, 65, 128), 128)
    img = O.norm_by_var(img + 1.0 - img.min())
    f2, (lo, hi) = O.feature_threshold(img)
    assert np.array_equal(f2, ref)
    tol = 1e-4 * (hi - lo)
    edge = (np.abs(img - lo) < tol) | (np.abs(img - hi) < tol)
    mism = fm != ref
    assert not np.any(mism & ~edge), int(np.sum(mism & ~edge))
    assert 0 < ref.sum() < 128 * 128
    assert np.array_equal(st, O.patch_starts(fm, px, py, xw, xh))
    if not mism.any():
        assert np.array_equal(st, O.patch_starts(ref, px, py, xw, xh))


@pytest.mark.parametrize("xw,xh,px,py", [(2882, 2048, 5, 5), (2048, 2048, 7, 5), (2882, 2048, 7, 5)])
def test_extract_patches_match_fft_of_windows(ctx, xw, xh, px, py):
    """CExtractPatch + the unnormalised patch R2C: every patch of every frame
    against the FP64 FFT of the same window (576 x 408 uses the register-pass
    kernels, 292 x 408 and 410 x 408 the stage kernels with the 73- and
    41-point prime stages)."""
    pw, ph = O.geometry(xw, xh, px, py)["patch"]
    rng = np.random.default_rng(px * 100 + xw)
    N = 3
    frames = rng.uniform(0.0, 4.0, (N, xh, xw)).astype(np.float32)
    P = px * py
    starts = np.stack([rng.integers(0, xw - pw + 1, P), rng.integers(0, xh - ph + 1, P)], 1)
    starts[0] = (0, 0)
    starts[-1] = (xw - pw, xh - ph)
    out = ctx.extract_patches(ctx.upload(frames), starts, pw, ph).numpy()
    assert out.shape == (P * N, pw // 2 + 1, ph)
    for q in range(P):
        for i in range(N):
            x0, y0 = starts[q]
            win = frames[i, y0:y0 + ph, x0:x0 + pw].astype(np.float64)
            ref = np.fft.rfft2(win)
            got = out[q * N + i].T
            assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 2e-6, (q, i)


@pytest.mark.parametrize("W,H,ft_bin", [(1152, 816, 1.0), (1152, 816, 2.0), (5760, 4092, 2.0)])
def test_finalize_sums_match_oracle(ctx, W, H, ft_bin):
    """CCorrectFullShift::mUnpadSums: crop to the -FtBin size, C2R, clamp at
    zero, against the oracle's resize and FP64 inverse of the same spectra."""
    rng = np.random.default_rng(W + int(ft_bin))
    K = 2
    specs = []
    for k in range(K):
        img = _blobs(W, H, 40, W + k) if W < 2000 else rng.uniform(0, 1, (H, W)) + _blobs(W, H, 4, k)
        img = img - 0.2 * img.mean()        # some negative values to clamp
        specs.append((np.fft.rfft2(img) / (W * H) * 50.0).astype(np.complex64))
    out = ctx.finalize_sums(ctx.upload(np.stack([_col_major(x) for x in specs])), W, ft_bin).numpy()
    ow, oh = out.shape[2], out.shape[1]
    if ft_bin == 1.0:
        assert (ow, oh) == (W, H)
    for k in range(K):
        ref = np.maximum(O.irfft2(O.fourier_resize(specs[k].astype(np.complex128), ow // 2 + 1, oh), ow), 0)
        assert np.linalg.norm(out[k] - ref) / np.linalg.norm(ref) < 1e-5, k
        assert out[k].min() >= 0
