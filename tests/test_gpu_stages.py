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


def test_local_warp(ctx):
    rng = np.random.default_rng(8)
    N, H, W, P = 3, 300, 420, 9
    frames = rng.standard_normal((N, H, W)).astype(np.float32)
    cen = np.array([[(i % 3 + 0.5) * W / 3, (i // 3 + 0.5) * H / 3] for i in range(P)])
    sh = rng.uniform(-4, 4, (N, P, 2))
    bad = (rng.uniform(0, 1, (N, P)) < 0.15).astype(np.uint8)
    out = ctx.local_warp(ctx.upload(frames), cen.astype(np.float32),
                         sh.astype(np.float32), bad).numpy()
    for i in range(N):
        ref = O.local_warp(frames[i].astype(np.float64), cen.astype(np.float32).astype(np.float64),
                           sh[i].astype(np.float32).astype(np.float64), bad[i])
        mism = np.mean(out[i] != ref)
        assert mism < 1e-3, mism


def test_dose_weight_sum(ctx):
    W, H, N = 512, 384, 30
    out = ctx.dose_weight_sum(W, H, N, 0.83, 300, 1.2).numpy().T
    ref = O.dose_weight_sum(W // 2 + 1, H, 0.83, 300, 1.2 * (np.arange(N) + 1))
    assert np.allclose(out, ref, rtol=2e-6, atol=1e-7)
