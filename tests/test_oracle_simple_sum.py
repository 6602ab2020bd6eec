"""-Align 0 in the oracle (Align/CSimpleSum.cpp:22-135 with
MrcUtil/CSumFFTStack.cpp:60-210): plain sums, the even-indexed / odd-indexed
split, no clamp at zero, and the -FtBin crop, pinned by closed forms (sums of
the frames themselves, and a band-limited cosine that a Fourier crop
resamples exactly)."""
import numpy as np

import oracle as O

N, H, W = 5, 48, 64


def test_sums_and_split_without_alignment():
    rng = np.random.default_rng(0)
    fr = rng.normal(0, 1, (N, H, W)).astype(np.float32)  # negative values too
    p = O.default_params(align=0, split_sum=1, pix_size=1.0, kv=300, fm_dose=1.0)
    r = O.align_movie(fr, p)
    f64 = fr.astype(np.float64)
    assert r["sums"].shape == (3, H, W) and r["global_iters"] == 0
    assert np.allclose(r["sums"][0], f64.sum(0), atol=1e-9)
    assert np.allclose(r["sums"][1], f64[0::2].sum(0), atol=1e-9)  # frames 1, 3, 5 (1-based): _ODD
    assert np.allclose(r["sums"][2], f64[1::2].sum(0), atol=1e-9)  # _EVN
    assert (r["sums"][0] < 0).any()


def test_ft_bin_resamples_a_band_limited_image():
    y, x = np.mgrid[0:H, 0:W]
    img = 2.0 + 0.5 * np.cos(2 * np.pi * (3 * x / W + 2 * y / H))
    fr = np.stack([img * (k + 1) for k in range(N)]).astype(np.float32)
    p = O.default_params(align=0, ft_bin=2.0, out_stack=(1, 1))
    r = O.align_movie(fr, p)
    yo, xo = np.mgrid[0:H // 2, 0:W // 2]
    half = 2.0 + 0.5 * np.cos(2 * np.pi * (3 * xo / (W // 2) + 2 * yo / (H // 2)))
    assert r["sums"].shape == (1, H // 2, W // 2)
    tot = fr.astype(np.float64).sum(0)[0, 0] / img[0, 0]
    assert np.allclose(r["sums"][0], tot * half, atol=1e-9)
    for k in range(N):
        assert np.allclose(r["stack"][k], (k + 1) * half.astype(np.float32), atol=1e-5)


def test_stack_and_zbin_without_alignment():
    rng = np.random.default_rng(1)
    fr = rng.normal(3, 1, (N, H, W)).astype(np.float32)
    one = O.align_movie(fr, O.default_params(align=0, out_stack=(1, 1)))["stack"]
    two = O.align_movie(fr, O.default_params(align=0, out_stack=(1, 2)))["stack"]
    assert np.allclose(one, fr, atol=1e-9)
    assert two.shape[0] == 3
    for k, idx in enumerate(([0, 1], [2, 3], [4])):
        assert np.allclose(two[k], fr[idx].astype(np.float64).sum(0), atol=1e-9)
