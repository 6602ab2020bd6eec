"""The whole path at the reference's production geometries against the FP64
oracle, including the Fourier crop to the alignment size and its bin of
about 2 (or 4): 5760 x 4092 frames with 5 x 5 patches (2882 x 2048 xcf,
576 x 408 patches), 4096 x 4096 with 7 x 5 patches (2048 x 2048 xcf,
292 x 408 patches) and 11520 x 8184 super-resolution frames with -FtBin 2
(xcf bin 4, 5760 x 4092 output).  Frames are few, so the oracle finishes in
tens of seconds; every geometric quantity is that of the full movies.

Bars: global and patch shifts within 0.05 px of the oracle (full-resolution
pixels); the feature-driven patch starts and the bad-patch flags equal; the
sums within 1e-4 relative L2 of the oracle's replay of the GPU's shifts.
FP16 and BF16 alignment stacks: patch-shift RMS within 0.1 px and global
shifts within 0.05 px of the FP64 oracle."""
import numpy as np
import pytest

import oracle as O
from motioncor3_amd import lib as L
from motioncor3_amd.synth import MovieRecipe

pytestmark = pytest.mark.gpu

CASES = {
    "5760x4092x6_p5x5": dict(W=5760, H=4092, N=6, nb=20000, patch=(5, 5), pix=0.83, ft_bin=1.0,
                             seed=4242),
    "4096x4096x6_p7x5": dict(W=4096, H=4096, N=6, nb=14000, patch=(7, 5), pix=0.83, ft_bin=1.0,
                             seed=4096),
    "11520x8184x6_p5x5_ftbin2": dict(W=11520, H=8184, N=6, nb=80000, patch=(5, 5), pix=0.415,
                                     ft_bin=2.0, seed=11520),
}
_CACHE = {}


def _kw(c):
    return dict(patch=c["patch"], bft=(500, 100), pix_size=c["pix"], kv=300, fm_dose=1.0,
                ft_bin=c["ft_bin"])


def _case(ctx, name):
    """Frames rendered on the GPU (device and host copies) and the oracle's
    own run on them, shared by the tests of one geometry."""
    if name not in _CACHE:
        c = CASES[name]
        r = MovieRecipe(c["W"], c["H"], c["N"], c["nb"], seed=c["seed"], dose=1.0, local_max=3.0)
        frames = ctx.empty((c["N"], c["H"], c["W"]), np.float32)
        ctx.synth_render(frames, r)
        host = frames.numpy()
        print("oracle run", name, flush=True)
        ores = O.align_movie(host, O.default_params(**_kw(c)))
        _CACHE[name] = (frames, host, ores)
    return _CACHE[name]


def _rms(d):
    return float(np.sqrt(np.mean(np.sum(np.asarray(d) ** 2, axis=-1))))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name", list(CASES))
def test_production_geometry_matches_oracle(ctx, name):
    c = CASES[name]
    print("geometry", name, flush=True)
    frames, host, ores = _case(ctx, name)
    gres = ctx.align_movie(frames, L.default_params(**_kw(c)))
    # discrete decisions: feature-driven placement and bad patches
    assert np.array_equal(gres["patch_starts"], ores["patch_starts"])
    assert np.array_equal(gres["patch_bad"], ores["patch_bad"])
    dg = gres["global_shift"] - ores["global_shift"]
    dp = gres["patch_shift"] - ores["patch_shift"]
    print("%s: global max %.2e rms %.2e px, patch max %.2e rms %.2e px, iters %d/%d" % (
        name, np.max(np.abs(dg)), _rms(dg), np.max(np.abs(dp)), _rms(dp),
        gres["global_iters"], ores["global_iters"]), flush=True)
    assert np.max(np.abs(dg)) <= 0.05
    assert np.max(np.abs(dp)) <= 0.05
    assert np.allclose(gres["patch_center"], ores["patch_center"], atol=1e-3)
    # the sums on the GPU's shifts (-InAln replay)
    orep = O.align_movie(host, O.default_params(**_kw(c)), replay_global=gres["global_shift"],
                         replay_patch=gres["patch_shift"], replay_centers=gres["patch_center"],
                         replay_bad=gres["patch_bad"])
    assert gres["num_sums"] == orep["sums"].shape[0] == 3
    for i in range(gres["num_sums"]):
        g = gres["sums"][i].numpy()
        assert g.shape == orep["sums"][i].shape
        rel = _rel(g, orep["sums"][i])
        print("  sum %d rel-L2 vs oracle replay %.2e, vs oracle's own run %.2e" % (
            i, rel, _rel(g, ores["sums"][i])), flush=True)
        assert rel < 1e-4, i


def _replay(host, c, res):
    return O.align_movie(host, O.default_params(**_kw(c)), replay_global=res["global_shift"],
                         replay_patch=res["patch_shift"], replay_centers=res["patch_center"],
                         replay_bad=res["patch_bad"])


@pytest.mark.timeout(600)
@pytest.mark.parametrize("fmt", [1, 2])
def test_reduced_precision_stacks_against_fp64(ctx, fmt):
    """FP16 (1) and BF16 (2) alignment stacks with per-transform scale at
    5760 x 4092 with 5 x 5 patches against the FP64 oracle: shift bars, and
    the sums (FP32 frame spectra) within 1e-4 of the oracle's replay of the
    run's own shifts.  The sums against the oracle's own run are printed:
    the nearest-pixel warp turns shift differences of 1e-5 px into whole
    pixels at the truncation boundaries, so they differ by more than the
    shifts suggest even for the FP32 path."""
    name = "5760x4092x6_p5x5"
    c = CASES[name]
    frames, host, ores = _case(ctx, name)
    red = ctx.align_movie(frames, L.default_params(spectra_fmt=fmt, **_kw(c)))
    assert red["spectra_fmt_used"] == (fmt, fmt) and red["frame_fmt_used"] == 0
    assert np.array_equal(red["patch_starts"], ores["patch_starts"])
    dg = red["global_shift"] - ores["global_shift"]
    dp = red["patch_shift"] - ores["patch_shift"]
    rep = _replay(host, c, red)
    rels = [_rel(red["sums"][i].numpy(), rep["sums"][i]) for i in range(red["num_sums"])]
    own = [_rel(red["sums"][i].numpy(), ores["sums"][i]) for i in range(red["num_sums"])]
    print("%s stacks: global max %.2e rms %.2e px, patch max %.2e rms %.2e px; sums rel-L2 vs "
          "oracle replay %s, vs oracle's own run %s" % (
              {1: "FP16", 2: "BF16"}[fmt], np.max(np.abs(dg)), _rms(dg), np.max(np.abs(dp)), _rms(dp),
              " ".join("%.2e" % v for v in rels), " ".join("%.2e" % v for v in own)), flush=True)
    assert np.max(np.abs(dg)) <= 0.05
    assert _rms(dp) <= 0.1
    assert max(rels) < 1e-4


@pytest.mark.timeout(600)
@pytest.mark.parametrize("fmt", [1, 2])
def test_reduced_precision_frames_against_fp64(ctx, fmt):
    """The full-frame spectra (forward transforms, local-correction output,
    what the sums read) stored in FP16 (1) or BF16 (2) with a per-frame
    scale, the alignment stacks in the same format: shifts against the FP64
    oracle (bars as above) and the sums' relative L2 against the oracle's
    replay of the run's own shifts, i.e. the storage error of the frames
    themselves (stated, with a bar of 2e-3 for FP16 and 1e-2 for BF16)."""
    name = "5760x4092x6_p5x5"
    c = CASES[name]
    frames, host, ores = _case(ctx, name)
    red = ctx.align_movie(frames, L.default_params(spectra_fmt=fmt, frame_fmt=fmt, **_kw(c)))
    assert red["spectra_fmt_used"] == (fmt, fmt) and red["frame_fmt_used"] == fmt
    assert np.array_equal(red["patch_starts"], ores["patch_starts"])
    dg = red["global_shift"] - ores["global_shift"]
    dp = red["patch_shift"] - ores["patch_shift"]
    rep = _replay(host, c, red)
    rels = [_rel(red["sums"][i].numpy(), rep["sums"][i]) for i in range(red["num_sums"])]
    print("%s frames + stacks: global max %.2e rms %.2e px, patch max %.2e rms %.2e px; sums rel-L2 "
          "vs oracle replay %s" % ({1: "FP16", 2: "BF16"}[fmt], np.max(np.abs(dg)), _rms(dg),
                                   np.max(np.abs(dp)), _rms(dp), " ".join("%.2e" % v for v in rels)),
          flush=True)
    assert np.max(np.abs(dg)) <= 0.05
    assert _rms(dp) <= 0.1
    assert max(rels) < (2e-3 if fmt == 1 else 1e-2)


def test_reduced_precision_frames_streams_and_pipeline(ctx):
    """FP16 frame spectra: two chunk streams give the bits of one, and a
    pipelined pair of movies the bits of separate calls."""
    c = CASES["5760x4092x6_p5x5"]
    frames, host, ores = _case(ctx, "5760x4092x6_p5x5")
    kw = dict(spectra_fmt=1, frame_fmt=1, **_kw(c))
    one = ctx.align_movie(frames, L.default_params(chunk_streams=1, **kw))
    two = ctx.align_movie(frames, L.default_params(chunk_streams=2, **kw))
    assert np.array_equal(one["patch_shift"], two["patch_shift"])
    for a, b in zip(one["sums"], two["sums"]):
        assert np.array_equal(a.numpy(), b.numpy())
    r2 = MovieRecipe(c["W"], c["H"], c["N"], c["nb"], seed=99, dose=1.0, local_max=3.0)
    f2 = ctx.empty((c["N"], c["H"], c["W"]), np.float32)
    ctx.synth_render(f2, r2)
    single = [two, ctx.align_movie(f2, L.default_params(**kw))]
    both = ctx.align_movies([frames, f2], L.default_params(movie_pipeline=1, **kw))
    for a, b in zip(both, single):
        assert np.array_equal(a["global_shift"], b["global_shift"])
        assert np.array_equal(a["patch_shift"], b["patch_shift"])
        for x, y in zip(a["sums"], b["sums"]):
            assert np.array_equal(x.numpy(), y.numpy())
