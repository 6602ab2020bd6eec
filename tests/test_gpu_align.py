"""Whole path (mc_align_movie) against the FP64 oracle on seeded movies.

Tolerances: shifts within 0.05 px of the oracle; sums within 1e-4 relative
L2 when both paths apply the same shifts (-InAln replay semantics), because
the |s| < 0.1 px skip rule and the (int) pixel truncation of the local warp
turn any shift difference into a discrete change of the sum."""
import numpy as np
import pytest

import oracle as O
from motioncor3_amd import lib as L
from motioncor3_amd.synth import MovieRecipe

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _params(**kw):
    return L.default_params(**kw), O.default_params(**kw)


@pytest.mark.parametrize("W,H,N,nb,it", [(512, 512, 10, 120, 7), (1024, 1024, 20, 400, 7)])
def test_global_alignment_matches_oracle(ctx, W, H, N, nb, it):
    print("global", W, H, N, flush=True)
    r = MovieRecipe(W, H, N, nb, seed=W + N)
    mv = r.movie()
    pg, po = _params(iter=it, pix_size=1.0, kv=300)
    gres = ctx.align_movie(ctx.upload(mv), pg)
    ores = O.align_movie(mv, po)
    assert np.max(np.abs(gres["global_shift"] - ores["global_shift"])) <= 0.05
    assert abs(gres["global_iters"] - ores["global_iters"]) <= 1
    # sum on identical shifts
    orep = O.align_movie(mv, po, replay_global=gres["global_shift"])
    assert _rel(gres["sums"][0].numpy(), orep["sums"][0]) < 1e-4
    # ground truth: reported shift = -(displacement) + const
    g = gres["global_shift"]
    e = (g - g.mean(0)) + (r.displacement - r.displacement.mean(0))
    assert np.sqrt((e ** 2).sum(1).mean()) < 0.3


def test_dose_weighted_and_split_sums(ctx):
    W, H, N = 512, 384, 12
    r = MovieRecipe(W, H, N, 100, seed=11)
    mv = r.movie()
    pg, po = _params(iter=5, pix_size=1.0, kv=200, fm_dose=2.5, split_sum=1)
    gres = ctx.align_movie(ctx.upload(mv), pg)
    assert gres["num_sums"] == 5
    orep = O.align_movie(mv, po, replay_global=gres["global_shift"])
    assert orep["sums"].shape[0] == 5
    for i in range(5):
        assert _rel(gres["sums"][i].numpy(), orep["sums"][i]) < 1e-4, i


def test_patch_alignment_matches_oracle(ctx):
    W, H, N = 1024, 1024, 12
    r = MovieRecipe(W, H, N, 500, seed=21, dose=4.0, local_max=2.0)
    mv = r.movie()
    kw = dict(patch=(3, 3), iter=10, pix_size=1.0, kv=300, fm_dose=1.0)
    pg, po = _params(**kw)
    gres = ctx.align_movie(ctx.upload(mv), pg)
    own = O.align_movie(mv, po)
    # discrete decisions first: the feature-driven patch placement
    assert np.array_equal(own["patch_starts"], gres["patch_starts"])
    ores = own
    assert np.max(np.abs(gres["global_shift"] - ores["global_shift"])) <= 0.05
    assert np.allclose(gres["patch_center"], ores["patch_center"], atol=1e-3)
    d = gres["patch_shift"] - ores["patch_shift"]
    assert np.max(np.abs(d)) <= 0.05
    assert np.array_equal(gres["patch_bad"], ores["patch_bad"])
    orep = O.align_movie(mv, po, replay_global=gres["global_shift"],
                         replay_patch=gres["patch_shift"], replay_centers=gres["patch_center"],
                         replay_bad=gres["patch_bad"])
    for i in range(gres["num_sums"]):
        assert _rel(gres["sums"][i].numpy(), orep["sums"][i]) < 1e-4, i
    # stage clock: every stage of a patch alignment ran on the GPU, and the
    # stages are intervals of the call
    st = gres["seconds_stage"]
    assert set(st) == set(L.STAGES)
    assert all(st[k] > 0 for k in ("global_align", "patch_align", "correct", "sums")), st
    assert sum(st.values()) <= gres["seconds_total"] * 1.05 + 1e-3, (st, gres["seconds_total"])


def test_production_geometry_ground_truth(ctx):
    """5760 x 4092 frames with 5x5 patches (the benchmark geometry, fewer
    frames): size-independent properties only."""
    W, H, N = 5760, 4092, 10
    r = MovieRecipe(W, H, N, 20000, seed=1234, dose=1.0, local_max=3.0)
    frames = ctx.empty((N, H, W), np.float32)
    ctx.synth_render(frames, r)
    pg = L.default_params(patch=(5, 5), bft=(500, 100), pix_size=0.83, kv=300, fm_dose=1.0)
    res = ctx.align_movie(frames, pg)
    assert res["num_sums"] == 3
    g = res["global_shift"]
    d = r.displacement
    e = (g - g.mean(0)) + (d - d.mean(0))
    assert np.sqrt((e ** 2).sum(1).mean()) < 1.0
    for s in res["sums"]:
        a = s.numpy()
        assert np.isfinite(a).all() and a.min() >= 0
    raw = res["sums"][0].numpy().astype(np.float64)
    assert abs(raw.mean() / (N * 1.0) - 1.0) < 0.05
    assert res["patch_shift"].shape == (N, 25, 2)
    assert np.isfinite(res["patch_shift"]).all()


def test_fourier_binned_output(ctx):
    """-FtBin 2: every sum is Fourier-cropped to half size before the C2R
    (Correct/CCorrectFullShift.cpp:146-185); 1152 x 816 frames give a
    576 x 408 output (cropping inverse column pass and 288-point rows)."""
    W, H, N = 1152, 816, 6
    r = MovieRecipe(W, H, N, 150, seed=33, dose=2.0)
    mv = r.movie()
    pg, po = _params(iter=5, pix_size=1.0, kv=300, fm_dose=1.0, ft_bin=2.0)
    gres = ctx.align_movie(ctx.upload(mv), pg)
    assert (gres["sums"][0].shape[0], gres["sums"][0].shape[1]) == (408, 576)
    orep = O.align_movie(mv, po, replay_global=gres["global_shift"])
    assert orep["sums"].shape[1:] == (408, 576)
    for i in range(gres["num_sums"]):
        assert _rel(gres["sums"][i].numpy(), orep["sums"][i]) < 1e-4, i


@pytest.mark.parametrize("fmt", [1, 2])
def test_reduced_precision_stacks(ctx, fmt):
    """FP16 (1) and BF16 (2) storage of the xcf and patch spectrum stacks
    (FP32 arithmetic, per-transform scale) on the 5760 x 4092 geometry with
    5 x 5 patches: the shifts stay within the north-star bars of the FP32
    path, which the other tests pin to the FP64 oracle."""
    W, H, N = 5760, 4092, 8
    r = MovieRecipe(W, H, N, 20000, seed=77, dose=1.0, local_max=3.0)
    frames = ctx.empty((N, H, W), np.float32)
    ctx.synth_render(frames, r)
    kw = dict(patch=(5, 5), bft=(500, 100), pix_size=0.83, kv=300, fm_dose=1.0)
    ref = ctx.align_movie(frames, L.default_params(**kw))
    red = ctx.align_movie(frames, L.default_params(spectra_fmt=fmt, **kw))
    assert ref["spectra_fmt_used"] == (0, 0)
    assert red["spectra_fmt_used"] == (fmt, fmt)
    assert np.max(np.abs(red["global_shift"] - ref["global_shift"])) <= 0.05
    assert np.array_equal(red["patch_starts"], ref["patch_starts"])
    d = red["patch_shift"] - ref["patch_shift"]
    assert float(np.sqrt(np.mean(np.sum(d ** 2, axis=2)))) <= 0.1
    # the sums use FP32 full spectra: replaying the reduced run's shift
    # tables through the FP32 path gives the same sums (the nearest-pixel
    # warp turns small shift differences into discrete changes, so the sums
    # of the two runs differ by more than their shifts suggest)
    rep = ctx.align_movie(frames, L.default_params(**kw), replay_global=red["global_shift"],
                          replay_patch=red["patch_shift"], replay_centers=red["patch_center"],
                          replay_bad=red["patch_bad"])
    for a, b in zip(red["sums"], rep["sums"]):
        assert _rel(a.numpy(), b.numpy()) < 1e-5
    print("fmt %d: sum rel-L2 vs the FP32 run %.2e" % (fmt, _rel(red["sums"][0].numpy(),
                                                                  ref["sums"][0].numpy())))


def test_iteration_graphs_match_direct_launches(ctx):
    """The alignment iterations replayed from captured HIP graphs give the
    same bits as direct launches, on the capturing call and on later
    replays with direct work in between."""
    W, H, N = 1024, 1024, 12
    r = MovieRecipe(W, H, N, 400, seed=5, dose=4.0, local_max=2.0)
    mv = ctx.upload(r.movie())
    kw = dict(patch=(3, 3), iter=10, pix_size=1.0, kv=300, fm_dose=1.0)
    direct = ctx.align_movie(mv, L.default_params(launch_graphs=0, **kw))
    for rep in range(3):
        g = ctx.align_movie(mv, L.default_params(launch_graphs=1, **kw))
        assert np.array_equal(g["global_shift"], direct["global_shift"]), rep
        assert np.array_equal(g["patch_shift"], direct["patch_shift"]), rep
        assert g["global_iters"] == direct["global_iters"]
        for a, b in zip(g["sums"], direct["sums"]):
            assert np.array_equal(a.numpy(), b.numpy()), rep
        # direct work between replays
        ctx.align_movie(mv, L.default_params(launch_graphs=0, **kw))


@pytest.mark.parametrize("pipeline,patch,fmt", [(0, (3, 3), 0), (1, (3, 3), 0), (1, (0, 0), 0),
                                                  (1, (3, 3), 1)])
def test_batched_movies_match_single_calls(ctx, pipeline, patch, fmt):
    """mc_align_movies: three movies of one size form a group (one batch of
    alignment jobs, or a pipeline whose stack slots are reused by the third
    movie), a fourth of another size goes in its own group; every result is
    the one a separate mc_align_movie call gives, bit for bit."""
    kw = dict(patch=patch, iter=10, pix_size=1.0, kv=300, fm_dose=1.0, spectra_fmt=fmt)
    shapes = [(1024, 1024, 12, 11), (1152, 816, 6, 12), (1024, 1024, 12, 13), (1024, 1024, 12, 14)]
    movies = [ctx.upload(MovieRecipe(W, H, N, 300, seed=sd, dose=4.0, local_max=2.0).movie())
              for W, H, N, sd in shapes]
    prm = L.default_params(movie_pipeline=pipeline, **kw)
    single = [ctx.align_movie(m, prm) for m in movies]
    batched = ctx.align_movies(movies, prm)
    keys = ("global_shift",) + (("patch_shift", "patch_bad", "patch_center", "patch_starts")
                                if patch[0] > 1 else ())
    for a, b in zip(batched, single):
        for k in keys:
            assert np.array_equal(a[k], b[k]), k
        assert a["global_iters"] == b["global_iters"]
        for x, y in zip(a["sums"], b["sums"]):
            assert np.array_equal(x.numpy(), y.numpy())


@pytest.mark.parametrize("case", ["1024", "production"])
def test_uneven_movies_share_iteration_launches(ctx, case):
    """mc_align_movies without the pipeline, on movies of different sizes
    whose alignment stacks share the column length: the global iterations of
    all movies run as one batch of jobs, and so do their patch iterations,
    each job with its own frame count, group count and row length (1024 x
    1024 x 12 with 1536 x 1024 x 8: 1024-point xcf and 340-point patch
    columns; 5760 x 4092 x 6 with 4096 x 4096 x 8: 2048-point xcf and
    408-point patch columns, the register-pass correlation).  Every result
    equals a separate mc_align_movie call's, bit for bit."""
    if case == "1024":
        kw = dict(patch=(3, 3), iter=10, pix_size=1.0, kv=300, fm_dose=1.0)
        shapes = [(1024, 1024, 12, 21), (1536, 1024, 8, 22), (1024, 1024, 12, 23)]
        movies = [ctx.upload(MovieRecipe(W, H, N, 300, seed=sd, dose=4.0, local_max=2.0).movie())
                  for W, H, N, sd in shapes]
    else:
        kw = dict(patch=(5, 5), bft=(500, 100), pix_size=0.83, kv=300, fm_dose=1.0)
        shapes = [(5760, 4092, 6, 31), (4096, 4096, 8, 32)]
        movies = []
        for W, H, N, sd in shapes:
            fr = ctx.empty((N, H, W), np.float32)
            ctx.synth_render(fr, MovieRecipe(W, H, N, 20000, seed=sd, dose=1.0, local_max=3.0))
            movies.append(fr)
    prm = L.default_params(movie_pipeline=0, **kw)
    print("uneven batch", case, shapes, flush=True)
    single = [ctx.align_movie(m, prm) for m in movies]
    batched = ctx.align_movies(movies, prm)
    keys = ("global_shift", "patch_shift", "patch_bad", "patch_center", "patch_starts", "patch_iters")
    for a, b in zip(batched, single):
        for k in keys:
            assert np.array_equal(a[k], b[k]), k
        assert a["global_iters"] == b["global_iters"]
        for x, y in zip(a["sums"], b["sums"]):
            assert np.array_equal(x.numpy(), y.numpy())


@pytest.mark.parametrize("W,H,N", [(5760, 4092, 6), (2048, 2048, 36)])
def test_chunk_streams_match_single_stream(ctx, W, H, N):
    """Frame chunks on two streams give the same bits as on one stream.  With
    36 frames (chunks of 10, 10, 10 and 6) the early local-correction inverse
    takes the first three and leaves the last to the correction stream,
    which is checked here too."""
    r = MovieRecipe(W, H, N, 20000, seed=9, dose=1.0, local_max=3.0)
    frames = ctx.empty((N, H, W), np.float32)
    ctx.synth_render(frames, r)
    kw = dict(patch=(5, 5), bft=(500, 100), pix_size=0.83, kv=300, fm_dose=1.0)
    one = ctx.align_movie(frames, L.default_params(chunk_streams=1, **kw))
    two = ctx.align_movie(frames, L.default_params(chunk_streams=2, **kw))
    assert np.array_equal(one["global_shift"], two["global_shift"])
    assert np.array_equal(one["patch_shift"], two["patch_shift"])
    for a, b in zip(one["sums"], two["sums"]):
        assert np.array_equal(a.numpy(), b.numpy())


def test_super_resolution_fourier_binned(ctx):
    """An 11520 x 8184 super-resolution movie with -FtBin 2 (the 8184-point
    columns and 5760-point rows): 5760 x 4092 output sums, the xcf stack of
    the 5760 x 4092 case, and the ground-truth drift recovered."""
    W, H, N = 11520, 8184, 6
    r = MovieRecipe(W, H, N, 40000, seed=5150, dose=1.0, local_max=3.0)
    frames = ctx.empty((N, H, W), np.float32)
    ctx.synth_render(frames, r)
    prm = L.default_params(patch=(5, 5), bft=(500, 100), pix_size=0.415, kv=300, fm_dose=1.0,
                           ft_bin=2.0)
    res = ctx.align_movie(frames, prm)
    assert res["num_sums"] == 3
    assert tuple(res["sums"][0].shape) == (4092, 5760)
    g = res["global_shift"]
    d = r.displacement
    e = (g - g.mean(0)) + (d - d.mean(0))
    assert np.sqrt((e ** 2).sum(1).mean()) < 1.0
    for s_ in res["sums"]:
        a = s_.numpy()
        assert np.isfinite(a).all() and a.min() >= 0
    assert np.isfinite(res["patch_shift"]).all()


def test_frame_weights_do_not_race_the_patch_preparation(ctx):
    """With an xcf bin of 1 the xcf stack is the frame stack.  -InFmMotion and
    -CorrInterp weigh the frame spectra in place on the second stream, which
    must not run before the patch preparation has read them unweighted:
    two chunk streams give the bits of one."""
    W, H, N = 1024, 1024, 10
    r = MovieRecipe(W, H, N, 400, seed=31, dose=4.0, local_max=2.0)
    mv = ctx.upload(r.movie())
    kw = dict(patch=(3, 3), iter=10, pix_size=1.0, kv=300, fm_dose=1.0, in_fm_motion=1,
              corr_interp=1)
    one = ctx.align_movie(mv, L.default_params(chunk_streams=1, **kw))
    for rep in range(2):
        two = ctx.align_movie(mv, L.default_params(chunk_streams=2, **kw))
        assert np.array_equal(one["patch_shift"], two["patch_shift"]), rep
        assert np.array_equal(one["global_shift"], two["global_shift"]), rep
        for a, b in zip(one["sums"], two["sums"]):
            assert np.array_equal(a.numpy(), b.numpy()), rep
