"""Outputs on the GPU path against the oracle: MRC sums with header
statistics; an .aln file written after an alignment and replayed (-InAln);
the -OutStack aligned stack and its _Stk.mrc file; -Align 0 simple sums; the
-InFmMotion / -CorrInterp frame weights."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
from motioncor3_amd import lib as L
from motioncor3_amd.synth import MovieRecipe

pytestmark = pytest.mark.gpu


def _read_mrc(path):
    raw = open(path, "rb").read()
    hi = np.frombuffer(raw[:1024], np.int32)
    hf = np.frombuffer(raw[:1024], np.float32)
    nx, ny, nz, mode = hi[:4]
    data = np.frombuffer(raw[1024 + hi[23]:], np.float32).reshape(nz, ny, nx)
    return hi, hf, data


def test_save_sums_mrc(ctx, tmp_path):
    W, H, N = 512, 512, 10
    mv = ctx.upload(MovieRecipe(W, H, N, 200, seed=91, dose=4.0).movie())
    pg = L.default_params(pix_size=1.1, kv=300, fm_dose=1.0, split_sum=1, ft_bin=1.0)
    res = ctx.align_movie(mv, pg)
    assert res["num_sums"] == 5
    out = str(tmp_path / "corr.MRC")
    ctx.save_sums(pg, res, out)
    names = ["corr", "corr_DW", "corr_DWS", "corr_ODD", "corr_EVN"]
    for i, nm in enumerate(names):
        path = str(tmp_path / (nm + ".mrc"))
        assert os.path.exists(path), nm
        hi, hf, data = _read_mrc(path)
        want = res["sums"][i].numpy()
        assert tuple(hi[:4]) == (W, H, 1, 2) and tuple(hi[7:10]) == (W, H, 1)
        assert np.allclose(hf[10:13], [W * np.float32(1.1), H * np.float32(1.1), np.float32(1.1)])
        assert tuple(hi[16:19]) == (1, 2, 3)
        assert np.array_equal(data[0], want)
        assert hf[19] == want.min() and hf[20] == want.max()
        assert abs(hf[21] - want.astype(np.float64).mean()) <= 1e-6 * abs(want.mean())


def test_aln_replay_matches_oracle(ctx, tmp_path):
    """-OutAln then -InAln: the GPU path replays the loaded shifts, and the
    sums agree with the oracle replaying the same loaded values."""
    W, H, N = 1024, 1024, 8
    mvh = MovieRecipe(W, H, N, 400, seed=93, dose=4.0, local_max=2.0).movie()
    mv = ctx.upload(mvh)
    kw = dict(patch=(3, 3), pix_size=1.0, kv=300, fm_dose=1.0)
    pg = L.default_params(**kw)
    res = ctx.align_movie(mv, pg)
    path = str(tmp_path / "movie.aln")
    L.save_aln(path, pg, W, H, N, res["global_shift"], res["patch_shift"], res["patch_center"])
    a = L.load_aln(path)
    assert a["n"] == N and a["patch"] == (3, 3)
    g = ctx.align_movie(mv, pg, replay_global=a["global_shift"], replay_patch=a["patch_shift"],
                        replay_centers=a["patch_center"])
    assert np.array_equal(g["global_shift"], a["global_shift"])
    po = O.default_params(**kw)
    o = O.align_movie(mvh, po, replay_global=a["global_shift"], replay_patch=a["patch_shift"],
                      replay_centers=a["patch_center"])
    for i in range(g["num_sums"]):
        s, r = g["sums"][i].numpy(), o["sums"][i]
        assert np.linalg.norm(s - r) / np.linalg.norm(r) < 1e-4, i


@pytest.mark.parametrize("patch,ft_bin,zbin", [(False, 1.0, 1), (False, 2.0, 3), (True, 1.0, 1),
                                               (True, 2.0, 4)])
def test_aligned_stack_matches_oracle(ctx, patch, ft_bin, zbin):
    """-OutStack: every corrected frame (global phase shift, or local warp in
    patch mode) before dose weighting at the -FtBin size, z-binned, against
    the oracle replaying the same shifts."""
    if patch:
        W, H, N = 1024, 1024, 8
        kw = dict(patch=(3, 3))
    else:
        W, H, N = 512, 384, 10
        kw = {}
    kw.update(pix_size=1.0, kv=300, fm_dose=1.0, ft_bin=ft_bin, out_stack=(1, zbin))
    mvh = MovieRecipe(W, H, N, 300, seed=95, dose=4.0, local_max=2.0 if patch else 0.0).movie()
    mv = ctx.upload(mvh)
    pg = L.default_params(**kw)
    res = ctx.align_movie(mv, pg)
    rep = dict(replay_global=res["global_shift"])
    if patch:
        rep.update(replay_patch=res["patch_shift"], replay_centers=res["patch_center"])
    g = ctx.align_movie(mv, pg, **rep)
    o = O.align_movie(mvh, O.default_params(**kw), **rep)
    st, want = g["stack"].numpy(), o["stack"]
    assert st.shape == want.shape == (ctx.stack_frames(pg, N),) + g["sums"][0].shape
    assert st.shape[0] == -(-N // zbin)
    for k in range(st.shape[0]):
        d = st[k] - want[k]
        # the local warp's nearest-pixel gather (Correct/GCorrectPatchShift.cu:52-65)
        # may pick the neighbouring pixel where an FP32 coordinate sits on a
        # pixel edge; such flips stay below 1 in 1000 pixels
        flip = np.abs(d) > 1e-3 * np.sqrt(np.mean(want[k] ** 2))
        assert flip.mean() < (1e-3 if patch else 1e-9), (k, flip.mean())
        assert np.linalg.norm(d[~flip]) / np.linalg.norm(want[k][~flip]) < 2e-5, k
    # the stack is the unclamped, unweighted sum split by frame
    s0 = g["sums"][0].numpy()
    assert np.linalg.norm(np.maximum(st.astype(np.float64).sum(0), 0) - s0) / np.linalg.norm(s0) < 1e-5


def test_save_stack_mrc(ctx, tmp_path):
    W, H, N = 384, 256, 7
    mv = ctx.upload(MovieRecipe(W, H, N, 100, seed=97, dose=4.0).movie())
    pg = L.default_params(pix_size=0.8, ft_bin=2.0, out_stack=(1, 2))
    res = ctx.align_movie(mv, pg)
    st = res["stack"].numpy()
    assert st.shape == (4, H // 2, W // 2)
    ctx.save_stack(pg, res, str(tmp_path / "corr.mrc"))
    path = str(tmp_path / "corr_Stk.mrc")
    raw = open(path, "rb").read()
    hi, hf, data = _read_mrc(path)
    pix = np.float32(0.8) * np.float32(2.0)
    assert tuple(hi[:4]) == (W // 2, H // 2, 4, 2) and tuple(hi[7:10]) == (W // 2, H // 2, 4)
    assert np.allclose(hf[10:13], [W // 2 * pix, H // 2 * pix, 4 * pix])
    assert hi[23] == 4 * 32 * 4 and np.frombuffer(raw[128:132], np.int16)[1] == 32
    ext = np.frombuffer(raw[1024:1024 + hi[23]], np.float32).reshape(4, 32)
    assert np.all(ext[:, 11] == pix) and np.count_nonzero(ext) == 4
    assert np.array_equal(data, st)
    assert hf[19] == st.min() and hf[20] == st.max()
    means = [np.float32(f.astype(np.float64).mean()) for f in st]
    assert abs(hf[21] - np.mean(np.array(means, np.float64))) <= 1e-6 * abs(np.mean(means))


@pytest.mark.parametrize("ft_bin,zbin", [(1.0, 1), (2.0, 2)])
def test_simple_sums_match_oracle(ctx, tmp_path, ft_bin, zbin):
    """-Align 0: sums without alignment (dose weighting does not apply), the
    _ODD / _EVN split, and the stack, against the oracle."""
    W, H, N = 512, 384, 7
    mvh = MovieRecipe(W, H, N, 200, seed=99, dose=4.0).movie()
    kw = dict(align=0, split_sum=1, ft_bin=ft_bin, out_stack=(1, zbin), pix_size=1.0, kv=300,
              fm_dose=1.0)
    pg = L.default_params(**kw)
    assert ctx.output_geometry(pg, W, H, N)[0] == 3
    g = ctx.align_movie(ctx.upload(mvh), pg)
    o = O.align_movie(mvh, O.default_params(**kw))
    assert g["num_sums"] == 3 and g["global_iters"] == 0
    assert not g["global_shift"].any()
    for i in range(3):
        s, r = g["sums"][i].numpy(), o["sums"][i]
        assert s.shape == r.shape
        assert np.linalg.norm(s - r) / np.linalg.norm(r) < 1e-5, i
    st, want = g["stack"].numpy(), o["stack"]
    assert st.shape == want.shape and st.shape[0] == -(-N // zbin)
    for k in range(st.shape[0]):
        assert np.linalg.norm(st[k] - want[k]) / np.linalg.norm(want[k]) < 1e-5, k
    ctx.save_sums(pg, g, str(tmp_path / "simple.mrc"))
    for i, nm in enumerate(["simple", "simple_ODD", "simple_EVN"]):
        _, _, data = _read_mrc(str(tmp_path / (nm + ".mrc")))
        assert np.array_equal(data[0], g["sums"][i].numpy()), nm
    assert not os.path.exists(str(tmp_path / "simple_DW.mrc"))


@pytest.mark.parametrize("patch", [False, True])
def test_in_frame_motion_and_interp_weights_match_oracle(ctx, patch):
    """-InFmMotion (and -CorrInterp in patch mode): the corrected frames are
    weighted before the sums and the stack, against the oracle replaying the
    same shifts."""
    if patch:
        W, H, N = 1024, 1024, 8
        kw = dict(patch=(3, 3), corr_interp=1)
    else:
        W, H, N = 512, 384, 10
        kw = {}
    kw.update(pix_size=1.0, kv=300, fm_dose=1.0, in_fm_motion=1, out_stack=(1, 1))
    mvh = MovieRecipe(W, H, N, 300, seed=101, dose=4.0, step_sigma=1.5,
                      local_max=2.0 if patch else 0.0).movie()
    mv = ctx.upload(mvh)
    pg = L.default_params(**kw)
    res = ctx.align_movie(mv, pg)
    assert (np.abs(np.diff(res["global_shift"], axis=0)).max(1) > 0.1).any()
    rep = dict(replay_global=res["global_shift"])
    if patch:
        rep.update(replay_patch=res["patch_shift"], replay_centers=res["patch_center"],
                   replay_bad=res["patch_bad"])
    g = ctx.align_movie(mv, pg, **rep)
    # the measured run weighs on the second stream beside the patch alignment;
    # the replay of its shifts and flags weighs in order: same sums
    for i in range(g["num_sums"]):
        a, b = res["sums"][i].numpy(), g["sums"][i].numpy()
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-5, i
    o = O.align_movie(mvh, O.default_params(**kw), **rep)
    plain = O.align_movie(mvh, O.default_params(**dict(kw, in_fm_motion=0, corr_interp=0)), **rep)
    for i in range(g["num_sums"]):
        s, r = g["sums"][i].numpy(), o["sums"][i]
        assert np.linalg.norm(s - r) / np.linalg.norm(r) < 1e-4, i
        assert np.linalg.norm(plain["sums"][i] - r) / np.linalg.norm(r) > 1e-3, i
    st, want = g["stack"].numpy(), o["stack"]
    for k in range(N):
        d = st[k] - want[k]
        flip = np.abs(d) > 1e-3 * np.sqrt(np.mean(want[k] ** 2))
        assert flip.mean() < (1e-3 if patch else 1e-9), (k, flip.mean())
        assert np.linalg.norm(d[~flip]) / np.linalg.norm(want[k][~flip]) < 2e-5, k


def test_stack_arguments_are_checked(ctx, tmp_path):
    """A result without a stack buffer under -OutStack, and saving a result
    with no stack, fail with an error instead of writing anywhere."""
    W, H, N = 256, 256, 4
    mv = ctx.upload(MovieRecipe(W, H, N, 50, seed=103, dose=4.0).movie())
    pg = L.default_params(out_stack=(1, 1))
    for align in (1, 0):
        pg.align = align
        mvs, res, out = ctx._prepare(mv, pg, None, None)
        res.stack = None
        with pytest.raises(RuntimeError, match="stack"):
            ctx._check(ctx.lib.mc_align_movie(ctx.ptr, ctypes.byref(pg), ctypes.byref(mvs), None,
                                              ctypes.byref(res)), "mc_align_movie")
    r = L.McResult()
    r.out_w, r.out_h = W, H
    with pytest.raises(RuntimeError, match="stack"):
        ctx._check(ctx.lib.mc_save_stack(ctx.ptr, ctypes.byref(pg), ctypes.byref(r), 1,
                                         str(tmp_path / "x.mrc").encode()), "mc_save_stack")
    assert not os.path.exists(str(tmp_path / "x_Stk.mrc"))


@pytest.mark.parametrize("patch", [False, True])
@pytest.mark.parametrize("pipeline", [1, 0])
def test_stack_of_batched_movies_matches_single_calls(ctx, pipeline, patch):
    """mc_align_movies fills every movie's stack as separate calls do, with
    the frame weights on (in patch mode the pipelined calls weigh on the
    second stream beside the patch alignment)."""
    W, H, N = (1024, 1024, 6) if patch else (512, 384, 6)
    movies = [ctx.upload(MovieRecipe(W, H, N, 150, seed=105 + k, dose=4.0,
                                     local_max=2.0 if patch else 0.0).movie()) for k in range(3)]
    kw = dict(patch=(3, 3), corr_interp=1) if patch else {}
    pg = L.default_params(out_stack=(1, 2), pix_size=1.0, kv=300, fm_dose=1.0, in_fm_motion=1,
                          movie_pipeline=pipeline, **kw)
    batch = ctx.align_movies(movies, pg)
    for k, mv in enumerate(movies):
        one = ctx.align_movie(mv, pg)
        assert np.array_equal(batch[k]["stack"].numpy(), one["stack"].numpy()), k
        assert np.array_equal(batch[k]["sums"][1].numpy(), one["sums"][1].numpy()), k


@pytest.mark.parametrize("patch", [False, True])
def test_shift_logs(ctx, tmp_path, patch):
    """-LogDir: the log files named as CAlignMain::mCreateLogFile does
    (Align/CAlignMain.cpp:77-107) hold the result's shifts in the printf
    formats of CFullAlign::LogShift (Align/CFullAlign.cpp:173-194) and
    CPatchShifts::LogFullShifts / LogPatchShifts / LogFrameShifts
    (Align/CPatchShifts.cpp:155-249)."""
    W, H, N = (1024, 1024, 6) if patch else (512, 384, 6)
    mv = ctx.upload(MovieRecipe(W, H, N, 200, seed=107, dose=4.0, local_max=2.0 if patch else 0.0).movie())
    kw = dict(patch=(3, 3)) if patch else {}
    pg = L.default_params(log_dir=str(tmp_path) + "/", movie_name="/data/Movie_007.TIFF", **kw)
    r = ctx.align_movie(mv, pg)
    stem = str(tmp_path / "Movie_007")
    gs = r["global_shift"]
    full = "".join("%4d  %8.2f  %8.2f\n" % (i + 1, gs[i, 0], gs[i, 1]) for i in range(N))
    if not patch:
        assert open(stem + "-Full.log").read() == "# full frame alignment\n" + full
        assert not os.path.exists(stem + "-Patch-Full.log")
        return
    assert open(stem + "-Patch-Full.log").read() == (
        "# Patch based alignment\n# Initial alignment based upon full frames\n\n" + full)
    ps, pc, bad = r["patch_shift"], r["patch_center"], r["patch_bad"]
    P = ps.shape[1]

    def row(j, i):
        return "%4d %8.2f %8.2f " % (j + 1, pc[i, 0], pc[i, 1]) + \
            "%8.2f %8.2f %4d\n" % (ps[j, i, 0], ps[j, i, 1], bad[j, i])

    head = "# Patch based alignment\n# Number of patches: %d\n" % P
    size = "# Movie size: %d  %d  %d\n\n" % (W, H, N)
    per_patch = "".join("#Patch %03d raw and fit shifts\n" % (i + 1) +
                        "".join(row(j, i) for j in range(N)) + "\n" for i in range(P))
    per_frame = "".join("".join(row(j, i) for i in range(P)) + "\n" for j in range(N))
    assert open(stem + "-Patch-Patch.log").read() == (
        head + "# Shifts are listed per patch.\n" + size + per_patch)
    assert open(stem + "-Patch-Frame.log").read() == (
        head + "# Shifts are listed per frame.\n" + size + per_frame)
