"""Ingest on the GPU: raw frames to corrected FP32 frames against the oracle
(oracle/ingest_oracle.py), bit for bit (one FP32 subtraction and one
multiplication per pixel, as in the reference); reference post-processing;
asynchronous loads overlapping alignments; an MRC file through the whole
path."""
import numpy as np
import pytest

import ingest_oracle as IO
from motioncor3_amd import lib as L
from motioncor3_amd.mrc import pack4, write_mrc
from motioncor3_amd.synth import MovieRecipe

pytestmark = pytest.mark.gpu

_DT = {0: np.uint8, 5: np.uint8, 1: np.int16, 6: np.uint16, 2: np.float32}


def _raw(rng, mode, n, sy, sx):
    if mode == 101:
        vals = rng.integers(0, 16, (n, sy, sx)).astype(np.uint8)
        return pack4(vals), vals
    if mode == 2:
        v = rng.normal(3, 2, (n, sy, sx)).astype(np.float32)
        return v, v
    lo, hi = {0: (0, 255), 5: (0, 255), 1: (-3000, 3000), 6: (0, 65535)}[mode]
    v = rng.integers(lo, hi + 1, (n, sy, sx)).astype(_DT[mode])
    return v, v


@pytest.mark.parametrize("mode", [0, 5, 1, 6, 2, 101])
@pytest.mark.parametrize("refs", ["none", "gain", "dark", "both"])
def test_apply_refs_matches_oracle(ctx, mode, refs):
    rng = np.random.default_rng(mode * 7 + len(refs))
    n, sy, sx = 3, 37, 301
    W, H = 290, 30   # centre crop (odd offsets in x)
    raw, vals = _raw(rng, mode, n, sy, sx)
    gain = rng.uniform(0.5, 1.5, (sy, sx)).astype(np.float32) if refs in ("gain", "both") else None
    dark = rng.uniform(0.0, 2.0, (sy, sx)).astype(np.float32) if refs in ("dark", "both") else None
    ctx.set_refs(gain, dark)
    try:
        pitch = (W // 2 + 1) * 2
        fr = ctx.empty((n, H, pitch), np.float32)
        ctx.load_frames(raw, mode, sx, sy, n, fr, width=W)
        got = fr.numpy()[:, :, :W]
        want = IO.apply_refs(vals, gain, dark, W, H)
        assert np.array_equal(got, want)
    finally:
        ctx.set_refs(None, None)


@pytest.mark.parametrize("rot,flip,inv", [(1, 0, 0), (2, 1, 0), (3, 2, 1), (0, 1, 0), (0, 0, 1)])
@pytest.mark.parametrize("quirks", [True, False])
def test_reference_post_processing(ctx, rot, flip, inv, quirks):
    rng = np.random.default_rng(rot * 9 + flip * 3 + inv)
    g = rng.uniform(0.5, 1.5, (11, 14)).astype(np.float32)
    g[2, 3] = 0.0
    d = rng.uniform(0, 1, (11, 14)).astype(np.float32)
    ctx.set_refs(g, d, rot=rot, flip=flip, inv=inv, quirks=quirks)
    try:
        gg, dd = ctx.get_refs()
        wg, wd = IO.post_process(g, d, rot, flip, inv, quirks)
        assert np.array_equal(gg, wg) and np.array_equal(dd, wd)
    finally:
        ctx.set_refs(None, None)


def _same_alignment(a, b):
    assert np.array_equal(a["global_shift"], b["global_shift"])
    assert np.array_equal(a["patch_shift"], b["patch_shift"])
    for i in range(a["num_sums"]):
        assert np.array_equal(a["sums"][i].numpy(), b["sums"][i].numpy())


def test_async_loads_overlap_alignments(ctx):
    """Movie m+1 is loaded asynchronously (pinned host memory, ingest stream)
    while movie m is aligned; every alignment must equal the one of frames
    loaded synchronously."""
    W = H = 1024
    N = 8
    pg = L.default_params(patch=(3, 3), pix_size=1.0, kv=300, fm_dose=1.0)
    movies = [MovieRecipe(W, H, N, 400, seed=70 + k, dose=4.0, local_max=2.0).movie().astype(np.uint8)
              for k in range(3)]
    host = [ctx.host_array(m.shape, np.uint8) for m in movies]
    for h, m in zip(host, movies):
        h.a[...] = m
    gain = np.random.default_rng(3).uniform(0.9, 1.1, (H, W)).astype(np.float32)
    ctx.set_refs(gain, None)
    try:
        bufs = [ctx.empty((N, H, W), np.float32) for _ in range(2)]
        ref = []
        for m in range(3):
            ctx.load_frames(host[m], 0, W, H, N, bufs[0])
            ref.append(ctx.align_movie(bufs[0], pg))
        ctx.load_frames(host[0], 0, W, H, N, bufs[0], async_=True)
        got = []
        for m in range(3):
            if m + 1 < 3:
                ctx.load_frames(host[m + 1], 0, W, H, N, bufs[(m + 1) % 2], async_=True)
            got.append(ctx.align_movie(bufs[m % 2], pg))
        for a, b in zip(got, ref):
            _same_alignment(a, b)
    finally:
        ctx.set_refs(None, None)
        for h in host:
            h.free()


def test_async_loads_feed_the_movie_pipeline(ctx):
    """Three movies loaded asynchronously, then aligned by one pipelined
    mc_align_movies call: each movie's forward transform (on the second
    stream) must wait for its own load, and every result must equal the
    alignment of frames loaded synchronously."""
    W = H = 1024
    N = 8
    pg = L.default_params(patch=(3, 3), pix_size=1.0, kv=300, fm_dose=1.0, movie_pipeline=1)
    movies = [MovieRecipe(W, H, N, 400, seed=80 + k, dose=4.0, local_max=2.0).movie().astype(np.uint8)
              for k in range(3)]
    host = [ctx.host_array(m.shape, np.uint8) for m in movies]
    for h, m in zip(host, movies):
        h.a[...] = m
    gain = np.random.default_rng(4).uniform(0.9, 1.1, (H, W)).astype(np.float32)
    ctx.set_refs(gain, None)
    try:
        bufs = [ctx.empty((N, H, W), np.float32) for _ in range(3)]
        ref = []
        for m in range(3):
            ctx.load_frames(host[m], 0, W, H, N, bufs[0])
            ref.append(ctx.align_movie(bufs[0], pg))
        for m in range(3):
            ctx.load_frames(host[m], 0, W, H, N, bufs[m], async_=True)
        got = ctx.align_movies(bufs, pg)
        for a, b in zip(got, ref):
            _same_alignment(a, b)
    finally:
        ctx.set_refs(None, None)
        for h in host:
            h.free()


def test_mrc_file_through_the_path(ctx, tmp_path):
    """A 4-bit MRC stack read with mc_mrc_read into pinned memory, corrected
    and aligned, equals aligning the oracle-corrected frames."""
    W = H = 1024
    N = 8
    counts = MovieRecipe(W, H, N, 400, seed=81, dose=2.0, local_max=2.0).movie()
    counts = np.minimum(counts, 15).astype(np.uint8)
    path = str(tmp_path / "movie.mrc")
    write_mrc(path, counts, 101, ext_bytes=1024)
    info = L.mrc_info(path)
    host = ctx.host_array((L.raw_frame_bytes(101, W, H) * N,), np.uint8)
    try:
        L.mrc_read(path, 0, N, out=host)
        gain = np.random.default_rng(4).uniform(0.8, 1.2, (H, W)).astype(np.float32)
        dark = np.random.default_rng(5).uniform(0.0, 0.1, (H, W)).astype(np.float32)
        ctx.set_refs(gain, dark)
        fr = ctx.empty((N, H, W), np.float32)
        ctx.load_frames(host, info["mode"], info["sx"], info["sy"], info["n"], fr)
        pg = L.default_params(patch=(3, 3), pix_size=1.0, kv=300, fm_dose=1.0)
        a = ctx.align_movie(fr, pg)
        want = IO.apply_refs(counts, gain, dark, W, H)
        assert np.array_equal(fr.numpy(), want)
        b = ctx.align_movie(ctx.upload(want), pg)
        _same_alignment(a, b)
    finally:
        ctx.set_refs(None, None)
        host.free()


@pytest.mark.parametrize("mode", [0, 101])
def test_integrated_frames_match_oracle(ctx, mode):
    """-FmIntFile ingest: unsigned-byte sums of raw frames (wrapping at 256),
    then dark and gain."""
    rng = np.random.default_rng(40 + mode)
    n, sy, sx = 11, 33, 70
    raw, vals = _raw(rng, mode, n, sy, sx)
    if mode == 0:
        vals = rng.integers(100, 256, (n, sy, sx)).astype(np.uint8)  # sums wrap
        raw = vals
    start = np.array([0, 2, 5, 9], np.int32)
    size = np.array([2, 3, 4, 2], np.int32)
    gain = rng.uniform(0.5, 1.5, (sy, sx)).astype(np.float32)
    dark = rng.uniform(0.0, 2.0, (sy, sx)).astype(np.float32)
    ctx.set_refs(gain, dark)
    try:
        W, H = 64, 31
        fr = ctx.empty((4, H, W), np.float32)
        ctx.load_frames(raw, mode, sx, sy, n, fr, int_start=start, int_size=size)
        summed = np.stack([(vals[s:s + k].astype(np.uint32).sum(0) & 0xFF).astype(np.uint8)
                           for s, k in zip(start, size)])
        assert np.array_equal(fr.numpy(), IO.apply_refs(summed, gain, dark, W, H))
    finally:
        ctx.set_refs(None, None)


def test_integrated_alignment_matches_oracle(ctx):
    """Frames integrated from raw frames (sizes 2 and 3): frame centres,
    grouping by raw-frame count and the accumulated dose of the dose
    weighting follow the table; GPU against the oracle with the same table."""
    import oracle as O
    W = H = 1024
    n_raw = 30
    t = L.fm_int_table(None, n_raw, 0.5)
    r = MovieRecipe(W, H, n_raw, 400, seed=97, dose=1.0, local_max=2.0)
    raw = np.minimum(r.movie(), 255).astype(np.uint8)
    start = np.array(list(range(0, 12, 2)) + list(range(12, 30, 3)), np.int32)
    size = np.array([2] * 6 + [3] * 6, np.int32)
    N = len(size)
    fr = ctx.empty((N, H, W), np.float32)
    ctx.load_frames(raw, 0, W, H, n_raw, fr, int_start=start, int_size=size)
    frames = fr.numpy()
    kw = dict(patch=(3, 3), pix_size=1.0, kv=300, fm_dose=0.5)
    pg = L.default_params(fm_int_start=start, fm_int_size=size, **kw)
    po = O.default_params(fm_int_start=start, fm_int_size=size, **kw)
    g = ctx.align_movie(fr, pg)
    own = O.align_movie(frames, po)
    assert np.array_equal(own["patch_starts"], g["patch_starts"])
    assert np.max(np.abs(g["global_shift"] - own["global_shift"])) <= 0.05
    assert np.max(np.abs(g["patch_shift"] - own["patch_shift"])) <= 0.05
    rep = O.align_movie(frames, po, replay_global=g["global_shift"], replay_patch=g["patch_shift"],
                        replay_centers=g["patch_center"], replay_bad=g["patch_bad"])
    for i in range(g["num_sums"]):
        s, o = g["sums"][i].numpy(), rep["sums"][i]
        assert np.linalg.norm(s - o) / np.linalg.norm(o) < 1e-4, i
    # the dose weighting really used the table: the plain-dose result differs
    plain = ctx.align_movie(fr, L.default_params(**kw), replay_global=g["global_shift"],
                            replay_patch=g["patch_shift"], replay_centers=g["patch_center"],
                            replay_bad=g["patch_bad"])
    assert not np.array_equal(plain["sums"][1].numpy(), g["sums"][1].numpy())
    assert t["size"].tolist() == [1] * n_raw


# ---- EER ----

def _eer_movie(rng, cam_w, cam_h, nfrac, bits, per_frac):
    import eer_oracle as E
    cam = cam_w * cam_h
    streams = []
    for _ in range(nfrac):
        q = np.sort(rng.choice(cam, size=per_frac, replace=False))
        if bits == 7:
            sub = rng.integers(0, 16, per_frac)
            streams.append(E.encode_7bit(list(zip(q.tolist(), sub.tolist())), cam))
        else:
            streams.append(E.encode_8bit(q.tolist(), cam))
    off = np.cumsum([0] + [len(s) for s in streams]).astype(np.int64)
    return np.concatenate(streams), off


@pytest.mark.parametrize("bits,sampling", [(7, 1), (7, 2), (7, 3), (8, 1), (8, 2)])
@pytest.mark.parametrize("integrate", [False, True])
def test_eer_render_matches_oracle(ctx, bits, sampling, integrate):
    """mc_load_eer renders EER fractions on the GPU (one thread per fraction
    and rendered frame) exactly as the oracle restates CDecodeEerFrame and
    CRenderMrcStack, then applies the gain and the centre crop; the 8-bit
    super-resolution case follows the reference's byte-0 walk."""
    import eer_oracle as E
    rng = np.random.default_rng(10 * bits + sampling)
    cam_w, cam_h = (4096, 6) if sampling > 1 else (96, 64)
    nfrac = 12
    data, off = _eer_movie(rng, cam_w, cam_h, nfrac, bits, 700)
    f = E.factor(sampling)
    fw, fh = cam_w * f, cam_h * f
    ist, isz = ([0, 2, 5, 9], [2, 3, 4, 3]) if integrate else (None, None)
    want = E.render(data, off, cam_w, cam_h, bits, sampling, ist, isz)
    gain = rng.uniform(0.9, 1.1, (fh, fw)).astype(np.float32)
    W, H = fw - 6, fh - 2
    n_out = want.shape[0]
    frames = ctx.empty((n_out, H, W), np.float32)
    ctx.set_refs(gain, None)
    try:
        ctx.load_eer(data, off, cam_w, cam_h, bits, sampling, frames, int_start=ist, int_size=isz)
    finally:
        ctx.set_refs(None, None)
    exp = IO.apply_refs(want, gain, None, W, H)
    assert np.array_equal(frames.numpy(), exp)
    assert want.sum() > 0


def test_eer_counts_wrap_at_256(ctx):
    """300 fractions integrated into one frame: a pixel hit by all of them
    reads 300 mod 256, as the reference's unsigned-byte accumulation."""
    import eer_oracle as E
    cam_w, cam_h = 64, 16
    streams = [E.encode_7bit([(7, 0), (9 + (k % 3), 0)], cam_w * cam_h) for k in range(300)]
    off = np.cumsum([0] + [len(s) for s in streams]).astype(np.int64)
    data = np.concatenate(streams)
    frames = ctx.empty((1, cam_h, cam_w), np.float32)
    ctx.load_eer(data, off, cam_w, cam_h, 7, 1, frames, int_start=[0], int_size=[300])
    got = frames.numpy()[0]
    assert got[0, 7] == 300 % 256
    assert got[0, 9] == 100 and got[0, 10] == 100 and got[0, 11] == 100
    assert np.array_equal(got, E.render(data, off, cam_w, cam_h, 7, 1, [0], [300])[0].astype(np.float32))


def test_eer_file_through_the_path(ctx, tmp_path):
    """An EER file (TIFF container, 7-bit codes) of a drifting synthetic
    movie: read, rendered with a frame-integration table asynchronously, then
    aligned; the alignment equals the one of the oracle-rendered frames."""
    import eer_oracle as E
    import tiffw
    rng = np.random.default_rng(21)
    W = H = 512
    nfrac, per = 40, 16
    truth = MovieRecipe(W, H, nfrac, 300, seed=77, dose=4.0, local_max=0.0).movie()
    streams = []
    for k in range(nfrac):
        p = truth[k].ravel().astype(np.float64)
        p /= p.sum()
        q = np.unique(rng.choice(W * H, size=W * H // per, p=p))
        streams.append(E.encode_7bit([(int(v), 0) for v in q], W * H))
    path = str(tmp_path / "m.eer")
    tiffw.write(path, [dict(w=W, h=H, bits=1, comp=65001, rps=H, blocks=[s.tobytes()])
                       for s in streams])
    info = L.eer_info(path)
    assert (info["cam_w"], info["cam_h"], info["n"], info["bits"]) == (W, H, nfrac, 7)
    data, off = L.eer_read(path)
    ist, isz = list(range(0, nfrac, 4)), [4] * (nfrac // 4)
    host = ctx.host_array((len(data),), np.uint8)
    host.a[...] = data
    try:
        frames = ctx.empty((len(isz), H, W), np.float32)
        ctx.load_eer(host, off, W, H, 7, 1, frames, async_=True, int_start=ist, int_size=isz)
        prm = L.default_params(patch=(0, 0), pix_size=1.0, kv=300, fm_dose=1.0)
        got = ctx.align_movie(frames, prm)
        ref_frames = E.render(data, off, W, H, 7, 1, ist, isz).astype(np.float32)
        ref = ctx.align_movie(ctx.upload(ref_frames), prm)
        assert np.array_equal(got["global_shift"], ref["global_shift"])
        for a, b in zip(got["sums"], ref["sums"]):
            assert np.array_equal(a.numpy(), b.numpy())
    finally:
        host.free()


def test_tiff_file_through_the_path(ctx, tmp_path):
    """An 8-bit LZW TIFF movie (Pillow's libtiff writer) decoded by
    mc_tiff_read into pinned memory, loaded with a gain reference and
    aligned: the same result as the flipped frames loaded from memory."""
    from PIL import Image
    W = H = 512
    N = 8
    mv = MovieRecipe(W, H, N, 300, seed=31, dose=4.0, local_max=0.0).movie()
    fr = np.clip(mv, 0, 255).astype(np.uint8)
    path = str(tmp_path / "m.tif")
    ims = [Image.fromarray(f) for f in fr]
    ims[0].save(path, save_all=True, append_images=ims[1:], compression="tiff_lzw")
    info = L.tiff_info(path)
    assert (info["sx"], info["sy"], info["n"], info["mode"]) == (W, H, N, 0)
    host = ctx.host_array((N, H, W), np.uint8)
    gain = np.random.default_rng(2).uniform(0.95, 1.05, (H, W)).astype(np.float32)
    ctx.set_refs(gain, None)
    try:
        L.tiff_read(path, out=host)
        assert np.array_equal(host.a, fr[:, ::-1])
        a = ctx.empty((N, H, W), np.float32)
        b = ctx.empty((N, H, W), np.float32)
        ctx.load_frames(host, 0, W, H, N, a)
        ctx.load_frames(np.ascontiguousarray(fr[:, ::-1]), 0, W, H, N, b)
        assert np.array_equal(a.numpy(), IO.apply_refs(fr[:, ::-1], gain, None, W, H))
        prm = L.default_params(patch=(0, 0), pix_size=1.0, kv=300, fm_dose=1.0)
        ra, rb = ctx.align_movie(a, prm), ctx.align_movie(b, prm)
        assert np.array_equal(ra["global_shift"], rb["global_shift"])
    finally:
        ctx.set_refs(None, None)
        host.free()


@pytest.mark.parametrize("f", [2, 4])
def test_references_enlarged_for_larger_frames(ctx, f):
    """A gain and dark f times smaller than the frames are enlarged by
    repeating each pixel f x f (CLoadRefs::AugmentRefs), for MRC frames and
    for EER frames rendered at super-resolution from a camera-size gain."""
    import eer_oracle as E
    rng = np.random.default_rng(40 + f)
    rh, rw = 12, 20
    gain = rng.uniform(0.9, 1.1, (rh, rw)).astype(np.float32)
    dark = rng.uniform(0.0, 2.0, (rh, rw)).astype(np.float32)
    big_g = np.repeat(np.repeat(gain, f, 0), f, 1)
    big_d = np.repeat(np.repeat(dark, f, 0), f, 1)
    sy, sx, n = rh * f, rw * f, 3
    raw = rng.integers(0, 255, (n, sy, sx)).astype(np.uint8)
    ctx.set_refs(gain, dark)
    try:
        out = ctx.empty((n, sy - 2, sx - 4), np.float32)
        ctx.load_frames(raw, 0, sx, sy, n, out)
        assert np.array_equal(out.numpy(), IO.apply_refs(raw, big_g, big_d, sx - 4, sy - 2))
        # a size that is not an integer multiple is refused
        with pytest.raises(RuntimeError):
            ctx.load_frames(raw[:, :, :sx - 1].copy(), 0, sx - 1, sy, n, ctx.empty((n, sy, sx - 1), np.float32))
    finally:
        ctx.set_refs(None, None)
    # EER: camera 4096 x 3, gain of camera size, rendering at 2x / 4x
    cam_w, cam_h = 4096, 3
    sampling = 2 if f == 2 else 3
    cg = rng.uniform(0.9, 1.1, (cam_h, cam_w)).astype(np.float32)
    data, off = _eer_movie(rng, cam_w, cam_h, 3, 7, 500)
    want = E.render(data, off, cam_w, cam_h, 7, sampling)
    ctx.set_refs(cg, None)
    try:
        fr = ctx.empty((3, cam_h * f, cam_w * f), np.float32)
        ctx.load_eer(data, off, cam_w, cam_h, 7, sampling, fr)
        exp = IO.apply_refs(want, np.repeat(np.repeat(cg, f, 0), f, 1), None, cam_w * f, cam_h * f)
        assert np.array_equal(fr.numpy(), exp)
    finally:
        ctx.set_refs(None, None)
