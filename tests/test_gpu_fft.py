"""Batched 2-D R2C/C2R on the GPU against numpy.fft in FP64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# frame, xcf, patch and feature-map sizes of the reference geometries
SIZES = [(64, 48), (128, 128), (340, 340), (576, 408), (292, 408), (410, 408), (1024, 1024),
         (2882, 2048), (4096, 4096), (5760, 4092), (11520, 8184)]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("W,H", SIZES)
def test_rfft2_matches_numpy(ctx, W, H):
    print("rfft2", W, H, flush=True)
    B = 2 if W * H < 5_000_000 else 1
    rng = np.random.default_rng(W + H)
    x = rng.standard_normal((B, H, W)).astype(np.float32)
    spec = ctx.rfft2(ctx.upload(x)).numpy()
    for b in range(B):
        ref = np.fft.rfft2(x[b].astype(np.float64))
        assert _rel(spec[b].T, ref) < 2e-6, (W, H)


@pytest.mark.parametrize("W,H", SIZES)
def test_irfft2_roundtrip(ctx, W, H):
    print("irfft2", W, H, flush=True)
    rng = np.random.default_rng(7 * W + H)
    x = rng.standard_normal((1, H, W)).astype(np.float32)
    d = ctx.upload(x)
    spec = ctx.rfft2(d, scale=1.0 / (W * H))
    back = ctx.irfft2(spec, W).numpy()
    assert _rel(back, x) < 2e-6
    # C2R of a non-Hermitian half spectrum follows numpy's irfft2 convention
    s = spec.numpy()[0].T.astype(np.complex128)
    s[0, 0] += 0.3j
    spec2 = ctx.upload(np.ascontiguousarray(s.T[None]).astype(np.complex64))
    out = ctx.irfft2(spec2, W).numpy()[0]
    ref = np.fft.irfft2(s, s=(H, W)) * (W * H)
    assert _rel(out, ref) < 2e-6


def test_clamped_c2r(ctx):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 64, 96)).astype(np.float32)
    spec = ctx.rfft2(ctx.upload(x), scale=1.0 / (64 * 96))
    out = ctx.irfft2(spec, 96, clamp=True).numpy()
    assert np.allclose(out, np.maximum(x, 0), atol=1e-5)


@pytest.mark.parametrize("W,H", [(5760, 4092), (576, 408), (2882, 2048)])
@pytest.mark.parametrize("clamp", [0, 1])
def test_c2r_odd_pitch_and_unaligned_output(ctx, W, H, clamp):
    """The C2R row passes store float pairs when the output allows it and
    single floats otherwise (odd pitch, 4-byte-aligned pointer).  Both forms
    must give the same bits and touch nothing outside the rows."""
    import ctypes as C
    print("c2r unaligned", W, H, clamp, flush=True)
    B = 2 if W * H < 5_000_000 else 1
    rng = np.random.default_rng(W + 3 * H + clamp)
    x = rng.standard_normal((B, H, W)).astype(np.float32)
    spec = ctx.rfft2(ctx.upload(x), scale=1.0 / (W * H))
    ref = ctx.irfft2(spec, W, clamp=bool(clamp)).numpy()
    P = W + 1
    n = B * H * P
    buf = ctx.upload(np.full(n + 2, np.nan, np.float32))
    ctx._check(ctx.lib.mc_irfft2(ctx.ptr, spec.vp(), W, H, B, buf.vp(1), C.c_longlong(P),
                                 C.c_longlong(H * P), clamp), "mc_irfft2")
    ctx.sync()
    got = buf.numpy()
    assert np.isnan(got[0]) and np.isnan(got[n + 1])
    rows = got[1:n + 1].reshape(B, H, P)
    assert np.all(np.isnan(rows[:, :, W]))
    assert np.array_equal(rows[:, :, :W], ref)


def _radix(name):
    return name.split("radix ")[1].split()[0]


@pytest.mark.parametrize("W,H", [(5760, 4092), (2882, 2048), (576, 408), (8192, 4096)])
def test_tile_variants_are_bit_identical(W, H):
    """Every compiled tile variant of a register-pass row or column kernel
    with the default radix order computes each row or column with the same
    arithmetic, so forward and inverse transforms must agree bit for bit
    with the default tile.  Variants of another radix order are the same
    transform with other rounding: within 2e-6 relative L2."""
    from motioncor3_amd.lib import Context
    c = Context(0)
    try:
        rng = np.random.default_rng(W + H)
        fr = c.upload(rng.normal(1.0, 0.5, (2, H, W)).astype(np.float32))
        counts = [c.fft_variant(W, H, cols)[0] for cols in (0, 1)]
        assert max(counts) >= 2, counts  # each size here has variants on some side
        ref_s = c.rfft2(fr).numpy()
        ref_r = c.irfft2(c.upload(ref_s), W).numpy()
        for cols in (0, 1):
            if counts[cols] == 0:
                continue
            base = _radix(c.fft_variant_name(W, H, cols, 0))
            for v in range(1, counts[cols]):
                name = c.fft_variant_name(W, H, cols, v)
                c.fft_variant(W, H, cols, set=v)
                s = c.rfft2(fr).numpy()
                r = c.irfft2(c.upload(ref_s), W).numpy()
                if _radix(name) == base:
                    assert np.array_equal(s, ref_s), (cols, v, name)
                    assert np.array_equal(r, ref_r), (cols, v, name)
                else:
                    assert _rel(s, ref_s) < 2e-6, (cols, v, name)
                    assert _rel(r, ref_r) < 2e-6, (cols, v, name)
            c.fft_variant(W, H, cols, set=0)
    finally:
        c.close()


@pytest.mark.parametrize("W,H", [(5760, 4092), (576, 408)])
def test_radix_orders_match_numpy(W, H):
    """The second radix order of a length (rows 2880 and 288, columns 4092
    and 408) against numpy's FP64 transform, forward and round trip."""
    from motioncor3_amd.lib import Context
    c = Context(0)
    try:
        rng = np.random.default_rng(3 * W + H)
        x = rng.standard_normal((1, H, W)).astype(np.float32)
        ref = np.fft.rfft2(x[0].astype(np.float64))
        seen = 0
        for cols in (0, 1):
            n = c.fft_variant(W, H, cols)[0]
            base = _radix(c.fft_variant_name(W, H, cols, 0)) if n else None
            for v in range(n):
                name = c.fft_variant_name(W, H, cols, v)
                if _radix(name) == base:
                    continue
                seen += 1
                print("radix order", W, H, name, flush=True)
                c.fft_variant(W, H, cols, set=v)
                spec = c.rfft2(c.upload(x))
                assert _rel(spec.numpy()[0].T, ref) < 2e-6, name
                back = c.irfft2(c.rfft2(c.upload(x), scale=1.0 / (W * H)), W).numpy()
                assert _rel(back, x) < 2e-6, name
            if n:
                c.fft_variant(W, H, cols, set=0)
        assert seen >= 2, (W, H, seen)
    finally:
        c.close()


def test_autotune_keeps_alignment_results():
    """mc_autotune picks tile variants by timing; the alignment results must
    not change.  With radix orders too (mc_autotune_schedules), the shifts
    stay within rounding and the sums within 1e-4 of the FP64 oracle's on
    the same shifts."""
    from motioncor3_amd import lib as L
    from motioncor3_amd.synth import MovieRecipe
    W, H, N = 5760, 4092, 6
    c = L.Context(0)
    try:
        r = MovieRecipe(W, H, N, 20000, seed=606, dose=1.0, local_max=3.0)
        fr = c.empty((N, H, W), np.float32)
        c.synth_render(fr, r)
        pg = L.default_params(patch=(5, 5), pix_size=0.83, kv=300, fm_dose=1.0)
        a = c.align_movie(fr, pg)
        rep = c.autotune(W, H, N, patch=(5, 5))
        assert "5760x4092 columns ->" in rep and "5760x4092 rows ->" in rep, rep
        b = c.align_movie(fr, pg)
        assert np.array_equal(a["global_shift"], b["global_shift"])
        assert np.array_equal(a["patch_shift"], b["patch_shift"])
        for i in range(a["num_sums"]):
            assert np.array_equal(a["sums"][i].numpy(), b["sums"][i].numpy())
        # radix orders as well: the same transforms up to FP32 rounding, so
        # the shifts move by rounding only.  The sums are compared with the
        # FP64 oracle on each run's own shifts: the local warp truncates its
        # source coordinate, so where a weighted local shift is near zero (or
        # any integer) a 1e-5 px shift difference moves whole regions of a
        # frame by one pixel, and two runs' sums differ by percents.
        rep = c.autotune(W, H, N, patch=(5, 5), radix=True)
        assert "radix" in rep, rep
        d = c.align_movie(fr, pg)
        assert np.max(np.abs(a["global_shift"] - d["global_shift"])) < 1e-3
        assert np.max(np.abs(a["patch_shift"] - d["patch_shift"])) < 1e-3
        assert np.array_equal(a["patch_bad"], d["patch_bad"])
        import oracle as O
        po = O.default_params(patch=(5, 5), pix_size=0.83, kv=300, fm_dose=1.0)
        host = fr.numpy()
        for res in (a, d):
            rep_o = O.align_movie(host, po, replay_global=res["global_shift"],
                                  replay_patch=res["patch_shift"], replay_centers=res["patch_center"],
                                  replay_bad=res["patch_bad"])
            for i in range(res["num_sums"]):
                x = res["sums"][i].numpy()
                assert np.linalg.norm(x - rep_o["sums"][i]) / np.linalg.norm(rep_o["sums"][i]) < 1e-4, i
    finally:
        c.close()
