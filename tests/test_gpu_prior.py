"""Learned shift prior on the GPU: the BF16 MFMA product, the network's
forward pass and gradient against the FP64 CPU restatement
(oracle/prior_oracle.py), training, and the guided patch alignment.

Tolerances: the GPU rounds layer operands to BF16 exactly as the oracle's
bf16 mode does and accumulates in FP32, so the products agree to FP32
accumulation error (relative 1e-5 of sum |a||b|); forward outputs and
gradients, where a BF16 rounding of an activation can flip when FP32 and
FP64 sums straddle a rounding midpoint, agree within relative L2 2e-3.
"""
import numpy as np
import pytest

import prior_oracle as PO
from motioncor3_amd import lib as L
from motioncor3_amd.synth import MovieRecipe

pytestmark = pytest.mark.gpu


def _bf16_bits(x):
    r = PO.bf16_round(np.asarray(x, np.float32)).astype(np.float32)
    return (r.view(np.uint32) >> 16).astype(np.uint16), r.astype(np.float64)


@pytest.mark.parametrize("M,N,K,ta,tb", [
    (37, 45, 19, False, False), (37, 45, 19, True, False), (37, 45, 19, False, True),
    (64, 128, 1090, False, False), (1090, 128, 64, True, False), (64, 1090, 128, False, True),
    (1, 1, 1, False, False), (100, 33, 8, True, True),
])
def test_gemm_bf16_matches_fp64(ctx, M, N, K, ta, tb):
    """Pins the MFMA operand and accumulator lane layout: any swap of rows,
    columns or k-slices shows as a large error."""
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    a16, a = _bf16_bits(rng.normal(0, 1, (M, K)))
    b16, b = _bf16_bits(rng.normal(0, 1, (K, N)))
    # storage: A as [M][K] (or transposed [K][M]), B as [K][N] (or [N][K])
    da = ctx.upload(np.ascontiguousarray(a16.T if ta else a16))
    db = ctx.upload(np.ascontiguousarray(b16.T if tb else b16))
    sa = (1, M) if ta else (K, 1)
    sb = (1, K) if tb else (N, 1)
    dc = ctx.empty((M, N), np.float32)
    L.gemm_bf16(ctx, da, sa, db, sb, dc, M, N, K)
    c = dc.numpy().astype(np.float64)
    ref = a @ b
    bound = 1e-5 * (np.abs(a) @ np.abs(b)) + 1e-30
    assert np.all(np.abs(c - ref) <= bound), np.max(np.abs(c - ref) / bound)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_and_gradient_match_oracle(ctx):
    frames, hid, n = 4, 32, 70
    pr = L.Prior(ctx, frames, hid, seed=3)
    assert (pr.in_dim, pr.out_dim) == (68 * frames + 2, 2 * frames)
    dims = [pr.in_dim, hid, hid, hid, pr.out_dim]
    assert pr.n_params == sum(dims[i] * dims[i + 1] + dims[i + 1] for i in range(4))
    rng = np.random.default_rng(8)
    # non-zero biases so that their gradients and the ReLU masks are exercised
    w = pr.params().astype(np.float64)
    for (W, b) in PO.split(w, dims):
        b[:] = rng.normal(0, 0.2, b.size)
    pr.set_params(w.astype(np.float32))
    w = pr.params().astype(np.float64)
    X = rng.normal(0, 1, (n, pr.in_dim)).astype(np.float32)
    T = rng.normal(0, 0.3, (n, pr.out_dim)).astype(np.float32)
    xd, td = ctx.upload(X), ctx.upload(T)
    layers = PO.split(w, dims)
    y = pr.forward(xd)
    yo = PO.forward(layers, X, bf16=True)
    assert _rel(y, yo) < 2e-3
    loss, g = pr.loss_grad(xd, td)
    lo, go = PO.loss_grad(layers, X, T, bf16=True)
    assert abs(loss - lo) / lo < 1e-4
    off = 0
    for l in range(4):
        nw = dims[l] * dims[l + 1]
        for part, sl in (("W", slice(off, off + nw)), ("b", slice(off + nw, off + nw + dims[l + 1]))):
            assert _rel(g[sl], go[sl]) < 2e-3, (l, part, _rel(g[sl], go[sl]))
        off += nw + dims[l + 1]
    # and the BF16 network's gradient stays close to the exact FP64 one
    _, g64 = PO.loss_grad(layers, X, T, bf16=False)
    assert _rel(g, g64) < 5e-2
    pr.close()


def test_training_is_deterministic_and_learns(ctx):
    frames, hid, n = 2, 64, 512
    rng = np.random.default_rng(21)
    pr = L.Prior(ctx, frames, hid, seed=5)
    X = rng.normal(0, 1, (n, pr.in_dim)).astype(np.float32)
    Wt = rng.normal(0, 0.3, (8, pr.out_dim))
    T = (np.tanh(X[:, :8] @ Wt)).astype(np.float32)
    xd, td = ctx.upload(X), ctx.upload(T)
    w0 = pr.params()
    l0, _ = pr.loss_grad(xd, td)
    hist = pr.train(xd, td, epochs=40, batch=64, lr=2e-3, seed=9)
    l1, _ = pr.loss_grad(xd, td)
    assert np.all(np.isfinite(hist))
    assert l1 < 0.2 * l0, (l0, l1, hist[:3], hist[-3:])
    w1 = pr.params()
    # same start, same shuffles: bit-identical parameters
    pr.set_params(w0)
    pr.train(xd, td, epochs=40, batch=64, lr=2e-3, seed=9)
    assert np.array_equal(pr.params(), w1)
    pr.close()


def _movie(ctx, W, H, N, seed, local_max):
    r = MovieRecipe(W, H, N, 500, seed=seed, dose=4.0, local_max=local_max)
    fr = ctx.empty((N, H, W), np.float32)
    ctx.synth_render(fr, r)
    return fr


def test_prior_features_layout(ctx):
    W = H = 1024
    N = 8
    fr = _movie(ctx, W, H, N, 41, 2.0)
    pg = L.default_params(patch=(3, 3), pix_size=1.0, kv=300, fm_dose=1.0)
    res = ctx.align_movie(fr, pg, prior_data=True)
    X, T = res["prior_features"], res["prior_targets"]
    assert X.shape == (9, 68 * N + 2) and T.shape == (9, N, 2)
    assert np.isfinite(X).all() and np.isfinite(T).all()
    # global trajectory: the same for every patch of the movie
    assert np.all(X[:, : 2 * N] == X[0, : 2 * N])
    # patch centre in [-1, 1] from the starts and the patch size
    geo = L.geometry(W, H, 3, 3)
    pw, ph = geo["patch"]
    xw, xh = geo["xcf"]
    st = res["patch_starts"]
    cx = (st[:, 0] + pw // 2) / xw * 2 - 1
    cy = (st[:, 1] + ph // 2) / xh * 2 - 1
    assert np.allclose(X[:, 2 * N], cx, atol=1e-6) and np.allclose(X[:, 2 * N + 1], cy, atol=1e-6)
    # pooled windows: each frame's block is scaled to a largest |value| of 1
    win = X[:, 4 * N + 2:].reshape(9, N, 64)
    assert np.allclose(np.abs(win).max(-1), 1.0, atol=1e-6)
    # the plain run is unchanged by asking for the features
    res0 = ctx.align_movie(fr, pg)
    assert np.array_equal(res0["patch_shift"], res["patch_shift"])


def _zero_output_prior(ctx, N, hid=32):
    pr = L.Prior(ctx, N, hid, seed=1)
    w = pr.params()
    dims = [pr.in_dim, hid, hid, hid, pr.out_dim]
    W3, b3 = PO.split(w.astype(np.float64), dims)[3]
    w[w.size - W3.size - b3.size:] = 0
    pr.set_params(w)
    return pr


def _same(a, b):
    assert np.array_equal(a["patch_shift"], b["patch_shift"])
    assert np.array_equal(a["global_shift"], b["global_shift"])
    for i in range(b["num_sums"]):
        assert np.array_equal(a["sums"][i].numpy(), b["sums"][i].numpy())


def test_guided_alignment_with_zero_correction_is_plain(ctx):
    """A prior whose output layer is zero adds nothing after the first
    iteration.  Patches that converge in two iterations then end exactly as
    in the plain path, and with prior_fallback 0 every other patch restarts
    from zero shifts, i.e. is solved by the plain path: bit-identical
    results, with the guided pass's two iterations added to the count."""
    N = 16
    pr = _zero_output_prior(ctx, N)
    pr.attach()
    try:
        for seed, lm, fb in ((43, 0.0, 0.2), (45, 4.0, 0.0)):
            fr = _movie(ctx, 1024, 1024, N, seed, lm)
            kw = dict(patch=(3, 3), pix_size=1.0, kv=300, fm_dose=1.0)
            plain = ctx.align_movie(fr, L.default_params(**kw))
            g = ctx.align_movie(fr, L.default_params(prior=1, prior_fallback=fb, **kw))
            slow = plain["patch_iters"] > 2
            if lm == 0.0:
                assert not slow.any()
            else:
                assert slow.any()  # the restart path is exercised
            assert g["prior_fallbacks"] == int(slow.sum())
            assert np.array_equal(g["patch_iters"], plain["patch_iters"] + 2 * slow)
            _same(g, plain)
    finally:
        pr.attach(False)
        pr.close()


def test_guided_alignment_needs_a_matching_prior(ctx):
    fr = _movie(ctx, 1024, 1024, 8, 44, 0.0)
    pgg = L.default_params(patch=(3, 3), pix_size=1.0, kv=300, fm_dose=1.0, prior=1)
    with pytest.raises(RuntimeError, match="learned prior"):
        ctx.align_movie(fr, pgg)
    pr = L.Prior(ctx, 6, 16, seed=1)
    pr.attach()
    try:
        with pytest.raises(RuntimeError, match="learned prior"):
            ctx.align_movie(fr, pgg)
    finally:
        pr.attach(False)
        pr.close()


def test_prior_ablation_cuts_iterations(ctx):
    """What the trained network gains: the prior is trained on the GPU from
    patch shifts the plain method produces on 160 seeded movies (1024 x 1024,
    16 frames, 3 x 3 patches, local motion up to 1-5 px), then 24 held-out
    movies are aligned with and without it.  Plain patches need 2-4
    iterations; guided ones stop after 2 with no fallback, and the guided
    shifts stay within 0.05 px RMS of the plain ones and within 0.1 px RMS of
    the FP64 oracle."""
    import oracle as O
    W = H = 1024
    N = 16
    fr = ctx.empty((N, H, W), np.float32)
    kw = dict(patch=(3, 3), pix_size=1.0, kv=300, fm_dose=1.0)
    pg = L.default_params(**kw)

    def render(seed):
        rng = np.random.default_rng(seed)
        r = MovieRecipe(W, H, N, 500, seed=seed, dose=4.0, local_max=float(rng.uniform(1.0, 5.0)))
        ctx.synth_render(fr, r)

    X, T = [], []
    for k in range(160):
        render(1000 + k)
        res = ctx.align_movie(fr, pg, prior_data=True)
        X.append(res["prior_features"])
        T.append(res["prior_targets"].reshape(9, -1))
    X, T = np.concatenate(X), np.concatenate(T)
    pr = L.Prior(ctx, N, 128, seed=1)
    hist = pr.train(ctx.upload(X), ctx.upload(T), epochs=150, batch=64, lr=5e-4, seed=0)
    assert hist[-1] < 0.2 * float((T ** 2).mean()), (hist[-1], float((T ** 2).mean()))
    pr.attach()
    try:
        pgg = L.default_params(prior=1, **kw)
        pit, git, fell, d = [], [], 0, []
        for k in range(24):
            render(50000 + k)
            a = ctx.align_movie(fr, pg)
            b = ctx.align_movie(fr, pgg)
            pit.append(a["patch_iters"])
            git.append(b["patch_iters"])
            fell += b["prior_fallbacks"]
            d.append(b["patch_shift"] - a["patch_shift"])
            if k == 0:
                mv = fr.numpy()
                po = O.default_params(**kw)
                own = O.align_movie(mv, po)
                assert np.array_equal(own["patch_starts"], b["patch_starts"])
                e = b["patch_shift"] - own["patch_shift"]
                assert np.sqrt((e ** 2).sum(-1).mean()) <= 0.1
        pit, git, d = np.concatenate(pit), np.concatenate(git), np.concatenate(d)
        print("patch iterations: plain mean %.2f max %d, guided mean %.2f max %d, fallbacks %d/%d, "
              "guided-plain RMS %.4f px" % (pit.mean(), pit.max(), git.mean(), git.max(), fell,
                                            git.size, np.sqrt((d ** 2).sum(-1).mean())))
        assert pit.max() > 2 and pit.mean() > 2.3
        assert fell <= git.size // 10
        assert git.mean() <= pit.mean() - 0.25
        assert np.sqrt((d ** 2).sum(-1).mean()) <= 0.05
    finally:
        pr.attach(False)
        pr.close()
