"""The oracle's -OutStack aligned stack (Correct/CCorrectFullShift.cpp:336-356
and Align/CZbinStack.cpp:27-70), pinned by a closed form (frames that are
integer circular shifts of one image align back onto it exactly), by
linearity against the unweighted sum, and by z-bin groups derived by hand."""
import numpy as np
import pytest

import oracle as O

N, H, W = 6, 64, 96
D = np.array([[0, 0], [1, 2], [3, -1], [-2, 4], [5, 5], [-3, -2]])  # (dx, dy) per frame


def _movie(seed=0):
    rng = np.random.default_rng(seed)
    base = rng.normal(5, 1, (H, W))
    fr = np.stack([np.roll(base, (D[i, 1], D[i, 0]), axis=(0, 1)) for i in range(N)])
    return base.astype(np.float32), fr.astype(np.float32)


@pytest.mark.parametrize("patch", [False, True])
def test_integer_shifts_align_onto_the_base_image(patch):
    base, fr = _movie()
    kw = dict(replay_global=-D.astype(float))  # reported shift = -displacement
    if patch:
        p = O.default_params(patch=(2, 2), out_stack=(1, 1))
        kw.update(replay_patch=np.zeros((N, 4, 2)),
                  replay_centers=np.array([[24, 16], [72, 16], [24, 48], [72, 48]], float))
    else:
        p = O.default_params(out_stack=(1, 1))
    r = O.align_movie(fr, p, **kw)
    assert r["stack"].shape == (N, H, W)
    # patch mode corrects relative to the reference frame N/2
    # (Align/CPatchAlign.cpp:47-65): every frame lands on that frame
    want = fr[N // 2] if patch else base
    assert np.abs(r["stack"] - want[None]).max() < 1e-9


@pytest.mark.parametrize("ft_bin", [1.0, 2.0])
def test_stack_sums_to_the_unweighted_sum(ft_bin):
    _, fr = _movie(1)
    rng = np.random.default_rng(2)
    sh = rng.uniform(-3, 3, (N, 2))
    p = O.default_params(out_stack=(1, 1), ft_bin=ft_bin, pix_size=1.0, kv=300, fm_dose=1.0)
    r = O.align_movie(fr, p, replay_global=sh)
    st = r["stack"]
    assert st.shape == (N, r["sums"].shape[1], r["sums"].shape[2])
    # the stack is taken before dose weighting and without the clamp at zero
    assert np.abs(np.maximum(st.sum(0), 0) - r["sums"][0]).max() < 1e-9 * np.abs(r["sums"][0]).max()
    assert not np.allclose(st.sum(0), r["sums"][1])


def test_zbin_groups():
    _, fr = _movie(3)
    sh = np.random.default_rng(4).uniform(-2, 2, (N, 2))
    one = O.align_movie(fr, O.default_params(out_stack=(1, 1)), replay_global=sh)["stack"]
    four = O.align_movie(fr, O.default_params(out_stack=(1, 4)), replay_global=sh)["stack"]
    # 6 frames in groups of 4 raw frames: [0..3], [4, 5]
    assert four.shape[0] == 2
    assert np.allclose(four[0], one[:4].sum(0), atol=1e-9)
    assert np.allclose(four[1], one[4:].sum(0), atol=1e-9)
    # integrated frames of 2, 2, 1, 1, 3, 1 raw frames: groups close once they
    # hold >= 4 x (smallest frame = 1) raw frames: [0, 1], [2, 3, 4], [5]
    sz = [2, 2, 1, 1, 3, 1]
    st = np.cumsum([0] + sz[:-1])
    pi = O.default_params(out_stack=(1, 4), fm_int_start=st, fm_int_size=sz)
    grp = O.align_movie(fr, pi, replay_global=sh)["stack"]
    p1 = O.default_params(out_stack=(1, 1), fm_int_start=st, fm_int_size=sz)
    one_i = O.align_movie(fr, p1, replay_global=sh)["stack"]
    assert grp.shape[0] == 3
    for k, idx in enumerate(([0, 1], [2, 3, 4], [5])):
        assert np.allclose(grp[k], one_i[idx].sum(0), atol=1e-9), k
