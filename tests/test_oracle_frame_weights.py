"""-InFmMotion and -CorrInterp frequency weights in the oracle
(MotionDecon/GMotionWeight.cu, Util/GCorrLinearInterp.cu), pinned by closed
forms: a motion along x is a sinc in kx alone; a general motion is a sinc of
the distance along the motion direction (a projection, derived independently
of the reference's slope form); the interpolation weight at radius 1/2; and
on the whole path, a constant motion scales the global-mode sum spectrum by
the weight once and the patch-mode one twice."""
import numpy as np
import pytest

import oracle as O

CX, CY = 65, 48


def test_motion_along_x_is_a_sinc_in_kx():
    w = O.frame_weight(CX, CY, 6.0, 0.0)
    t = 6.0 * np.arange(CX) / (2 * (CX - 1))
    want = np.where(t > 1, 0.0, np.sinc(t))
    want[0] = 1.0
    assert np.allclose(w, want[None, :], atol=1e-12)


@pytest.mark.parametrize("mx,my", [(3.0, -4.0), (-1.5, 7.0), (0.4, 0.3)])
def test_general_motion_is_a_sinc_of_the_distance_along_the_motion(mx, my):
    w = O.frame_weight(CX, CY, mx, my)
    m = np.hypot(mx, my)
    k = np.arange(CX)[None, :]
    yy = np.arange(CY)[:, None]
    yy = np.where(yy > CY // 2, yy - CY, yy)
    d = np.abs(k * mx + yy * my) / m
    t = m * d / (2 * (CX - 1))
    want = np.where(d < 0.1, 1.0, np.where(t > 1, 0.0, np.sinc(t)))
    want[0, 0] = 1.0
    assert np.allclose(w, want, atol=1e-9)


def test_small_motion_leaves_the_spectrum():
    assert np.array_equal(O.frame_weight(CX, CY, 0.06, -0.07), np.ones((CY, CX)))


def test_corr_interp_values():
    w = O.frame_weight(CX, CY, 0.0, 0.0, motion=False, corr_interp=True)
    q = (np.pi / 2) ** 0.125  # radius 1/2: sinc = 2 / pi
    assert w[0, 0] == 1.0 and (w >= 1.0).all()
    assert np.isclose(w[0, CX - 1], q) and np.isclose(w[CY // 2, 0], q)
    # radial: (k / (2 (CX - 1)), y / CY) = (0.25, 0) and (0, 0.25)
    assert np.isclose(w[0, (CX - 1) // 2], w[CY // 4, 0])


@pytest.mark.parametrize("patch", [False, True])
def test_constant_motion_scales_the_sum_spectrum(patch):
    N, H, W = 6, 64, 96
    rng = np.random.default_rng(5)
    base = rng.normal(5, 1, (H, W))
    d = np.array([[2 * i, 0] for i in range(N)])
    fr = np.stack([np.roll(base, (d[i, 1], d[i, 0]), axis=(0, 1)) for i in range(N)]).astype(np.float32)
    kw = dict(replay_global=-d.astype(float))
    pk = {}
    if patch:
        pk = dict(patch=(2, 2))
        kw.update(replay_patch=np.zeros((N, 4, 2)),
                  replay_centers=np.array([[24, 16], [72, 16], [24, 48], [72, 48]], float))
    off = O.align_movie(fr, O.default_params(**pk), **kw)["sums"][0]
    on = O.align_movie(fr, O.default_params(in_fm_motion=1, **pk), **kw)["sums"][0]
    assert off.min() > 0 and on.min() > 0  # the clamp at zero stays inactive
    ratio = np.fft.rfft2(on) / np.fft.rfft2(off)
    w = O.frame_weight(W // 2 + 1, H, -2.0, 0.0)
    want = w * w if patch else w
    keep = np.abs(np.fft.rfft2(off)) > 1e-6 * np.abs(np.fft.rfft2(off)).max()
    assert np.allclose(ratio[keep], want[keep], atol=1e-9)
