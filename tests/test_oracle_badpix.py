"""Pins of the bad-pixel oracle (oracle/badpix_oracle.py)."""
import numpy as np

import badpix_oracle as BP


def test_template_closed_form():
    # 6 x 6: 20 border pixels at -1, 16 inside at +1; mean -1/9, and
    # E[t^2] = 1 gives std = sqrt(1 - 1/81)
    t = BP.template(6, 6)
    m, s = -1.0 / 9.0, np.sqrt(1.0 - 1.0 / 81.0)
    assert np.allclose(t[0, 0], (-1 - m) / s, rtol=1e-6)
    assert np.allclose(t[2, 2], (1 - m) / s, rtol=1e-6)
    assert abs(t.astype(np.float64).mean()) < 1e-6 and abs(t.astype(np.float64).std() - 1) < 1e-6


def _frames(seed=0, n=4, H=64, W=80):
    rng = np.random.default_rng(seed)
    return rng.poisson(5.0, (n, H, W)).astype(np.float32)


def test_hot_pixels_found():
    f = _frames()
    f[:, 10, 20] += 400.0
    f[:, 40, 7] += 300.0
    bad, (hot, _) = BP.detect(f)
    assert bad[10, 20] == 1 and bad[40, 7] == 1
    s = BP.frame_sum(f)
    assert hot > s.mean() and np.array_equal(bad[s > hot], np.ones(int((s > hot).sum()), np.uint8))


def test_defect_block_found_by_correlation():
    f = _frames(1)
    # a dark 4 x 4 defect inside a 6 x 6 window: strongest template response
    f[:, 21:25, 31:35] = 0.0
    f[:, 20:26, 30:36] += 0.0
    bad, _ = BP.detect(f)
    assert bad[21:25, 31:35].all()
    assert bad.sum() < 200


def test_rectangles_and_map():
    f = _frames(2)
    dm = np.zeros((64, 80), np.uint8)
    dm[5, 5] = 3
    bad, _ = BP.detect(f, rects=[(-2, 60, 5, 10)], defect_map=dm)
    assert bad[60:64, 0:3].all() and bad[5, 5] == 1


def test_correction_search_hand_traced():
    f = _frames(3, n=2)
    bad = np.zeros((64, 80), np.uint8)
    bad[30, 40] = 1
    out = BP.correct(f, bad)
    # first step: next = (30 * 82 + 40) * 109 + 619, times 997, mod 31^2
    nxt = ((30 * 82 + 40) * 109 + 619) * 997 % 2 ** 32
    ix = nxt % 961
    iy, ix = ix // 31 - 15 + 30, ix % 31 - 15 + 40
    assert np.array_equal(out[:, 30, 40], f[:, iy, ix])
    mask = np.ones((64, 80), bool)
    mask[30, 40] = False
    assert np.array_equal(out[:, mask], f[:, mask])


def test_correction_uses_good_pixels_in_the_window():
    f = _frames(4, n=1)
    bad = np.zeros((64, 80), np.uint8)
    bad[10:14, 10:14] = 1
    out = BP.correct(f, bad)
    good = f[0][bad == 0]
    for y in range(10, 14):
        for x in range(10, 14):
            assert out[0, y, x] in good


def test_local_cc_against_brute_force():
    rng = np.random.default_rng(9)
    img = rng.normal(10, 3, (13, 17)).astype(np.float32)
    cc = BP.local_cc(img, 6, 6)
    t = BP.template(6, 6).astype(np.float64)
    for y in range(13):
        for x in range(17):
            if x + 6 >= 17 or y + 6 >= 13:
                assert cc[y, x] == 0
                continue
            w = img[y:y + 6, x:x + 6].astype(np.float64)
            want = abs((t * w).mean()) / w.std()
            assert abs(cc[y, x] - want) <= 1e-4 * want
