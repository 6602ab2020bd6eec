"""Bad-pixel detection and correction on the GPU against the oracle
(oracle/badpix_oracle.py): the map bit for bit (planted hot pixels, a
defect block, a defect file and a defect map), the thresholds, and the
corrected frames bit for bit."""
import numpy as np
import pytest

import badpix_oracle as BP

pytestmark = pytest.mark.gpu


def test_detect_and_correct_match_oracle(ctx, tmp_path):
    rng = np.random.default_rng(17)
    N, H, W = 6, 256, 320
    f = rng.poisson(3.0, (N, H, W)).astype(np.float32)
    f[:, 17, 33] += 200.0
    f[:, 200, 301] += 150.0
    f[:, 100:104, 150:154] = 0.0
    dm = np.zeros((H, W), np.uint8)
    dm[5, 9] = 1
    dm[250, 3] = 255
    path = str(tmp_path / "defects.txt")
    open(path, "w").write("10 240 6 3\n# comment\n318 0 5 2\n")
    pitch = (W // 2 + 1) * 2
    padded = np.zeros((N, H, pitch), np.float32)
    padded[:, :, :W] = f
    d = ctx.upload(padded)
    bad, nbad, th = ctx.detect_bad(d, defect_file=path, defect_map=dm, width=W)
    want, wth = BP.detect(f, rects=[(10, 240, 6, 3), (318, 0, 5, 2)], defect_map=dm)
    got = bad.numpy()
    assert th == (float(wth[0]), float(wth[1]))
    assert np.array_equal(got, want)
    assert nbad == int(want.sum())
    assert got[17, 33] == 1 and got[100:104, 150:154].all() and got[5, 9] == 1
    ctx.correct_bad(d, bad, width=W)
    out = d.numpy()[:, :, :W]
    assert np.array_equal(out, BP.correct(f, want))
