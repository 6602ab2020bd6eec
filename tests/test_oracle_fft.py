"""The oracle's FFT against numpy.fft golden vectors (tests/golden)."""
import os

import numpy as np
import pytest

import oracle as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("n", [8, 12, 51, 62, 124, 146, 205, 292, 408, 410, 1441, 2046, 2882])
def test_c2c_golden(n):
    x = GOLD["c2c_in_%d" % n]
    assert _rel(O.fft_c2c(x, -1), GOLD["c2c_fwd_%d" % n]) < 1e-12
    assert _rel(O.fft_c2c(x, +1), GOLD["c2c_inv_%d" % n]) < 1e-12


@pytest.mark.parametrize("wh", [(64, 62), (146, 34), (292, 24), (410, 34), (576, 17), (2882, 4)])
def test_rfft2_golden_and_roundtrip(wh):
    w, h = wh
    x = GOLD["r2_in_%dx%d" % (w, h)]
    f = O.rfft2(x)
    assert _rel(f, GOLD["r2_fwd_%dx%d" % (w, h)]) < 1e-12
    back = O.irfft2(f, w) / (w * h)
    assert np.max(np.abs(back - x)) < 1e-12


def test_parseval_and_delta():
    rng = np.random.default_rng(1)
    x = rng.standard_normal(4092)
    f = O.fft_c2c(x, -1)
    assert abs(np.sum(np.abs(f) ** 2) / 4092 - np.sum(x ** 2)) < 1e-8 * np.sum(x ** 2)
    d = np.zeros(131, complex)
    d[0] = 1
    assert np.allclose(O.fft_c2c(d, -1), 1.0)
