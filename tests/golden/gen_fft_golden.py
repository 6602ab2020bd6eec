"""Writes tests/golden/fft_golden.npz: seeded inputs and numpy.fft (pocketfft,
FP64) outputs at lengths with the prime factors the MotionCor3 sizes produce
(5760x4092 frames: 4092 = 2^2*3*11*31; xcf 2882 = 2*11*131; patches
576x408, 408 = 2^3*3*17; 410 = 2*5*41; 292 = 2^2*73).

Run:  python tests/golden/gen_fft_golden.py
"""
import os

import numpy as np

C2C_SIZES = [8, 12, 51, 62, 124, 146, 205, 292, 408, 410, 1441, 2046, 2882]
R2_SIZES = [(64, 62), (146, 34), (292, 24), (410, 34), (576, 17), (2882, 4)]


def main():
    rng = np.random.default_rng(20241202)
    out = {}
    for n in C2C_SIZES:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out["c2c_in_%d" % n] = x
        out["c2c_fwd_%d" % n] = np.fft.fft(x)
        out["c2c_inv_%d" % n] = np.fft.ifft(x) * n
    for (w, h) in R2_SIZES:
        x = rng.standard_normal((h, w))
        out["r2_in_%dx%d" % (w, h)] = x
        out["r2_fwd_%dx%d" % (w, h)] = np.fft.rfft2(x)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
