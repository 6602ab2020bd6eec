"""CPU restatement of the ingest step before the alignment boundary (test
infrastructure: only tests, the smoke check and the benchmark's CPU leg use
it).

apply_refs follows MrcUtil/GApplyRefsToFrame.cu:15-77 as driven by
MrcUtil/CApplyRefs.cpp:101-212 (centre crop of the stack frame, v = raw,
v -= dark, v *= gain, in FP32); unpack4 follows mGUnpack / mGDoPkd
(:15-52, even x in the low nibble); post_process follows
CLoadRefs::PostProcess (CLoadRefs.cpp:204-228) with G90Rotate2D
(MrcUtil/G90Rotate2D.cu:17-33,80-91), GFlip2D and GInverse2D.

Pinned in tests/test_oracle_ingest.py: rotations against numpy.rot90,
flips against flipud / fliplr, the nibble order against a bit-by-bit loop,
the crop and arithmetic against hand-computed small cases.
"""
import math

import numpy as np


def unpack4(packed, sx):
    """[n, sy, (sx + 1) // 2] bytes -> [n, sy, sx] values 0..15."""
    p = np.asarray(packed, np.uint8)
    lo = p & 0xF
    hi = p >> 4
    out = np.empty(p.shape[:2] + (2 * p.shape[2],), np.uint8)
    out[:, :, 0::2] = lo
    out[:, :, 1::2] = hi
    return out[:, :, :sx]


def apply_refs(raw, gain, dark, W, H):
    """raw [n, sy, sx] (numeric) -> float32 [n, H, W]."""
    raw = np.asarray(raw)
    n, sy, sx = raw.shape
    x0, y0 = (sx - W) // 2, (sy - H) // 2
    v = raw[:, y0:y0 + H, x0:x0 + W].astype(np.float32)
    if dark is not None:
        v = v - np.asarray(dark, np.float32)[y0:y0 + H, x0:x0 + W]
    if gain is not None:
        v = v * np.asarray(gain, np.float32)[y0:y0 + H, x0:x0 + W]
    return v.astype(np.float32)


def rotate(img, k):
    """k x 90 degrees by the reference's nearest-pixel formula."""
    if k == 0:
        return np.array(img, np.float32)
    h, w = img.shape
    ang = math.atan(1.0) * 4 / 180.0 * 90.0 * k
    fc, fs = np.float32(math.cos(ang)), np.float32(math.sin(ang))
    ic = int(fc + np.float32(0.1)) if fc > 0 else int(fc - np.float32(0.1))
    is_ = int(fs + np.float32(0.1)) if fs > 0 else int(fs - np.float32(0.1))
    rw, rh = (w, h) if k % 2 == 0 else (h, w)
    out = np.empty((rh, rw), np.float32)
    f32 = np.float32
    for y in range(rh):
        for x in range(rw):
            nx = f32(x) - f32(rw - 1) * f32(0.5)
            ny = f32(y) - f32(rh - 1) * f32(0.5)
            fx = nx * f32(ic) + ny * f32(is_)
            fy = -nx * f32(is_) + ny * f32(ic)
            sx = int(fx + f32(w - 1) * f32(0.5) + f32(0.5))
            sy = int(fy + f32(h - 1) * f32(0.5) + f32(0.5))
            out[y, x] = img[sy, sx]
    return out


def flip(img, f):
    if f == 1:
        return img[::-1, :].copy()
    if f == 2:
        return img[:, ::-1].copy()
    return img.copy()


def post_process(gain, dark, rot, flp, inv, quirks=True):
    """CLoadRefs::PostProcess: returns (gain, dark)."""
    g = None if gain is None else np.array(gain, np.float32)
    d = None if dark is None else np.array(dark, np.float32)
    if g is None or (rot == 0 and flp == 0 and inv == 0):
        return g, d
    g = flip(rotate(g, rot), flp)
    if inv:
        with np.errstate(divide="ignore"):
            g = np.where(g == 0, np.float32(0), np.float32(1) / g).astype(np.float32)
    if d is not None and ((rot != 0 or flp == 0) if quirks else (rot != 0 or flp != 0)):
        d = flip(rotate(d, rot), flp)
    return g, d
