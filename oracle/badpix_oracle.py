"""CPU restatement of bad-pixel detection and correction (test
infrastructure: only tests, the smoke check and the benchmark's CPU leg use
it).

detect() follows CDetectMain::DoIt (BadPixel/CDetectMain.cpp:54-245): frame
sum (MrcUtil/CSumFFTStack.cpp:65-117, FP32, frames in order), hot pixels
(BadPixel/GDetectHot.cu:25-55), the local correlation map with the border
template (BadPixel/GLocalCC.cu:15-53, BadPixel/CTemplate.cpp:19-57) and its
threshold and block marking (BadPixel/GDetectPatch.cu:9-105), defect
rectangles and an 8-bit map.  Moments are taken in FP64 and the thresholds
then formed in FP32 as the reference does; the correlation map is FP32 with
the window sums accumulated in the reference's order.  correct() follows
GCorrectBad (BadPixel/GCorrectBad.cu:13-64), pseudo-random search included.

Pinned in tests/test_oracle_badpix.py: the template against its closed
form, hot pixels and a planted defect block found at their positions, the
search against a hand-traced sequence and brute-force properties.
"""
import numpy as np


def frame_sum(frames):
    f = np.asarray(frames, np.float32)
    s = f[0].copy()
    for n in range(1, f.shape[0]):
        s = (s + f[n]).astype(np.float32)
    return s


def template(mw, mh):
    t = np.ones((mh, mw), np.float32)
    t[0, :] = t[-1, :] = -1.0
    t[:, 0] = t[:, -1] = -1.0
    t = t.ravel()
    n = t.size
    dm = float(np.sum(t.astype(np.float64)))
    ds = float(np.sum((t * t).astype(np.float64)))
    fm = np.float32(dm / n)
    ds = ds / n - float(np.float32(fm * fm))
    fs = np.float32(0.0) if ds <= 0 else np.float32(np.sqrt(ds))
    fs = np.float32(fs + np.float32(1e-30))
    return ((t - fm) / fs).astype(np.float32).reshape(mh, mw)


def _thr(img):
    v = img.astype(np.float64)
    m = np.float32(v.mean())
    e2 = np.float32((v * v).mean())
    return m, e2


def local_cc(img, mw, mh):
    img = np.asarray(img, np.float32)
    H, W = img.shape
    tpl = template(mw, mh)
    cc = np.zeros((H, W), np.float32)
    ny, nx = H - mh, W - mw  # window must end before the last row / column
    if ny <= 0 or nx <= 0:
        return cc
    mean = np.zeros((ny, nx), np.float32)
    sq = np.zeros((ny, nx), np.float32)
    c = np.zeros((ny, nx), np.float32)
    for j in range(mh):
        for i in range(mw):
            v = img[j:j + ny, i:i + nx]
            mean = (mean + v).astype(np.float32)
            sq = (sq + (v * v).astype(np.float32)).astype(np.float32)
            c = (c + (tpl[j, i] * v).astype(np.float32)).astype(np.float32)
    n = np.float32(mw * mh)
    c = (c / n).astype(np.float32)
    mean = (mean / n).astype(np.float32)
    sd = ((sq / n).astype(np.float32) - (mean * mean).astype(np.float32)).astype(np.float32)
    sd = np.where(sd < 0, np.float32(0), np.sqrt(np.maximum(sd, 0))).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        out = np.where(sd == 0, np.float32(0), (np.abs(c) / sd)).astype(np.float32)
    cc[:ny, :nx] = out
    return cc


def detect(frames, threshold=6.0, mw=6, mh=6, rects=(), defect_map=None):
    """-> (bad map uint8 [H, W], (hot threshold, correlation threshold))."""
    s = frame_sum(frames)
    H, W = s.shape
    f32 = np.float32
    m, e2 = _thr(s)
    sd = f32(np.sqrt(max(f32(e2 - f32(m * m)), f32(0))))
    delta = f32(f32(threshold) * sd)
    if delta < 10:
        delta = f32(10)
    hot = f32(m + delta)
    bad = (s > hot).astype(np.uint8)
    cc = local_cc(s, mw, mh)
    cm, ce2 = _thr(cc)
    csd = f32(np.sqrt(max(f32(0), f32(ce2 - f32(cm * cm)))))
    cthr = f32(cm + f32(f32(threshold) * csd))
    ys, xs = np.nonzero(cc > cthr)
    for y, x in zip(ys, xs):
        if x + mw < W and y + mh < H:
            bad[y:y + mh, x:x + mw] = 1
    for (x, y, w, h) in rects:
        x0, y0 = max(x, 0), max(y, 0)
        bad[y0:min(y + h, H), x0:min(x + w, W)] = 1
    if defect_map is not None:
        bad |= (np.asarray(defect_map, np.uint8) & 1)
    return bad, (hot, cthr)


def correct(frames, bad, win=7):
    """Every bad pixel of every frame replaced as GCorrectBad does."""
    out = np.array(frames, np.float32)
    n, H, W = out.shape
    win0 = max(win, 31)
    padx = (W // 2 + 1) * 2
    M = 0xFFFFFFFF
    for y, x in zip(*np.nonzero(bad)):
        src = None
        nxt = ((y * padx + x) * 109 + 619) & M
        w = win0
        for _ in range(10):
            nk = nxt
            w2, half = w * w, w // 2
            for _k in range(500):
                nk = (nk * 997) & M
                ix = nk % w2
                iy = ix // w - half + y
                if iy < 0 or iy >= H:
                    continue
                ix = ix % w - half + x
                if ix < 0 or ix >= W:
                    continue
                if bad[iy, ix] == 1:
                    continue
                src = (iy, ix)
                break
            if src is not None:
                break
            w += 50
            nxt = (nxt * 997) & M
        if src is None:
            idx = nxt % (W * H)
            src = (idx // W, idx % W)
        out[:, y, x] = out[:, src[0], src[1]]
    return out
