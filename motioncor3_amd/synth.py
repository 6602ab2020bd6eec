"""Seeded synthetic cryo-EM movies with known drift (tests and benchmark input).

Recipe:
  * phantom: `nblobs` Gaussian blobs, sigma ~ U(3, 8) px, amplitude ~
    U(0.5, 1.5), uniform positions, on a unit background; the frame is scaled
    so its noiseless mean equals `dose` e-/px/frame;
  * global drift: per-frame steps ~ N(0, step_sigma) px per axis, cumulative
    sum, 5-frame box smoothing, frame 0 at zero displacement;
  * local deformation (optional): sum over a+b <= 2, c = 1..3 of
    k_abc * x^a * y^b * t^c with x, y in [-0.5, 0.5] and t in [0, 1], seeded
    coefficients scaled so the largest displacement over the frame is
    `local_max` px;
  * each blob is rendered analytically at its displaced centre (periodic
    wrap), which equals an exact Fourier shift for these band-limited blobs;
  * counts ~ Poisson(noiseless frame) as float32.

`displacement[t]` is the content displacement of frame t (x, y); the
alignment reports shifts of the opposite sign up to a constant.
"""
import numpy as np


class MovieRecipe:
    def __init__(self, W, H, N, nblobs, seed, dose=1.0, step_sigma=0.5,
                 local_max=0.0):
        self.W, self.H, self.N = int(W), int(H), int(N)
        self.dose = float(dose)
        self.seed = int(seed)
        rng = np.random.default_rng(seed)
        self.bx = rng.uniform(0, W, nblobs)
        self.by = rng.uniform(0, H, nblobs)
        self.bs = rng.uniform(3.0, 8.0, nblobs)
        self.ba = rng.uniform(0.5, 1.5, nblobs)
        steps = rng.normal(0.0, step_sigma, (N, 2))
        steps[0] = 0.0
        cum = np.cumsum(steps, axis=0)
        pad = np.concatenate([np.repeat(cum[:1], 2, 0), cum, np.repeat(cum[-1:], 2, 0)])
        sm = np.stack([np.convolve(pad[:, k], np.ones(5) / 5.0, mode="valid")
                       for k in range(2)], axis=1)
        self.displacement = sm - sm[0]
        self.local_coef = None
        self.local_max = float(local_max)
        if local_max > 0:
            self.local_coef = rng.normal(0.0, 1.0, (2, 6, 3))
            xs = np.linspace(-0.5, 0.5, 33)
            X, Y = np.meshgrid(xs, xs)
            peak = 0.0
            for t in np.linspace(0.0, 1.0, 11):
                dx, dy = self._local_raw(X, Y, t)
                peak = max(peak, float(np.max(np.hypot(dx, dy))))
            self.local_coef *= local_max / max(peak, 1e-12)
        # noiseless mean of the unscaled phantom: 1 + sum of blob integrals / area
        integ = np.sum(self.ba * 2.0 * np.pi * self.bs ** 2)
        self.scale = dose / (1.0 + integ / (W * H))

    _MONO = [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2)]

    def _local_raw(self, x, y, t):
        dx = np.zeros_like(np.asarray(x, dtype=np.float64))
        dy = np.zeros_like(dx)
        for m, (a, b) in enumerate(self._MONO):
            base = (x ** a) * (y ** b)
            for c in range(3):
                tc = t ** (c + 1)
                dx = dx + self.local_coef[0, m, c] * base * tc
                dy = dy + self.local_coef[1, m, c] * base * tc
        return dx, dy

    def blob_displacement(self, frame):
        """Per-blob (dx, dy) of frame `frame`: global drift + local field."""
        gx, gy = self.displacement[frame]
        dx = np.full(self.bx.shape, gx)
        dy = np.full(self.by.shape, gy)
        if self.local_coef is not None:
            t = frame / max(self.N - 1, 1)
            lx, ly = self._local_raw(self.bx / self.W - 0.5, self.by / self.H - 0.5, t)
            dx = dx + lx
            dy = dy + ly
        return dx, dy

    def local_field(self, x, y, frame):
        """Local displacement at full-res pixel coordinates (no global part)."""
        if self.local_coef is None:
            return np.zeros_like(x, dtype=np.float64), np.zeros_like(y, dtype=np.float64)
        t = frame / max(self.N - 1, 1)
        return self._local_raw(np.asarray(x) / self.W - 0.5, np.asarray(y) / self.H - 0.5, t)

    def mean_local(self, frame, grid=64):
        """Frame mean of the local displacement (over a grid x grid lattice of
        pixel centres): the part of the local field a whole-frame alignment
        sees as drift."""
        if self.local_coef is None:
            return 0.0, 0.0
        xs = (np.arange(grid) + 0.5) * self.W / grid
        ys = (np.arange(grid) + 0.5) * self.H / grid
        X, Y = np.meshgrid(xs, ys)
        dx, dy = self.local_field(X, Y, frame)
        return float(dx.mean()), float(dy.mean())

    def global_truth(self):
        """N x 2 content displacement of each frame as a whole-frame alignment
        sees it: the global drift plus the frame mean of the local field."""
        return self.displacement + np.array([self.mean_local(t) for t in range(self.N)])

    def render(self, frame):
        """Noiseless frame (H x W float64), periodic wrap."""
        W, H = self.W, self.H
        img = np.ones((H, W), np.float64)
        dx, dy = self.blob_displacement(frame)
        cx = self.bx + dx
        cy = self.by + dy
        for k in range(len(self.bx)):
            s = self.bs[k]
            r = int(np.ceil(4.0 * s))
            x0 = int(np.floor(cx[k])) - r
            y0 = int(np.floor(cy[k])) - r
            xs = np.arange(x0, x0 + 2 * r + 2)
            ys = np.arange(y0, y0 + 2 * r + 2)
            gx = np.exp(-((xs - cx[k]) ** 2) / (2 * s * s))
            gy = np.exp(-((ys - cy[k]) ** 2) / (2 * s * s))
            img[np.ix_(ys % H, xs % W)] += self.ba[k] * np.outer(gy, gx)
        return img * self.scale

    def movie(self, noise=True):
        """N x H x W float32 movie (Poisson counts when noise=True)."""
        rng = np.random.default_rng(self.seed + 7919)
        out = np.empty((self.N, self.H, self.W), np.float32)
        for t in range(self.N):
            img = self.render(t)
            out[t] = rng.poisson(img).astype(np.float32) if noise else img.astype(np.float32)
        return out


def eer_stream_7bit(positions, subpix, cam_pixels):
    """A 7-bit EER code stream (compression 65001, as EerUtil/CDecodeEerFrame.cpp
    reads it) of electrons at sorted, distinct pixel indices `positions` with
    sub-pixel nibbles `subpix` (benchmark input).  Each electron is its zero
    run (127 = run continues, no electron) then its nibble ^ 0x0A, least
    significant bit first; a final run reaches the end of the camera."""
    q = np.asarray(positions, np.int64)
    sub = np.asarray(subpix, np.int64) & 15
    prev = np.concatenate([[-1], q[:-1]]) if len(q) else np.zeros(0, np.int64)
    runs = q - prev - 1
    rest = int(cam_pixels - (q[-1] + 1 if len(q) else 0))
    full = np.concatenate([runs // 127, [rest // 127]])
    last = np.concatenate([runs % 127, [rest % 127]])
    # fields per electron: full x (127, 7 bits), (run, 7 bits), (nibble, 4 bits);
    # then the final run
    nf = full + 2
    nf[-1] = full[-1] + 1
    vals = np.full(int(nf.sum()), 127, np.int64)
    wid = np.full(len(vals), 7, np.int64)
    end = np.cumsum(nf)
    start = end - nf
    lastpos = start + full
    vals[lastpos] = last
    nib = lastpos[:-1] + 1
    vals[nib] = sub ^ 0x0A
    wid[nib] = 4
    bitpos = np.cumsum(wid) - wid
    total = int(wid.sum())
    bits = np.zeros(total + 40, np.uint8)
    for j in range(7):
        m = j < wid
        bits[bitpos[m] + j] = (vals[m] >> j) & 1
    return np.packbits(bits, bitorder="little")[:(total + 7) // 8 + 4]
