"""EER decoding and rendering oracle (test infrastructure only).

Sequential restatement of the reference's EER frame decoders and renderer:
  - CDecodeEerFrame::mDo7BitsCounted / mDo7BitsSuperRes / mDo8BitsCounted /
    mDo8BitsSuperRes / mFindElectron (EerUtil/CDecodeEerFrame.cpp:72-215),
    with the sub-pixel tables of CDecodeEerFrame::Setup (:19-43);
  - CRenderMrcStack::mRender / mRenderInt / mRenderFrame
    (EerUtil/CRenderMrcStack.cpp:36-101): every rendered frame is the
    unsigned-byte event count of its fractions.
Two departures, both only for malformed input: reads stop at the end of a
fraction, and events outside the rendered frame are dropped (the reference
reads and writes past both).

`encode_7bit` / `encode_8bit` write code streams from electron lists; they
are the inverse of the format (a 7- or 8-bit zero run, then the sub-pixel
nibble) written independently of the decoders, so decode(encode(events)) ==
events pins the decoders.  The 8-bit super-resolution walk reads its second
code from byte 0 of the fraction, as the reference does; that case is pinned
by a hand-traced stream in the tests.
"""
import numpy as np

SUPERRES = {2: (2, 8, 1, 1, 3), 3: (3, 12, 2, 0, 2)}  # and0, and1, shift0, shift1, shift2


def factor(sampling):
    return {1: 1, 2: 2, 3: 4}[sampling]


def _byte(b, k):
    return int(b[k]) if k < len(b) else 0


def _hit(num, sub, cam_w, sampling, fw, fh, out):
    if sampling == 1:
        x, y = num % fw, num // fw
    else:
        a0, a1, s0, s1, s2 = SUPERRES[sampling]
        x = ((num & 4095) << s0) | ((sub & a0) >> s1)
        y = ((num >> 12) << s0) | ((sub & a1) >> s2)
    if x < fw and y < fh:
        out.append((x, y))


def decode(frame, cam_w, cam_h, bits, sampling):
    """Electron positions (x, y) of one EER fraction, in stream order."""
    b = np.asarray(frame, np.uint8)
    nb = len(b)
    cam = cam_w * cam_h
    fw, fh = cam_w * factor(sampling), cam_h * factor(sampling)
    out = []
    num = 0
    if bits == 7:
        bit = 0
        while True:
            b0 = bit >> 3
            if b0 >= nb:
                break
            chunk = (_byte(b, b0) | (_byte(b, b0 + 1) << 8) | (_byte(b, b0 + 2) << 16)
                     | (_byte(b, b0 + 3) << 24)) >> (bit & 7)
            for lo in (0, 11):
                r = (chunk >> lo) & 127
                bit += 7
                num += r
                if num >= cam:
                    return out
                if r == 127:
                    break
                bit += 4
                _hit(num, ((chunk >> (lo + 7)) & 15) ^ 0x0A, cam_w, sampling, fw, fh, out)
                num += 1
        return out
    byte0 = _byte(b, 0)
    pos = 0
    while pos < nb:
        c0, c1, c2 = _byte(b, pos), _byte(b, pos + 1), _byte(b, pos + 2)
        num += c0
        if num >= cam:
            break
        if c0 < 255:
            _hit(num, c1 & 0x0A, cam_w, sampling, fw, fh, out)
            num += 1
        if sampling == 1:
            r2, s2 = ((c1 >> 4) | (c2 << 4)) & 255, 0
        else:
            r2, s2 = ((c1 >> 4) | byte0) & 255, (byte0 >> 4) ^ 0x0A
        num += r2
        if num >= cam:
            break
        if r2 < 255:
            _hit(num, s2, cam_w, sampling, fw, fh, out)
            num += 1
        pos += 3
    return out


def render(data, offsets, cam_w, cam_h, bits, sampling, int_start=None, int_size=None):
    """uint8 rendered frames [n_out, cam_h*f, cam_w*f]."""
    n = len(offsets) - 1
    if int_size is None:
        int_start, int_size = list(range(n)), [1] * n
    f = factor(sampling)
    out = np.zeros((len(int_size), cam_h * f, cam_w * f), np.int64)
    for o, (st, sz) in enumerate(zip(int_start, int_size)):
        for k in range(st, st + sz):
            for x, y in decode(data[offsets[k]:offsets[k + 1]], cam_w, cam_h, bits, sampling):
                out[o, y, x] += 1
    return (out & 255).astype(np.uint8)


class _Bits:
    """Fields packed least significant bit first, in order."""

    def __init__(self):
        self.vals, self.widths = [], []

    def put(self, value, nbits):
        self.vals.append(value & ((1 << nbits) - 1))
        self.widths.append(nbits)

    def bytes(self):
        v = np.array(self.vals, np.int64)
        w = np.array(self.widths, np.int64)
        start = np.cumsum(w) - w
        total = int(w.sum())
        bits = np.zeros(total + 32 + 8, np.uint8)
        for j in range(int(w.max()) if len(w) else 0):
            m = j < w
            bits[start[m] + j] = (v[m] >> j) & 1
        return np.packbits(bits, bitorder="little")[:(total + 7) // 8 + 4].copy()


def encode_7bit(events, cam_pixels):
    """Code stream of events [(pixel index, sub-pixel nibble)], pixel indices
    strictly increasing: zero runs of at most 126 (127 = run continues), each
    electron's run followed by its nibble (stored ^ 0x0A), and a final run to
    the end of the camera."""
    bs = _Bits()
    cur = 0
    for q, sub in events:
        skip = q - cur
        assert skip >= 0
        while skip >= 127:
            bs.put(127, 7)
            skip -= 127
        bs.put(skip, 7)
        bs.put((sub ^ 0x0A) & 15, 4)
        cur = q + 1
    rest = cam_pixels - cur
    while rest >= 127:
        bs.put(127, 7)
        rest -= 127
    if rest > 0 or cur < cam_pixels:
        bs.put(rest, 7)
    return bs.bytes()


def encode_8bit(events, cam_pixels):
    """8-bit counted stream: codes of 8 bits (255 = run continues), two per
    three bytes (byte0 = first run, byte1 high nibble + byte2 low nibble =
    second run), then a final run to the end of the camera."""
    codes = []
    cur = 0
    for q in events:
        skip = q - cur
        while skip >= 255:
            codes.append(255)
            skip -= 255
        codes.append(skip)
        cur = q + 1
    rest = cam_pixels - cur
    while rest >= 255:
        codes.append(255)
        rest -= 255
    codes.append(rest)
    if len(codes) % 2:
        codes.append(255)
    out = []
    for i in range(0, len(codes), 2):
        a, b = codes[i], codes[i + 1]
        out += [a, (b & 15) << 4, (b >> 4) & 15]
    return np.array(out, np.uint8)
