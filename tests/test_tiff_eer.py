"""TIFF and EER containers (host side of the C ABI: mc_tiff_info / mc_tiff_read
/ mc_eer_info / mc_eer_read) and the EER oracle (oracle/eer_oracle.py).

TIFF files come from Pillow's libtiff writer (an independent encoder: LZW,
Deflate, PackBits, multi-strip, horizontal predictor) and, for what it cannot
write (tiles, big-endian samples, BigTIFF), from tests/tiffw.py.  The
reference stores TIFF rows bottom-up (TiffUtil/CLoadTiffImage.cpp:72-160),
so every decoded frame must equal the flipped image."""
import numpy as np
import pytest
from PIL import Image

import eer_oracle as E
import tiffw
from motioncor3_amd import lib as L

DT = {"L": np.uint8, "I;16": np.uint16, "F": np.float32}
MODE = {"L": 0, "I;16": 6, "F": 2}


def _frames(mode, n=3, h=37, w=53, seed=0):
    rng = np.random.default_rng(seed)
    hi = {"L": 256, "I;16": 65536}.get(mode)
    if hi:
        return [rng.integers(0, hi, (h, w)).astype(DT[mode]) for _ in range(n)]
    return [rng.normal(size=(h, w)).astype(np.float32) for _ in range(n)]


@pytest.mark.parametrize("mode", ["L", "I;16", "F"])
@pytest.mark.parametrize("comp", [None, "tiff_lzw", "tiff_adobe_deflate", "packbits"])
def test_pillow_tiff_stacks(tmp_path, mode, comp):
    fr = _frames(mode)
    path = str(tmp_path / "m.tif")
    ims = [Image.fromarray(x) for x in fr]
    ims[0].save(path, save_all=True, append_images=ims[1:], compression=comp)
    info = L.tiff_info(path)
    assert (info["sx"], info["sy"], info["n"], info["mode"]) == (53, 37, 3, MODE[mode])
    got = L.tiff_read(path).view(DT[mode]).reshape(3, 37, 53)
    for k in range(3):
        assert np.array_equal(got[k], np.flipud(fr[k])), k
    one = L.tiff_read(path, 1, 1).view(DT[mode]).reshape(37, 53)
    assert np.array_equal(one, np.flipud(fr[1]))


@pytest.mark.parametrize("mode", ["L", "I;16"])
def test_pillow_multistrip_and_predictor(tmp_path, mode):
    fr = _frames(mode, n=2, h=45, w=61, seed=3)
    ims = [Image.fromarray(x) for x in fr]
    bs = np.dtype(DT[mode]).itemsize
    for kw in (dict(strip_size=61 * bs * 8), dict(tiffinfo={317: 2}),
               dict(tiffinfo={317: 2}, strip_size=61 * bs * 5)):
        path = str(tmp_path / "p.tif")
        ims[0].save(path, save_all=True, append_images=ims[1:], compression="tiff_lzw", **kw)
        got = L.tiff_read(path).view(DT[mode]).reshape(2, 45, 61)
        for k in range(2):
            assert np.array_equal(got[k], np.flipud(fr[k])), (kw, k)


@pytest.mark.parametrize("big_endian,bigtiff", [(False, False), (True, False), (False, True), (True, True)])
def test_tiles_byte_order_and_bigtiff(tmp_path, big_endian, bigtiff):
    rng = np.random.default_rng(5)
    h, w, tw, th = 37, 53, 16, 16
    img = rng.integers(0, 65536, (h, w)).astype(np.uint16)
    dt = ">u2" if big_endian else "<u2"
    tiles = []
    for ty in range(0, h, th):
        for tx in range(0, w, tw):
            t = np.zeros((th, tw), np.uint16)
            blk = img[ty:ty + th, tx:tx + tw]
            t[:blk.shape[0], :blk.shape[1]] = blk
            tiles.append(t.astype(dt).tobytes())
    strips = [img[y:y + 8].astype(dt).tobytes() for y in range(0, h, 8)]
    path = str(tmp_path / "t.tif")
    tiffw.write(path, [dict(w=w, h=h, bits=16, comp=1, blocks=tiles, tile=(tw, th)),
                       dict(w=w, h=h, bits=16, comp=1, blocks=strips, rps=8)],
                big_endian=big_endian, bigtiff=bigtiff)
    info = L.tiff_info(path)
    assert (info["n"], info["mode"]) == (2, 6)
    got = L.tiff_read(path).view(np.uint16).reshape(2, h, w)
    assert np.array_equal(got[0], np.flipud(img))
    assert np.array_equal(got[1], np.flipud(img))


def test_tiff_errors(tmp_path):
    bad = tmp_path / "x.tif"
    bad.write_bytes(b"not a tiff file at all")
    with pytest.raises(RuntimeError):
        L.tiff_info(str(bad))
    with pytest.raises(RuntimeError):
        L.tiff_info(str(tmp_path / "missing.tif"))
    # a strip that runs past the end of the file, and a truncated file
    good = tmp_path / "g.tif"
    tiffw.write(str(good), [dict(w=8, h=4, bits=8, comp=1, rps=4, blocks=[bytes(32)])])
    raw = bytearray(good.read_bytes())
    short = tmp_path / "s.tif"
    short.write_bytes(bytes(raw[:12]))
    with pytest.raises(RuntimeError):
        L.tiff_info(str(short))
    # StripByteCounts (tag 279, LONG, count 1) patched to a huge value
    import struct
    ifd = struct.unpack_from("<I", raw, 4)[0]
    n = struct.unpack_from("<H", raw, ifd)[0]
    for e in range(n):
        p = ifd + 2 + 12 * e
        if struct.unpack_from("<H", raw, p)[0] == 279:
            struct.pack_into("<I", raw, p + 8, 0x7fffffff)
    huge = tmp_path / "h.tif"
    huge.write_bytes(bytes(raw))
    with pytest.raises(RuntimeError):
        L.tiff_info(str(huge))
    assert L.tiff_info(str(good))["n"] == 1


# ---- EER ----

def _events(rng, cam, count):
    return np.sort(rng.choice(cam, size=count, replace=False))


@pytest.mark.parametrize("cam_w,cam_h", [(64, 48), (4096, 3)])
def test_eer_7bit_counted_round_trip(cam_w, cam_h):
    rng = np.random.default_rng(1)
    cam = cam_w * cam_h
    for count in (0, 1, 37, 400):
        q = _events(rng, cam, count)
        sub = rng.integers(0, 16, count)
        s = E.encode_7bit(list(zip(q.tolist(), sub.tolist())), cam)
        got = E.decode(s, cam_w, cam_h, 7, 1)
        assert got == [(int(p) % cam_w, int(p) // cam_w) for p in q], count


def test_eer_7bit_long_runs_and_edges():
    cam_w, cam_h = 4096, 2
    cam = cam_w * cam_h
    # runs of exactly 127, 254 and a multiple of 127 to the end; first and last pixel
    q = [0, 127, 128, 128 + 254 + 1, cam - 1 - 127 * 3, cam - 1]
    s = E.encode_7bit([(p, 0) for p in q], cam)
    assert E.decode(s, cam_w, cam_h, 7, 1) == [(p % cam_w, p // cam_w) for p in q]


@pytest.mark.parametrize("sampling", [2, 3])
def test_eer_7bit_superres_round_trip(sampling):
    rng = np.random.default_rng(2)
    cam_w, cam_h = 4096, 4
    f = E.factor(sampling)
    q = _events(rng, cam_w * cam_h, 300)
    sub = rng.integers(0, 16, 300)
    s = E.encode_7bit(list(zip(q.tolist(), sub.tolist())), cam_w * cam_h)
    got = E.decode(s, cam_w, cam_h, 7, sampling)
    # sub-pixel nibble: sampling 2 uses bit 1 (x) and bit 3 (y); sampling 3
    # uses bits 0-1 (x) and bits 2-3 (y)
    want = []
    for p, sb in zip(q.tolist(), sub.tolist()):
        col, row = p % 4096, p // 4096
        if sampling == 2:
            want.append((col * 2 + ((sb >> 1) & 1), row * 2 + ((sb >> 3) & 1)))
        else:
            want.append((col * 4 + (sb & 3), row * 4 + ((sb >> 2) & 3)))
    assert got == want
    assert all(x < cam_w * f and y < cam_h * f for x, y in got)


def test_eer_8bit_counted_round_trip():
    rng = np.random.default_rng(4)
    cam_w, cam_h = 100, 70
    for count in (0, 5, 500):
        q = _events(rng, cam_w * cam_h, count)
        s = E.encode_8bit(q.tolist(), cam_w * cam_h)
        assert E.decode(s, cam_w, cam_h, 8, 1) == [(int(p) % cam_w, int(p) // cam_w) for p in q]


def test_eer_8bit_superres_follows_the_reference_walk():
    # hand trace of CDecodeEerFrame::mDo8BitsSuperRes on bytes
    # [0x05, 0x2A, 0x00, 0x03, 0x10, 0x00] with a 4096 x 2 camera, sampling 2:
    # triple 0: p1 = 5 -> pixel 5, sub = 0x2A & 0x0A = 0x0A -> x = 10 | 1,
    #   y = 0 | 1; p2 = (0x2A >> 4) | byte0 = 2 | 5 = 7 -> pixel 13, sub =
    #   (5 >> 4) ^ 0x0A = 0x0A -> (27, 1)
    # triple 1: p1 = 3 -> pixel 17, sub = 0x10 & 0x0A = 0 -> (34, 0);
    #   p2 = (0x10 >> 4) | 5 = 5 -> pixel 23, sub 0x0A -> (47, 1)
    s = np.array([0x05, 0x2A, 0x00, 0x03, 0x10, 0x00], np.uint8)
    assert E.decode(s, 4096, 2, 8, 2) == [(11, 1), (27, 1), (34, 0), (47, 1)]


def test_eer_render_integration_and_byte_wrap():
    cam_w, cam_h = 32, 8
    cam = cam_w * cam_h
    streams = [E.encode_7bit([(7, 0), (9, 0)], cam) for _ in range(300)]
    off = np.cumsum([0] + [len(s) for s in streams])
    data = np.concatenate(streams)
    fr = E.render(data, off, cam_w, cam_h, 7, 1, int_start=[0, 100], int_size=[100, 200])
    assert fr.shape == (2, 8, 32)
    assert fr[0, 0, 7] == 100 and fr[1, 0, 7] == 200 - 0 and fr[1, 0, 9] == 200
    full = E.render(data, off, cam_w, cam_h, 7, 1, int_start=[0], int_size=[300])
    assert full[0, 0, 7] == 300 % 256 and full.sum() == 2 * (300 % 256)


@pytest.mark.parametrize("bits", [7, 8])
def test_eer_container(tmp_path, bits):
    rng = np.random.default_rng(6)
    cam_w, cam_h = 64, 40
    streams = []
    for k in range(5):
        q = _events(rng, cam_w * cam_h, 50 + 10 * k)
        streams.append(E.encode_7bit([(int(p), 3) for p in q], cam_w * cam_h) if bits == 7
                       else E.encode_8bit(q.tolist(), cam_w * cam_h))
    # each fraction split over two strips, as a camera may write it
    dirs = [dict(w=cam_w, h=cam_h, bits=1, comp=65001 if bits == 7 else 65000, rps=cam_h // 2,
                 blocks=[s[:len(s) // 2].tobytes(), s[len(s) // 2:].tobytes()]) for s in streams]
    path = str(tmp_path / "m.eer")
    tiffw.write(path, dirs)
    info = L.eer_info(path)
    assert (info["cam_w"], info["cam_h"], info["n"], info["bits"]) == (cam_w, cam_h, 5, bits)
    assert info["bytes"] == sum(len(s) for s in streams)
    data, off = L.eer_read(path)
    assert np.array_equal(off, np.cumsum([0] + [len(s) for s in streams]))
    assert np.array_equal(data, np.concatenate(streams))
    # a plain TIFF is not an EER file
    p2 = str(tmp_path / "plain.tif")
    Image.fromarray(np.zeros((4, 4), np.uint8)).save(p2)
    with pytest.raises(RuntimeError):
        L.eer_info(p2)


def test_rows_per_strip_default_is_one_strip(tmp_path):
    """RowsPerStrip = 2^32 - 1 (the TIFF default: one strip for the whole
    image) reads the image's own rows, not 2^30 rows of buffer."""
    rng = np.random.default_rng(8)
    h, w = 23, 41
    fr = [rng.integers(0, 256, (h, w)).astype(np.uint8) for _ in range(2)]
    path = str(tmp_path / "one.tif")
    tiffw.write(path, [dict(w=w, h=h, bits=8, comp=1, blocks=[x.tobytes()], rps=0xFFFFFFFF)
                       for x in fr])
    got = L.tiff_read(path).view(np.uint8).reshape(2, h, w)
    for k in range(2):
        assert np.array_equal(got[k], np.flipud(fr[k])), k
