"""Pins of the ingest oracle (oracle/ingest_oracle.py) and the host-only MRC
reader of the C ABI (mc_mrc_info / mc_mrc_read)."""
import numpy as np
import pytest

import ingest_oracle as IO
from motioncor3_amd import lib as L
from motioncor3_amd.mrc import pack4, write_mrc


@pytest.mark.parametrize("h,w", [(5, 7), (6, 4), (3, 3)])
def test_rotation_is_clockwise_quarter_turns(h, w):
    img = np.arange(h * w, dtype=np.float32).reshape(h, w)
    for k in range(4):
        assert np.array_equal(IO.rotate(img, k), np.rot90(img, -k)), k


def test_flips():
    img = np.arange(12, dtype=np.float32).reshape(3, 4)
    assert np.array_equal(IO.flip(img, 1), np.flipud(img))
    assert np.array_equal(IO.flip(img, 2), np.fliplr(img))
    assert np.array_equal(IO.flip(img, 0), img)


def test_unpack4_nibble_order():
    rng = np.random.default_rng(1)
    vals = rng.integers(0, 16, (2, 3, 7)).astype(np.uint8)
    packed = pack4(vals)
    assert packed.shape == (2, 3, 4)
    # bit by bit: byte b of a row holds x = 2b in bits 0-3 and x = 2b+1 in bits 4-7
    for n in range(2):
        for y in range(3):
            for x in range(7):
                byte = int(packed[n, y, x // 2])
                want = sum(((byte >> (4 * (x % 2) + i)) & 1) << i for i in range(4))
                assert want == vals[n, y, x]
    assert np.array_equal(IO.unpack4(packed, 7), vals)


def test_apply_refs_small_case():
    raw = np.arange(2 * 4 * 6, dtype=np.uint8).reshape(2, 4, 6)
    gain = np.full((4, 6), 2.0, np.float32)
    dark = np.ones((4, 6), np.float32)
    out = IO.apply_refs(raw, gain, dark, W=2, H=2)
    # centre crop starts at x = 2, y = 1
    assert out.shape == (2, 2, 2) and out.dtype == np.float32
    assert out[0].tolist() == [[(8 - 1) * 2, (9 - 1) * 2], [(14 - 1) * 2, (15 - 1) * 2]]
    assert out[1, 0, 0] == (24 + 8 - 1) * 2
    assert np.array_equal(IO.apply_refs(raw, None, None, 6, 4), raw.astype(np.float32))


def test_post_process_dark_rule():
    g = np.arange(12, dtype=np.float32).reshape(3, 4) + 1
    d = np.arange(12, dtype=np.float32).reshape(3, 4)
    # flip only: the reference leaves the dark as it is (quirk), else flips it
    g1, d1 = IO.post_process(g, d, 0, 1, 0, quirks=True)
    assert np.array_equal(g1, np.flipud(g)) and np.array_equal(d1, d)
    g2, d2 = IO.post_process(g, d, 0, 1, 0, quirks=False)
    assert np.array_equal(d2, np.flipud(d))
    # rotation turns both; inverse maps 0 to 0
    g3, d3 = IO.post_process(np.where(g > 5, g, 0), d, 1, 0, 1)
    assert np.array_equal(d3, np.rot90(d, -1))
    gg = np.rot90(np.where(g > 5, g, 0), -1)
    with np.errstate(divide="ignore"):
        want = np.where(gg == 0, 0, np.float32(1) / gg).astype(np.float32)
    assert np.array_equal(g3, want)
    # no gain: nothing happens to the dark
    assert np.array_equal(IO.post_process(None, d, 1, 1, 0)[1], d)


@pytest.mark.parametrize("mode", [0, 5, 1, 6, 2, 101])
@pytest.mark.parametrize("ext,swap", [(0, False), (192, False), (0, True)])
def test_mrc_reader_round_trip(tmp_path, mode, ext, swap):
    if swap and mode == 101:
        pytest.skip("4-bit data has no byte order")
    rng = np.random.default_rng(mode + ext)
    n, sy, sx = 3, 5, 7
    hi = {0: 255, 5: 255, 1: 3000, 6: 60000, 2: 100, 101: 15}[mode]
    frames = rng.integers(0, hi + 1, (n, sy, sx))
    if mode == 1:
        frames = frames - 1500
    if mode == 2:
        frames = rng.normal(0, 10, (n, sy, sx)).astype(np.float32)
    path = str(tmp_path / ("m%d.mrc" % mode))
    write_mrc(path, frames, mode, ext_bytes=ext, swap=swap)
    info = L.mrc_info(path)
    assert (info["sx"], info["sy"], info["n"], info["mode"]) == (sx, sy, n, mode)
    assert info["data_offset"] == 1024 + ext
    raw = L.mrc_read(path, 1, 2)
    if mode == 101:
        assert np.array_equal(IO.unpack4(raw.reshape(2, sy, 4), sx), frames[1:3].astype(np.uint8))
    else:
        dt = {0: np.uint8, 5: np.uint8, 1: np.int16, 6: np.uint16, 2: np.float32}[mode]
        assert np.array_equal(raw.view(dt).reshape(2, sy, sx), frames[1:3].astype(dt))
    with pytest.raises(RuntimeError):
        L.mrc_read(path, 2, 2)


@pytest.mark.parametrize("swap", [False, True])
def test_mrc_stacks(tmp_path, swap):
    """-StackZ (MrcUtil/CLoadStack.cpp:40-55,95-112): the flag splits the
    file into nz / StackZ movies; without it a header mz above 1 bounds the
    one movie; otherwise the movie is every section."""
    import struct
    path = str(tmp_path / "s.mrc")
    write_mrc(path, np.zeros((6, 4, 8)), 2, swap=swap)
    assert L.mrc_stacks(path) == (6, 1)
    assert L.mrc_stacks(path, 3) == (3, 2)
    assert L.mrc_stacks(path, 4) == (4, 1)
    with pytest.raises(RuntimeError):
        L.mrc_stacks(path, 7)
    raw = bytearray(open(path, "rb").read())
    for mz, want in ((2, (2, 1)), (1, (6, 1))):
        struct.pack_into(">i" if swap else "<i", raw, 36, mz)
        open(path, "wb").write(bytes(raw))
        assert L.mrc_stacks(path) == want, mz
        assert L.mrc_stacks(path, 3) == (3, 2)
