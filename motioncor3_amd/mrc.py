"""MRC stack files: a writer for raw camera stacks and float images.

Header layout of the 1024-byte MRC main header as LibSrc/Mrcfile reads it
(Header/CMainHeader.cpp, Header/CLoadMainHeader.cpp:18-40): words 1-4 nx, ny,
nz, mode; 8-10 mx, my, mz; 11-13 cell size; 14-16 axis order (1 2 3);
20-22 min, max, mean; 24 nsymbt (extended header bytes); 53 "MAP "; 54
machine stamp; 55 rms.  Mode 101 stores 4-bit values two per byte, even x
in the low nibble, rows of (nx + 1) // 2 bytes.  Reading goes through the
library (motioncor3_amd.lib.mrc_info / mrc_read).
"""
import numpy as np

_DTYPES = {0: np.uint8, 5: np.uint8, 1: np.int16, 6: np.uint16, 2: np.float32}


def pack4(frames):
    """[n, sy, sx] values 0..15 -> [n, sy, (sx + 1) // 2] packed bytes."""
    f = np.asarray(frames, np.uint8)
    n, sy, sx = f.shape
    if sx % 2:
        f = np.concatenate([f, np.zeros((n, sy, 1), np.uint8)], axis=2)
    return (f[:, :, 0::2] & 0xF) | ((f[:, :, 1::2] & 0xF) << 4)


def write_mrc(path, frames, mode, ext_bytes=0, swap=False):
    """Write [n, sy, sx] frames (values of the mode's type; 0..15 for 101)."""
    frames = np.asarray(frames)
    n, sy, sx = frames.shape
    if mode == 101:
        data = pack4(frames)
    else:
        data = np.ascontiguousarray(frames, _DTYPES[mode])
    fv = frames.astype(np.float64)
    hdr = np.zeros(256, np.int32)
    hdr[0:4] = (sx, sy, n, mode)
    hdr[7:10] = (sx, sy, n)
    hdr[13:16] = (1, 2, 3)
    hdr[23] = ext_bytes
    hf = hdr.view(np.float32)
    hf[10:13] = (sx, sy, n)
    hf[19:22] = (fv.min(), fv.max(), fv.mean())
    hf[54] = fv.std()
    hdr[52] = np.frombuffer(b"MAP ", np.int32)[0]
    hdr[53] = np.frombuffer(bytes([0x44, 0x44, 0, 0]), np.int32)[0]
    body = data
    if swap:
        hdr = hdr.byteswap()
        body = data.byteswap() if data.dtype.itemsize > 1 else data
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(b"\0" * ext_bytes)
        f.write(np.ascontiguousarray(body).tobytes())
