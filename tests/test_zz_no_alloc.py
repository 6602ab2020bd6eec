"""mc_local_warp_noalloc (no device allocation; its tables live in a
file-scope device array) against mc_local_warp, bit for bit: two calls in a
row, the second with a smaller input, then one without bad flags after one
with them (the array keeps the previous call's contents); and the size
limit."""
import ctypes as C

import numpy as np
import pytest

from motioncor3_amd import lib as L

pytestmark = pytest.mark.gpu


def _noalloc(ctx, frames, centers, shifts, bad):
    N, H, W = frames.shape
    P = len(centers)
    out = ctx.empty(frames.shape, np.float32)
    cen = np.ascontiguousarray(centers, dtype=np.float32)
    sh = np.ascontiguousarray(shifts, dtype=np.float32)
    bp = None
    if bad is not None:
        bad = np.ascontiguousarray(bad, dtype=np.uint8)
        bp = bad.ctypes.data_as(C.POINTER(C.c_ubyte))
    rc = ctx.lib.mc_local_warp_noalloc(ctx.ptr, frames.vp(), out.vp(), W, H, N, P,
                                       cen.ctypes.data_as(C.POINTER(C.c_float)),
                                       sh.ctypes.data_as(C.POINTER(C.c_float)), bp)
    return rc, out


def _case(rng, N, H, W, P, with_bad):
    frames = rng.standard_normal((N, H, W)).astype(np.float32)
    g = int(np.sqrt(P))
    cen = np.array([[(i % g + 0.5) * W / g, (i // g + 0.5) * H / g] for i in range(P)], np.float32)
    sh = rng.uniform(-4, 4, (N, P, 2)).astype(np.float32)
    bad = (rng.uniform(0, 1, (N, P)) < 0.2).astype(np.uint8) if with_bad else None
    return frames, cen, sh, bad


def test_noalloc_warp_equals_allocating_warp(ctx):
    rng = np.random.default_rng(17)
    cases = [_case(rng, 6, 400, 520, 25, True),    # first call
             _case(rng, 3, 210, 300, 9, True),     # smaller: fewer frames, patches, pixels
             _case(rng, 4, 256, 256, 16, False)]   # no flags after flags: all cleared
    for k, (frames, cen, sh, bad) in enumerate(cases):
        print("no-alloc warp case", k, frames.shape, len(cen), flush=True)
        d = ctx.upload(frames)
        ref = ctx.local_warp(d, cen, sh, bad).numpy()
        rc, out = _noalloc(ctx, d, cen, sh, bad)
        assert rc == L.MC_OK, ctx.lib.mc_last_error(ctx.ptr)
        assert np.array_equal(out.numpy(), ref), k


def test_noalloc_warp_rejects_tables_beyond_its_scratch(ctx):
    """64 patches x 8000 frames need 4.6 MB of tables, more than the 4 MiB
    array: refused before any GPU work, with a message."""
    N, P = 8000, 64
    frames = ctx.empty((N, 2, 2), np.float32)
    cen = np.zeros((P, 2), np.float32)
    sh = np.zeros((N, P, 2), np.float32)
    rc, _ = _noalloc(ctx, frames, cen, sh, None)
    assert rc == L.MC_E_ARG
    assert b"scratch" in ctx.lib.mc_last_error(ctx.ptr)
