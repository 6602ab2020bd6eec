"""Batched 2-D R2C/C2R on the GPU against numpy.fft in FP64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# frame, xcf, patch and feature-map sizes of the reference geometries
SIZES = [(64, 48), (128, 128), (340, 340), (576, 408), (292, 408), (410, 408), (1024, 1024),
         (2882, 2048), (4096, 4096), (5760, 4092), (11520, 8184)]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("W,H", SIZES)
def test_rfft2_matches_numpy(ctx, W, H):
    print("rfft2", W, H, flush=True)
    B = 2 if W * H < 5_000_000 else 1
    rng = np.random.default_rng(W + H)
    x = rng.standard_normal((B, H, W)).astype(np.float32)
    spec = ctx.rfft2(ctx.upload(x)).numpy()
    for b in range(B):
        ref = np.fft.rfft2(x[b].astype(np.float64))
        assert _rel(spec[b].T, ref) < 2e-6, (W, H)


@pytest.mark.parametrize("W,H", SIZES)
def test_irfft2_roundtrip(ctx, W, H):
    print("irfft2", W, H, flush=True)
    rng = np.random.default_rng(7 * W + H)
    x = rng.standard_normal((1, H, W)).astype(np.float32)
    d = ctx.upload(x)
    spec = ctx.rfft2(d, scale=1.0 / (W * H))
    back = ctx.irfft2(spec, W).numpy()
    assert _rel(back, x) < 2e-6
    # C2R of a non-Hermitian half spectrum follows numpy's irfft2 convention
    s = spec.numpy()[0].T.astype(np.complex128)
    s[0, 0] += 0.3j
    spec2 = ctx.upload(np.ascontiguousarray(s.T[None]).astype(np.complex64))
    out = ctx.irfft2(spec2, W).numpy()[0]
    ref = np.fft.irfft2(s, s=(H, W)) * (W * H)
    assert _rel(out, ref) < 2e-6


def test_clamped_c2r(ctx):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 64, 96)).astype(np.float32)
    spec = ctx.rfft2(ctx.upload(x), scale=1.0 / (64 * 96))
    out = ctx.irfft2(spec, 96, clamp=True).numpy()
    assert np.allclose(out, np.maximum(x, 0), atol=1e-5)
