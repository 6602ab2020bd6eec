"""FP64 CPU restatement of the learned shift prior's network (test
infrastructure: only tests, the smoke check and the benchmark's CPU leg use
it).

The prior has no counterpart in the reference (which starts every patch's
iterations from zero shifts, Align/CIterativeAlign.cpp:68), so nothing here
follows a reference file.  It restates motioncor3_amd/csrc/mc_prior.hip:
four dense layers in -> H -> H -> H -> out with ReLU between them, the MSE
loss mean((y - t)^2) and its gradient.  With ``bf16=True`` every operand of a
layer product (inputs, activations, weights, back-propagated errors) is
rounded to BF16 first, as the MFMA kernels see them; sums and products are
then FP64.  Bias gradients use the unrounded errors, as on the GPU.

Pinned by tests/test_oracle_prior.py: the FP64 gradient against central
finite differences of the FP64 loss (every parameter block), the BF16
rounding against hand-computed round-to-nearest-even cases, and the forward
pass against a direct evaluation of the layer formula on a 2-2-2-2-1 net.
"""
import numpy as np


def bf16_round(x):
    """Round float32 values to BF16 (round to nearest, ties to even), returned
    as float64.  Same bit arithmetic as mc_bf16_bits (NaNs stay NaN)."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    r = (r << 16).astype(np.uint32)
    out = r.view(np.float32).astype(np.float64)
    out[np.isnan(a)] = np.nan
    return out


def split(w, dims):
    """Parameter vector -> [(W0, b0), ..., (W3, b3)] (W row-major din x dout)."""
    w = np.asarray(w, np.float64)
    out, off = [], 0
    for l in range(len(dims) - 1):
        n = dims[l] * dims[l + 1]
        W = w[off: off + n].reshape(dims[l], dims[l + 1])
        off += n
        b = w[off: off + dims[l + 1]]
        off += dims[l + 1]
        out.append((W, b))
    assert off == w.size
    return out


def _r(x, bf16):
    return bf16_round(x) if bf16 else np.asarray(x, np.float64)


def forward(layers, X, bf16=False, keep=False):
    """Rows of X -> outputs; with keep also the per-layer (input, Z)."""
    a = _r(X, bf16)
    cache = []
    L = len(layers)
    for l, (W, b) in enumerate(layers):
        Z = a @ _r(W, bf16) + b
        cache.append((a, Z))
        if l < L - 1:
            a = _r(np.maximum(Z, 0.0), bf16)
        else:
            a = Z
    return (a, cache) if keep else a


def loss_grad(layers, X, T, bf16=False):
    """MSE loss and the gradient vector (same layout as the parameters)."""
    Y, cache = forward(layers, X, bf16, keep=True)
    T = np.asarray(T, np.float64)
    cnt = Y.size
    loss = float(np.sum((Y - T) ** 2) / cnt)
    dZ = 2.0 * (Y - T) / cnt
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        a, Z = cache[l]
        W = layers[l][0]
        d16 = _r(dZ, bf16)
        gW = a.T @ d16
        gb = dZ.sum(0)
        grads[l] = (gW, gb)
        if l == 0:
            break
        dH = d16 @ _r(W, bf16).T
        dZ = dH * (cache[l - 1][1] > 0)
    vec = np.concatenate([np.concatenate([g.ravel(), b]) for g, b in grads])
    return loss, vec
