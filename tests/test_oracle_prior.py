"""Pins of the prior network's CPU restatement (oracle/prior_oracle.py):
finite differences for the gradient, hand-computed BF16 roundings, and a
closed-form forward pass."""
import numpy as np

import prior_oracle as PO


def _net(rng, dims):
    ws = []
    for l in range(len(dims) - 1):
        ws.append(rng.normal(0, 1.0 / np.sqrt(dims[l]), dims[l] * dims[l + 1]))
        ws.append(rng.normal(0, 0.1, dims[l + 1]))
    return np.concatenate(ws)


def test_bf16_round_to_nearest_even():
    cases = [
        (1.0, 1.0),
        (1.0 + 2.0 ** -8, 1.0),                       # halfway, ties to even (down)
        (1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -6),       # halfway, ties to even (up)
        (1.0 + 2.0 ** -8 + 2.0 ** -20, 1.0 + 2.0 ** -7),  # just above halfway
        (-(1.0 + 2.0 ** -9), -1.0),
        (2.0 ** 100 * (1.0 + 2.0 ** -8), 2.0 ** 100),  # large values keep their exponent
        (2.0 ** -100 * (1.0 + 3 * 2.0 ** -8), 2.0 ** -100 * (1.0 + 2.0 ** -6)),
        (0.0, 0.0),
    ]
    for x, want in cases:
        got = PO.bf16_round(np.array([x], np.float32))[0]
        assert got == want, (x, got, want)
    # every result is a BF16 value (low 16 bits zero) within half a BF16 ulp
    rng = np.random.default_rng(3)
    x = rng.normal(0, 10, 10000).astype(np.float32)
    r = PO.bf16_round(x).astype(np.float32)
    assert np.all(r.view(np.uint32) & 0xFFFF == 0)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 + 1e-38)


def test_forward_closed_form():
    # 2-2-2-2-1 network evaluated by hand
    dims = [2, 2, 2, 2, 1]
    W0, b0 = np.array([[1.0, -1.0], [2.0, 0.5]]), np.array([0.0, 1.0])
    W1, b1 = np.eye(2), np.array([-1.0, 0.0])
    W2, b2 = np.array([[1.0, 1.0], [1.0, -1.0]]), np.zeros(2)
    W3, b3 = np.array([[2.0], [3.0]]), np.array([0.5])
    w = np.concatenate([W0.ravel(), b0, W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])
    x = np.array([[1.0, 1.0]])
    # layer 0: z = [1+2, -1+0.5+1] = [3, 0.5]; layer 1: [2, 0.5]; layer 2: [2.5, 1.5]
    # output: 2*2.5 + 3*1.5 + 0.5 = 10
    y = PO.forward(PO.split(w, dims), x)
    assert np.allclose(y, [[10.0]])
    x = np.array([[-1.0, 0.0]])
    # layer 0: [-1, 2] -> relu [0, 2]; layer 1: [-1, 2] -> [0, 2]; layer 2: [2, -2] -> [2, 0]
    # output: 4 + 0 + 0.5
    assert np.allclose(PO.forward(PO.split(w, dims), x), [[4.5]])


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(11)
    dims = [7, 6, 6, 6, 3]
    w = _net(rng, dims)
    X = rng.normal(0, 1, (9, 7))
    T = rng.normal(0, 1, (9, 3))
    loss, g = PO.loss_grad(PO.split(w, dims), X, T)
    assert loss > 0
    h = 1e-6
    # every parameter of every block (the net is small)
    num = np.zeros_like(w)
    for i in range(w.size):
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        lp = PO.loss_grad(PO.split(wp, dims), X, T)[0]
        lm = PO.loss_grad(PO.split(wm, dims), X, T)[0]
        num[i] = (lp - lm) / (2 * h)
    assert np.max(np.abs(num - g)) <= 1e-7 * max(1.0, np.max(np.abs(g))), np.max(np.abs(num - g))


def test_bf16_emulation_is_close_to_fp64():
    rng = np.random.default_rng(5)
    dims = [40, 16, 16, 16, 8]
    w = _net(rng, dims)
    X = rng.normal(0, 1, (20, 40))
    T = rng.normal(0, 1, (20, 8))
    l64, g64 = PO.loss_grad(PO.split(w, dims), X, T)
    l16, g16 = PO.loss_grad(PO.split(w, dims), X, T, bf16=True)
    assert abs(l16 - l64) / l64 < 2e-2
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 3e-2
    assert not np.array_equal(g16, g64)
