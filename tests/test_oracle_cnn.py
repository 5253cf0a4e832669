"""The learned stand-in's FP64 network oracle pinned by central finite
differences (every layer's weights and biases), by im2col/col2im being
adjoint, and by a one-layer case against a direct 3^3 convolution sum
(CPU only)."""
from __future__ import annotations

import numpy as np

from oracle import cnn


def test_im2col_and_col2im_are_adjoint():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 6, 6, 6, 3))
    c = rng.normal(size=(2 * 4 * 4 * 4, 27 * 3))
    assert np.isclose((cnn.im2col(x) * c).sum(), (x * cnn.col2im(c, 2, 6, 3)).sum(), rtol=1e-12)


def test_one_layer_is_a_direct_convolution():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 5, 5, 5, 2))
    W = rng.normal(size=(27 * 2, 4))
    y = (cnn.im2col(x) @ W).reshape(3, 3, 3, 4)
    z, yy, xx, co = 1, 2, 0, 3
    s = sum(x[0, z + dz, yy + dy, xx + dx, c] * W[((dz * 3 + dy) * 3 + dx) * 2 + c, co]
            for dz in range(3) for dy in range(3) for dx in range(3) for c in range(2))
    assert np.isclose(y[z, yy, xx, co], s, rtol=1e-12)


def test_cic_weights_interpolate_linear_fields_exactly():
    """The 2^3 combination reproduces a linear field sampled at the corners
    (cloud-in-cell), and its weights sum to 1."""
    rng = np.random.default_rng(5)
    frac = rng.uniform(0, 1, (20, 3))
    a, g = rng.normal(size=3), rng.normal(size=(3, 3))
    corners = np.array([[c & 1, (c >> 1) & 1, c >> 2] for c in range(8)], np.float64)
    out = (a[None, None, :] + corners[None, :, :] @ g[None, :, :]).repeat(20, 0)
    assert np.allclose(cnn.cic_weights(frac).sum(1), 1.0)
    assert np.allclose(cnn.cic(out, frac), a + frac @ g, rtol=1e-13)


def test_backward_matches_central_differences():
    """Through the cloud-in-cell combination of the 2^3 outputs, as trained."""
    rng = np.random.default_rng(2)
    params = cnn.init_params(3)
    params = [(W, rng.normal(0, 0.1, b.shape)) for W, b in params]
    x = rng.uniform(0, 2, (3, 10, 10, 10, 4))
    t = rng.normal(size=(3, 3))
    frac = rng.uniform(0, 1, (3, 3))
    w = cnn.cic_weights(frac)
    out, cache = cnn.forward(params, x)
    _, dpred = cnn.mse(cnn.cic(out, frac), t)
    dout = (w[:, :, None] * dpred[:, None, :]).reshape(3, 24)
    grads = cnn.backward(params, cache, dout)

    def loss(p):
        return cnn.mse(cnn.cic(cnn.forward(p, x)[0], frac), t)[0]

    h = 1e-6
    for l in range(4):
        for which in (0, 1):
            arr = params[l][which]
            for idx in [tuple(rng.integers(0, s) for s in arr.shape) for _ in range(4)]:
                keep = arr[idx]
                arr[idx] = keep + h
                lp = loss(params)
                arr[idx] = keep - h
                lm = loss(params)
                arr[idx] = keep
                fd = (lp - lm) / (2 * h)
                g = grads[l][which][idx]
                assert abs(fd - g) <= 1e-6 * max(1.0, abs(fd)), (l, which, idx, fd, g)
