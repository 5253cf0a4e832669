"""FP64 CPU oracle of the learned PP stand-in's network (TEST INFRASTRUCTURE
ONLY: loaded by tests/ as the checker of cube2_amd/learned.py).

The network has no counterpart in the reference.  Its definition: patches of
10^3 fine cells with 4 channels (log(1 + count) and the mean offset of the
cell's particles from its centre), channels-last [patch][z][y][x][c]; four
3^3 valid convolutions 4 -> 32 -> 32 -> 32 -> 3 channels (10 -> 8 -> 6 -> 4
-> 2 cells), ReLU after the first three; layer weights W[k][c_out] with
k = ((dz*3 + dy)*3 + dx)*c_in + c, biases b[c_out].  A sample's prediction
is the cloud-in-cell combination of its 2^3 outputs with its fractions
(``cic``).  Forward and backward are plain numpy in FP64; the tests pin the
backward against central finite differences of the forward.
"""
from __future__ import annotations

import numpy as np

CHANNELS = (4, 32, 32, 32, 3)
PATCH = 10


def im2col(x: np.ndarray) -> np.ndarray:
    """x [B][D][D][D][C] -> col [B*(D-2)^3][27*C] (offset-major, channel fastest)."""
    B, D, _, _, C = x.shape
    P = D - 2
    cols = np.empty((B, P, P, P, 27, C), x.dtype)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                cols[:, :, :, :, (dz * 3 + dy) * 3 + dx, :] = x[:, dz:dz + P, dy:dy + P, dx:dx + P, :]
    return cols.reshape(B * P * P * P, 27 * C)


def col2im(dcol: np.ndarray, B: int, D: int, C: int) -> np.ndarray:
    """Adjoint of im2col."""
    P = D - 2
    d = dcol.reshape(B, P, P, P, 27, C)
    dx_ = np.zeros((B, D, D, D, C), dcol.dtype)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                dx_[:, dz:dz + P, dy:dy + P, dx:dx + P, :] += d[:, :, :, :, (dz * 3 + dy) * 3 + dx, :]
    return dx_


def init_params(seed: int = 0):
    """He-uniform weights, zero biases (FP64)."""
    rng = np.random.default_rng(seed)
    params = []
    for cin, cout in zip(CHANNELS[:-1], CHANNELS[1:]):
        lim = np.sqrt(6.0 / (27 * cin))
        params.append((rng.uniform(-lim, lim, (27 * cin, cout)), np.zeros(cout)))
    return params


def cic_weights(frac: np.ndarray) -> np.ndarray:
    """[n][8] weights of the 2^3 corners (corner c = x + 2y + 4z) for
    fractions frac [n][3]."""
    f = np.asarray(frac, np.float64)
    w = np.empty((f.shape[0], 8))
    for c in range(8):
        w[:, c] = ((f[:, 0] if c & 1 else 1 - f[:, 0]) * (f[:, 1] if c & 2 else 1 - f[:, 1])
                   * (f[:, 2] if c & 4 else 1 - f[:, 2]))
    return w


def cic(out: np.ndarray, frac: np.ndarray) -> np.ndarray:
    """Predictions [n][3] from outputs [n][2][2][2][3] (or [n][24])."""
    o = np.asarray(out).reshape(-1, 8, 3)
    return np.einsum("nc,ncd->nd", cic_weights(frac), o)


def forward(params, x: np.ndarray):
    """x [B][D][D][D][c_in] -> (out [B][(D-8)^3 * 3], cache for backward)."""
    B = x.shape[0]
    acts, cols = [x], []
    h = x
    for l, (W, b) in enumerate(params):
        D, C = h.shape[1], h.shape[4]
        col = im2col(h)
        y = col @ W + b
        if l < len(params) - 1:
            y = np.maximum(y, 0.0)
        P = D - 2
        h = y.reshape(B, P, P, P, W.shape[1])
        acts.append(h)
        cols.append(col)
    return h.reshape(B, -1), (acts, cols)


def backward(params, cache, dout: np.ndarray):
    """Gradients [(dW, db)] of sum(dout * out)."""
    acts, cols = cache
    B = dout.shape[0]
    g = dout.reshape(acts[-1].shape)
    grads = [None] * len(params)
    for l in range(len(params) - 1, -1, -1):
        W, _ = params[l]
        y = acts[l + 1].reshape(-1, W.shape[1])
        dy = g.reshape(-1, W.shape[1])
        if l < len(params) - 1:
            dy = dy * (y > 0)
        grads[l] = (cols[l].T @ dy, dy.sum(0))
        if l > 0:
            D, C = acts[l].shape[1], acts[l].shape[4]
            g = col2im(dy @ W.T, B, D, C)
    return grads


def mse(out: np.ndarray, target: np.ndarray):
    r = out - target
    return float((r * r).mean()), 2.0 * r / r.size
