"""NumPy restatement of the amplitude network's forward and backward passes.

TEST INFRASTRUCTURE ONLY (the checker of csrc/hd_mlp.hip).  There is no
reference counterpart (the reference has no learned start); the math is pinned
by central finite differences in float64 (tests/test_mlp_oracle.py).

Rounding model of the device kernels: BF16 (round to nearest even) for the
input features, the weights, the hidden activations and the backward deltas;
FP32 accumulation and FP32 biases, pre-activations and outputs.  With
`bf16=False` everything stays in float64 (for finite differences).
"""
from __future__ import annotations

import numpy as np

SHAPES = [(256, 64), (256, 256), (256, 256), (64, 256)]


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 to bfloat16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def unflatten(flat):
    out, off = [], 0
    for r, c in SHAPES:
        W = flat[off:off + r * c].reshape(r, c)
        off += r * c
        b = flat[off:off + r]
        off += r
        out.append((W, b))
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def features(states, params, n_bits, n_params):
    n = len(states)
    X = np.zeros((n, 64), dtype=np.float64)
    s = np.asarray(states, dtype=np.uint64)
    for k in range(n_bits):
        X[:, k] = ((s >> np.uint64(k)) & np.uint64(1)).astype(np.float64)
    X[:, n_bits:n_bits + n_params] = np.asarray(params, dtype=np.float64).reshape(n, n_params)
    return X


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(flat, states, params, n_bits, n_params, bf=True):
    layers = unflatten(np.asarray(flat, dtype=np.float64))
    r = (lambda a: bf16(a).astype(np.float64)) if bf else (lambda a: a)
    h = r(features(states, params, n_bits, n_params))
    hs, zs = [h], []
    for l in range(3):
        W, b = layers[l]
        z = h @ r(W).T + b
        zs.append(z)
        h = r(silu(z))
        hs.append(h)
    W, b = layers[3]
    out = h @ r(W).T + b
    return out, hs, zs


def loss_and_grads(flat, states, params, targets, n_bits, n_params, lam=1.0, bf=True):
    """Loss = mean_b sum_s w_bs [ (o_2s - a_bs)^2 + lam * softplus(-s_bs o_2s+1) ]
    over the S amplitudes of targets (n x 3S: a, s, w per amplitude) and its
    gradient in the flat layout."""
    layers = unflatten(np.asarray(flat, dtype=np.float64))
    r = (lambda a: bf16(a).astype(np.float64)) if bf else (lambda a: a)
    out, hs, zs = forward(flat, states, params, n_bits, n_params, bf)
    n = len(states)
    tg = np.asarray(targets, dtype=np.float64).reshape(n, -1, 3)
    loss = 0.0
    d = np.zeros_like(out)
    for st in range(tg.shape[1]):
        a, s, w = tg[:, st, 0], tg[:, st, 1], tg[:, st, 2]
        o0, o1 = out[:, 2 * st], out[:, 2 * st + 1]
        m = -s * o1
        loss += np.sum(w * ((o0 - a) ** 2 + lam * np.logaddexp(0.0, m))) / n
        d[:, 2 * st] = 2 * w * (o0 - a) / n
        d[:, 2 * st + 1] = -lam * w * s / (1.0 + np.exp(-m)) / n
    d = r(d)
    grads = [None] * 4
    grads[3] = (d.T @ hs[3], d.sum(axis=0))
    for l in (2, 1, 0):
        W_next = r(layers[l + 1][0])
        d = r((d @ W_next) * dsilu(zs[l]))
        grads[l] = (d.T @ hs[l], d.sum(axis=0))
    return loss, flatten(grads)
