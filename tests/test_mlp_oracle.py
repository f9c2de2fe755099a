"""The NumPy network oracle against central finite differences (float64,
no rounding): pins the backward pass the device gradients are checked with."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def _random_net(rng, n_bits, n_params, n_states=1):
    layers = []
    for l, (r, c) in enumerate(mo.SHAPES):
        W = np.zeros((r, c))
        live_r = 2 * n_states if l == 3 else r
        live_c = n_bits + n_params if l == 0 else c
        W[:live_r, :live_c] = rng.normal(scale=np.sqrt(2.0 / live_c), size=(live_r, live_c))
        b = np.zeros(r)
        b[:live_r] = rng.normal(scale=0.1, size=live_r)
        layers.append((W, b))
    return mo.flatten(layers)


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.0e-5, 65504.0], dtype=np.float32)
    r = mo.bf16(x)
    # bf16 keeps 8 mantissa bits: 1 + 2^-8 is a tie and rounds to even (1.0)
    assert r[0] == 1.0 and r[1] == 1.0
    assert r[2] == np.float32(1.0078125)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)


@pytest.mark.parametrize("n_states", [1, 3])
def test_gradients_match_finite_differences(n_states):
    """One or several amplitudes (2 outputs each; targets n x 3 n_states)."""
    rng = np.random.default_rng(n_states - 1)
    n_bits, n_params = 10, 3
    flat = _random_net(rng, n_bits, n_params, n_states)
    n = 24
    states = np.array([int(rng.choice(1 << n_bits)) for _ in range(n)], dtype=np.uint64)
    params = rng.random((n, n_params))
    targets = np.concatenate([np.stack([rng.normal(size=n), rng.choice([-1.0, 1.0], size=n), rng.random(n) + 0.5], 1)
                              for _ in range(n_states)], axis=1)
    loss, g = mo.loss_and_grads(flat, states, params, targets, n_bits, n_params, lam=0.7, bf=False)
    layers = mo.unflatten(np.arange(flat.size))
    picks = []
    for l, (W, b) in enumerate(layers):
        live_r = 2 * n_states if l == 3 else W.shape[0]
        live_c = n_bits + n_params if l == 0 else W.shape[1]
        for _ in range(6):
            picks.append(int(W[rng.integers(live_r), rng.integers(live_c)]))
        picks.append(int(b[rng.integers(live_r)]))
    eps = 1e-6
    for i in picks:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += eps
        fm[i] -= eps
        lp, _ = mo.loss_and_grads(fp, states, params, targets, n_bits, n_params, lam=0.7, bf=False)
        lm, _ = mo.loss_and_grads(fm, states, params, targets, n_bits, n_params, lam=0.7, bf=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
