"""GPU: the BF16 MFMA GEMM, the network's forward and backward passes against
the NumPy oracle, training, and the learned start vector."""
import sys

import numpy as np
import pytest

import hamiltonian_diagonalization_amd as hd
from hamiltonian_diagonalization_amd import learned, models
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _bf16_bits(x):
    return mo.bf16(x).view(np.uint32) >> 16


@pytest.mark.parametrize("M,N,K", [(64, 64, 16), (128, 192, 64), (64, 256, 4096)])
def test_mfma_gemm(M, N, K):
    say("mfma gemm", M, N, K)
    rng = np.random.default_rng(M + N + K)
    A = mo.bf16(rng.normal(size=(M, K)).astype(np.float32))
    B = mo.bf16(rng.normal(size=(N, K)).astype(np.float32))
    C = learned.gemm_check(_bf16_bits(A).astype(np.uint16), _bf16_bits(B).astype(np.uint16))
    want = A.astype(np.float64) @ B.astype(np.float64).T
    assert np.max(np.abs(C - want)) <= 1e-5 * np.sqrt(K) * np.max(np.abs(want))


def _data(rng, n, n_bits):
    states = np.array([int(rng.integers(0, 1 << n_bits)) for _ in range(n)], dtype=np.uint64)
    params = rng.random((n, 3)).astype(np.float32)
    targets = np.stack([rng.normal(size=n), rng.choice([-1.0, 1.0], size=n),
                        rng.random(n) + 0.5], 1).astype(np.float32)
    return states, params, targets


@pytest.mark.parametrize("n_states", [1, 3])
def test_forward_and_gradients_match_oracle(n_states):
    say("mlp forward/backward vs oracle, amplitudes", n_states)
    rng = np.random.default_rng(1)
    n_bits = 34
    net = learned.AmplitudeNet(n_bits, 3, seed=5, n_states=n_states)
    w = net.get_weights()
    states, params, targets = _data(rng, 300, n_bits)  # 300 pads to 320 rows
    targets = np.concatenate([targets] + [_data(rng, 300, n_bits)[2] for _ in range(n_states - 1)], axis=1)
    out = net.forward(states, params)
    assert out.shape == (300, 2 * n_states)
    want, _, _ = mo.forward(w, states, params, n_bits, 3, bf=True)
    live = want[:, :2 * n_states]
    assert np.max(np.abs(out - live)) <= 2e-2 * max(1.0, np.max(np.abs(live)))
    assert np.all(want[:, 2 * n_states:] == 0)
    loss, g = net.gradients(states, params, targets, lam=0.5)
    lref, gref = mo.loss_and_grads(w, states, params, targets, n_bits, 3, lam=0.5, bf=True)
    assert abs(loss - lref) <= 1e-3 * abs(lref)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    say("gradient relative error", rel)
    assert rel < 2e-2
    # per layer too (no layer may hide behind a larger one)
    for (gw, gb), (rw, rb) in zip(learned.unflatten(g), learned.unflatten(gref.astype(np.float32))):
        assert np.linalg.norm(gw - rw) <= 5e-2 * np.linalg.norm(rw) + 1e-6
        assert np.linalg.norm(gb - rb) <= 5e-2 * np.linalg.norm(rb) + 1e-6


def test_weight_roundtrip_and_padding():
    net = learned.AmplitudeNet(22, 3, seed=1)
    w = net.get_weights()
    layers = learned.unflatten(w)
    assert np.all(layers[0][0][:, 25:] == 0) and np.all(layers[3][0][2:] == 0)
    w2 = w.copy()
    w2[:10] += 0.5
    net.set_weights(w2)
    assert np.array_equal(net.get_weights(), w2)


def test_training_and_learned_start():
    """Train on plain-solver ground states of a small sphere at a few
    parameter points; the network's start vector must overlap the held-out
    ground state better than before training, and Lanczos from it must reach
    the same eigenvalues."""
    say("mlp training on N=6 2S=15 ground states")
    n, two_s = 6, 15
    basis = hd.FockBasis(hd.Sector.lz(0, two_s), n, two_s + 1)
    rng = np.random.default_rng(3)
    pts = learned.sample_params(rng, 8)
    data = learned.make_dataset(basis, two_s, pts, per_point=2048, seed=0, log=None)
    net = learned.AmplitudeNet(two_s + 1, 3, seed=0)
    held = learned.sample_params(np.random.default_rng(99), 1)[0]
    h = learned.sphere_hamiltonian(basis, two_s, held)
    ref = h.lanczos(nev=3, tol=1e-10, seed=0)
    psi = h.eigenvector(0)
    ov0 = learned.overlap(net, basis, held, psi)
    losses = learned.train(net, data, steps=600, batch=1024, lr=2e-3, log=None)
    say("loss", losses[0], "->", losses[-1])
    assert losses[-1] < 0.3 * losses[0]
    ov1 = learned.overlap(net, basis, held, psi)
    say("overlap untrained", ov0, "trained", ov1)
    assert ov1 > max(0.9, ov0)
    start = net.start_vector(basis, held)
    r = h.lanczos(nev=3, tol=1e-10, start=start)
    assert r.converged
    assert np.max(np.abs(r.evals - ref.evals)) < 1e-9
    # the same start vector as the reference's --use-initial file, read by the
    # C++ drop-in driver: same ground state, fewer products than a random start
    import json
    import os
    import subprocess
    import tempfile
    from hamiltonian_diagonalization_amd import fileio
    driver = os.path.join(os.path.dirname(hd.__file__), "lib", "hd_diagonalize")
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "initial_vector.bin")
        v = net.write_initial_vector(basis, held, f)
        assert np.array_equal(fileio.read_initial_vector(f), v) and len(v) == basis.dim
        vm = models.coulomb_list(two_s, {1: float(held[0]), 3: float(held[1]), 5: float(held[2])})
        base = [driver, "--nbr-particles", str(n), "--nbr-orbitals", str(two_s + 1), "--lz", "0",
                "--pseudopotentials", ",".join("%.17g" % x for x in vm), "--nbr-eigenvalues", "1", "--tol", "1e-10"]
        out = [json.loads(subprocess.run(base + extra, capture_output=True, text=True, check=True,
                                         timeout=300).stdout.strip().splitlines()[-1])
               for extra in ([], ["--initial-vector", f])]
    say("drop-in matvecs random start", out[0]["matvecs"], "learned start", out[1]["matvecs"])
    assert abs(out[1]["eigenvalues"][0] - ref.evals[0]) < 1e-9
    assert out[1]["matvecs"] < out[0]["matvecs"]



def test_multi_amplitude_start_vector():
    """A network with 3 amplitudes trained on the 3 lowest eigenvectors of a
    small sphere: each amplitude overlaps its eigenvector well, the start
    vector (the sum of the normalised amplitudes) overlaps all three, and
    Lanczos from it reaches the random start's eigenvalues."""
    say("mlp with 3 amplitudes on N=6 2S=15")
    n, two_s = 6, 15
    basis = hd.FockBasis(hd.Sector.lz(0, two_s), n, two_s + 1)
    pts = learned.sample_params(np.random.default_rng(3), 8)
    data = learned.make_dataset(basis, two_s, pts, per_point=2048, seed=0, log=None, n_states=3)
    assert data.targets.shape[1] == 9
    net = learned.AmplitudeNet(two_s + 1, 3, seed=0, n_states=3)
    losses = learned.train(net, data, steps=800, batch=1024, lr=2e-3, log=None)
    say("loss", losses[0], "->", losses[-1])
    held = learned.sample_params(np.random.default_rng(99), 1)[0]
    h = learned.sphere_hamiltonian(basis, two_s, held)
    ref = h.lanczos(nev=3, tol=1e-10, seed=0)
    vecs = [h.eigenvector(k) for k in range(3)]
    amp = learned.amplitude_overlaps(net, basis, held, vecs)
    v = net.start_vector(basis, held).to_array(np.float64, basis.dim)
    v /= np.linalg.norm(v)
    over = [abs(v @ x) for x in vecs]
    say("amplitude overlaps", amp, "start vector overlaps", over)
    assert min(amp) > 0.5 and min(over) > 0.2
    # the forward outputs assemble the same start vector
    out = net.forward(basis.states(), np.repeat(np.asarray(held, dtype=np.float32)[None, :], basis.dim, axis=0))
    want = np.zeros(basis.dim)
    for s in range(3):
        a = np.exp((out[:, 2 * s] - out[:, 2 * s].max()).astype(np.float64)) * np.where(out[:, 2 * s + 1] >= 0, 1.0, -1.0)
        want += a / np.linalg.norm(a)
    assert np.max(np.abs(v - want / np.linalg.norm(want))) < 1e-12
    r = h.lanczos(nev=3, tol=1e-10, start=net.start_vector(basis, held))
    assert r.converged and np.max(np.abs(r.evals - ref.evals)) < 1e-9
