"""Eigenvector observables (struct Observables, fermion_observables.hpp): the
oracle's restatement pinned by closed forms and an independent operator
calculation on the CPU; the device kernels (hd_obs_*) against the oracle."""
import sys

import numpy as np
import pytest

import hamiltonian_diagonalization_amd as hd
from hamiltonian_diagonalization_amd import models
from oracle import observables as ob


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _sector_states(oracle, kind, a, b, c, d, n, norb):
    return oracle.basis(oracle.sector(kind, a, b, c, d), n, norb)


def _rand_psi(rng, dim, cplx):
    v = rng.normal(size=dim) + (1j * rng.normal(size=dim) if cplx else 0)
    return v / np.linalg.norm(v)


# ------------------------------------------------------------------ oracle pins (CPU)

def _apply_c(k, vec):
    """c_k on a {state: amplitude} vector, standard second quantisation with
    orbitals ordered by index (sign = (-1)^(occupied orbitals below k))."""
    out = {}
    for s, a in vec.items():
        if (s >> k) & 1:
            sign = -1 if bin(s & ((1 << k) - 1)).count("1") % 2 else 1
            t = s ^ (1 << k)
            out[t] = out.get(t, 0) + sign * a
    return out


def _apply_cdag(k, vec):
    out = {}
    for s, a in vec.items():
        if not (s >> k) & 1:
            sign = -1 if bin(s & ((1 << k) - 1)).count("1") % 2 else 1
            t = s | (1 << k)
            out[t] = out.get(t, 0) + sign * a
    return out


def test_translational_oracle_matches_operator_algebra(oracle):
    """sum_i p_i <i| sum_{k1,k2} c+_{k2-q} c_{k2} c+_{k1+q} c_{k1} |i> / N, the
    operators composed one by one."""
    X, Y, n = 2, 3, 3
    norb = X * Y
    add, sub, _ = models.lattice_tables(X, Y)
    states = _sector_states(oracle, oracle.SECTOR_K2D, 1, 2, X, Y, n, norb)
    psi = _rand_psi(np.random.default_rng(0), len(states), True)
    got = ob.translational_dd(states, psi, n, norb, add, sub)
    want = np.zeros(norb)
    for i, s in enumerate(states.tolist()):
        p = abs(psi[i]) ** 2
        for q in range(norb):
            acc = 0.0
            for k1 in range(norb):
                for k2 in range(norb):
                    v = _apply_c(k1, {s: 1.0})
                    v = _apply_cdag(int(add[k1, q]), v)
                    v = _apply_c(k2, v)
                    v = _apply_cdag(int(sub[k2, q]), v)
                    acc += v.get(s, 0.0)
            want[q] += p * acc
    assert np.allclose(got, want / n, rtol=1e-13, atol=1e-15)
    # closed forms: S(0) = N for a normalised vector; a single Fock state gives
    # #{k occupied, k+q empty} / N for q != 0
    assert abs(got[0] - n) < 1e-12
    i = len(states) - 1
    s = int(states[i])
    e = np.zeros(len(states))
    e[i] = 1.0
    one = ob.translational_dd(states, e, n, norb, add, sub)
    occ = [k for k in range(norb) if (s >> k) & 1]
    for q in range(1, norb):
        assert one[q] * n == sum(1 for k in occ if not (s >> int(add[k, q])) & 1)


def test_rotational_oracle_closed_form(oracle):
    X, Y, n = 3, 3, 3
    norb = X * Y
    _, _, inv = models.lattice_tables(X, Y)
    # InverseR60 applied six times is the identity; R(0) counts matches over alpha
    k = np.arange(norb)
    r = k.copy()
    for _ in range(6):
        r = inv[r, 1]
    assert np.array_equal(r, k) and np.array_equal(inv[:, 0], k)
    states = _sector_states(oracle, oracle.SECTOR_K2D, 0, 0, X, Y, n, norb)
    s = int(states[1])
    e = np.zeros(len(states))
    e[1] = 1.0
    got = ob.rotational_dd(states, e, n, inv)
    occ = [q for q in range(norb) if (s >> q) & 1]
    C = [sum(1 for k1 in occ for k2 in occ if int(inv[k2, a]) == k1) for a in range(6)]
    assert C[0] == n
    want = [sum(C[a] * np.exp(1j * np.pi / 3 * m * a) for a in range(6)) / (6 * n) for m in range(6)]
    assert np.allclose(got, want, rtol=1e-14, atol=1e-15)


def test_density_oracle_invariants(oracle):
    n, norb = 4, 11
    states = _sector_states(oracle, oracle.SECTOR_LZ, 0, 10, 0, 0, n, norb)
    psi = _rand_psi(np.random.default_rng(1), len(states), False)
    d = ob.density(states, psi, norb)
    dd = ob.density_density(states, psi, norb)
    assert abs(d.sum() - n) < 1e-12
    assert np.allclose(np.diag(dd), d, rtol=1e-14) and np.allclose(dd, dd.T, rtol=1e-14)
    assert np.allclose(dd.sum(axis=1), n * d, rtol=1e-12)
    # the L_z sector is symmetric under m -> -m (orbital k <-> norb-1-k) only in
    # distribution; a single Fock state gives its occupations exactly
    e = np.zeros(len(states))
    e[5] = 1.0
    s = int(states[5])
    assert np.array_equal(ob.density(states, e, norb), [(s >> k) & 1 for k in range(norb)])
    assert ob.participation_ratio(np.full(len(states), 1 / np.sqrt(len(states)))) == pytest.approx(len(states))
    assert ob.participation_ratio(e) == 1.0


# ------------------------------------------------------------------ device parity

@pytest.mark.gpu
def test_gpu_observables_sphere_ground_state(oracle):
    say("observables sphere N=6 2S=15 lz=0: ground state and a random vector")
    b, t = models.sphere_problem(6, 15)
    h = hd.Hamiltonian(b, t, values=(hd.F64,))
    r = h.lanczos(nev=1, tol=1e-12, seed=0)
    psi0 = h.eigenvector(0)
    states = b.states()
    norb = 16
    for psi in (psi0, _rand_psi(np.random.default_rng(2), b.dim, False)):
        assert np.allclose(b.density(psi), ob.density(states, psi, norb), rtol=1e-12, atol=1e-15)
        assert np.allclose(b.density_density(psi), ob.density_density(states, psi, norb), rtol=1e-12, atol=1e-15)
        assert b.participation_ratio(psi) == pytest.approx(ob.participation_ratio(psi), rel=1e-12)
    # the L=0 ground state is rotationally invariant: uniform density N/Norb
    assert np.allclose(b.density(psi0), 6 / 16, atol=1e-8)


@pytest.mark.gpu
@pytest.mark.parametrize("cplx", [False, True])
def test_gpu_observables_lattice(oracle, cplx):
    say("observables flux lattice 3x4 N=4, complex", cplx)
    X, Y, n = 3, 4, 4
    norb = X * Y
    add, sub, inv = models.lattice_tables(X, Y)
    b = hd.FockBasis(hd.Sector.k2d(1, 2, X, Y), n, norb)
    states = b.states()
    psi = _rand_psi(np.random.default_rng(3), b.dim, cplx)
    assert np.allclose(b.density(psi), ob.density(states, psi, norb), rtol=1e-12, atol=1e-15)
    assert np.allclose(b.translational_dd(psi, add, sub), ob.translational_dd(states, psi, n, norb, add, sub),
                       rtol=1e-12, atol=1e-14)
    assert np.allclose(b.rotational_dd(psi, inv), ob.rotational_dd(states, psi, n, inv), rtol=1e-12, atol=1e-14)
    st, amp = b.most_probable(psi, 10)
    wst, wamp = ob.most_probable(states, psi, 10)
    assert np.array_equal(st, wst) and np.array_equal(amp, wamp)


@pytest.mark.gpu
def test_gpu_most_probable_ties_and_large_sector():
    say("most probable: ties, n > distinct values, N=8 2S=21")
    b = hd.FockBasis(hd.Sector.lz(0, 21), 8, 22)
    v = np.zeros(b.dim)
    v[[7, 3, 100, 5000]] = [0.5, -0.5, 0.5, 0.1]
    st, amp = b.most_probable(v, 5)
    states = b.states()
    # ties at |0.5| in ascending index order, then 0.1, then the first zero
    assert np.array_equal(st, states[[3, 7, 100, 5000, 0]])
    assert np.array_equal(amp, [-0.5, 0.5, 0.5, 0.1, 0.0])


@pytest.mark.gpu
def test_gpu_dropin_observables(oracle, tmp_path):
    """The C++ drop-in's Observables through the driver, against the oracle on
    the ground state the driver wrote (raw --final-vector file)."""
    import json
    import os
    import subprocess
    say("drop-in observables: flux lattice 3x4 N=5 complex, sector (0,0)")
    driver = os.path.join(os.path.dirname(hd.__file__), "lib", "hd_diagonalize")
    X, Y, n = 3, 4, 5
    norb = X * Y
    kt, vt, eps = models.ofl_tables(X, Y, 11)
    kt.tofile(tmp_path / "k.bin")
    vt.astype(np.complex128).tofile(tmp_path / "v.bin")
    eps.astype(np.complex128).tofile(tmp_path / "e.bin")
    f = tmp_path / "gs.bin"
    p = subprocess.run([driver, "--nbr-particles", str(n), "--kmax", str(norb), "--complex", "1",
                        "--ktable", str(tmp_path / "k.bin"), "--vtable", str(tmp_path / "v.bin"),
                        "--eps", str(tmp_path / "e.bin"), "--sector", f"k2d:0,0,{X},{Y}", "--nbr-eigenvalues", "2",
                        "--tol", "1e-11", "--dense-below", "10", "--final-vector", str(f), "--observables", "1",
                        "--dim-x", str(X), "--dim-y", str(Y)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    r = json.loads(p.stdout.strip().splitlines()[-1])
    states = oracle.basis(oracle.sector(oracle.SECTOR_K2D, 0, 0, X, Y), n, norb)
    psi = np.fromfile(f, dtype=np.complex128)
    assert len(psi) == len(states) == r["dim"]
    add, sub, inv = models.lattice_tables(X, Y)
    assert np.allclose(r["density"], ob.density(states, psi, norb), rtol=1e-12, atol=1e-15)
    assert r["participation_ratio"] == pytest.approx(ob.participation_ratio(psi), rel=1e-12)
    assert np.allclose(r["translational"], ob.translational_dd(states, psi, n, norb, add, sub), rtol=1e-12, atol=1e-14)
    rot = np.array(r["rotational"][0::2]) + 1j * np.array(r["rotational"][1::2])
    assert np.allclose(rot, ob.rotational_dd(states, psi, n, inv), rtol=1e-12, atol=1e-14)
    wst, _ = ob.most_probable(states, psi, 5)
    assert r["most_probable"] == [int(x) for x in wst]
