"""The host eigensolver of the projected matrices (hd_sym_eig: Householder +
implicit QL) against numpy's LAPACK eigvalsh, including degenerate spectra,
diagonal input and negligible off-diagonals.  CPU only."""
import numpy as np
import pytest

import hamiltonian_diagonalization_amd as hd


@pytest.mark.parametrize("n", [1, 2, 3, 8, 31, 64, 128, 160])
def test_random_symmetric(n):
    rng = np.random.default_rng(n)
    a = rng.normal(size=(n, n))
    a = a + a.T
    w, z = hd.sym_eig(a)
    assert np.all(np.diff(w) >= 0)
    assert np.max(np.abs(w - np.linalg.eigvalsh(a))) <= 1e-13 * max(1.0, np.max(np.abs(w)))
    assert np.linalg.norm(a @ z - z * w) <= 1e-12 * np.linalg.norm(a)
    assert np.linalg.norm(z.T @ z - np.eye(n)) <= 1e-12 * n


def test_degenerate_and_special():
    rng = np.random.default_rng(7)
    q, _ = np.linalg.qr(rng.normal(size=(40, 40)))
    d = np.repeat([1.0, 2.0, 3.0, 4.0], 10)
    a = q @ np.diag(d) @ q.T
    a = (a + a.T) / 2
    w, z = hd.sym_eig(a)
    assert np.max(np.abs(w - np.sort(d))) < 1e-13
    assert np.linalg.norm(a @ z - z * w) < 1e-12
    w, _ = hd.sym_eig(np.diag([3.0, 1.0, 1.0, 2.0, 0.0, 0.0]))
    assert np.array_equal(w, [0.0, 0.0, 1.0, 1.0, 2.0, 3.0])
    t = 2 * np.eye(9) + np.diag(np.full(8, 1e-310), 1) + np.diag(np.full(8, 1e-310), -1)
    w, _ = hd.sym_eig(t)
    assert np.allclose(w, 2.0)
    # a Lanczos-like arrow + tridiagonal matrix
    n, k = 31, 18
    a = np.zeros((n, n))
    a[np.arange(k), np.arange(k)] = np.sort(rng.normal(size=k))
    a[:k, k] = a[k, :k] = rng.normal(size=k) * 1e-3
    for j in range(k, n):
        a[j, j] = rng.normal()
        if j + 1 < n:
            a[j, j + 1] = a[j + 1, j] = abs(rng.normal())
    w, z = hd.sym_eig(a)
    assert np.max(np.abs(w - np.linalg.eigvalsh(a))) < 1e-13
