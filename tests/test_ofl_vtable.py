"""Flux-lattice quartic terms from Bloch coefficients
(InteractingOflModel::GenerateQuarticTerms, interacting_ofl_model.cpp:457-590):
the oracle's convolution restatement pinned on the CPU by a direct
cross-correlation, the defining sum over reciprocal lattice vectors, and closed
forms; the device kernel (hd_ofl_vtable) against the oracle bit for bit."""
import sys

import numpy as np
import pytest

import hamiltonian_diagonalization_amd as hd


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _bloch(rng, K, cut):
    """Random normalised band-0 blocks: [K][(2c+1)^2][2 spins]."""
    n = (2 * cut + 1) ** 2
    b = rng.normal(size=(K, n, 2)) + 1j * rng.normal(size=(K, n, 2))
    return b / np.linalg.norm(b.reshape(K, -1), axis=1)[:, None, None]


def _entries(K):
    for k4 in range(K):
        for k3 in range(k4, K):
            for k2 in range(K):
                yield k4, k3, k2


# ------------------------------------------------------------------ oracle pins (CPU)

@pytest.mark.parametrize("X,Y", [(3, 3), (5, 5), (3, 5)])
def test_conv2d_is_a_cross_correlation(oracle, X, Y):
    """out[i][j] = sum_p left[p] right[p + (X-1-i, Y-1-j)] over the overlap;
    the offset arguments zero the border rows and columns."""
    rng = np.random.default_rng(X * 10 + Y)
    L = rng.normal(size=X * Y) + 1j * rng.normal(size=X * Y)
    R = rng.normal(size=X * Y) + 1j * rng.normal(size=X * Y)
    got = oracle.conv2d(R, L, X, Y)
    L2, R2 = L.reshape(X, Y), R.reshape(X, Y)
    want = np.zeros((2 * X - 1, 2 * Y - 1), dtype=complex)
    for i in range(2 * X - 1):
        for j in range(2 * Y - 1):
            dx, dy = X - 1 - i, Y - 1 - j
            for px in range(X):
                for py in range(Y):
                    if 0 <= px + dx < X and 0 <= py + dy < Y:
                        want[i, j] += L2[px, py] * R2[px + dx, py + dy]
    assert np.allclose(got, want, rtol=1e-14, atol=1e-14)
    z = oracle.conv2d(R, L, X, Y, 1, 2, 1, 0)
    mask = np.zeros_like(want, dtype=bool)
    mask[1:2 * X - 2, 2:] = True
    assert np.array_equal(z[mask], got[mask]) and np.all(z[~mask] == 0)


def _brute_v(bloch, dim_x, dim_y, cut, k1, k2, k3, k4, gx, gy):
    """sum_sigma sum_{GA+GB = GC+GD-g} conj(c1_GA,s) conj(c2_GB,-s) c3_GC,-s c4_GD,s."""
    X = 2 * cut + 1
    G = [(a - cut, b - cut) for a in range(X) for b in range(X)]
    idx = {g: i for i, g in enumerate(G)}
    v = 0j
    for s in (0, 1):
        A, B, C, D = np.conj(bloch[k1, :, s]), np.conj(bloch[k2, :, 1 - s]), bloch[k3, :, 1 - s], bloch[k4, :, s]
        for ia, ga in enumerate(G):
            for ib, gb in enumerate(G):
                for ic, gc in enumerate(G):
                    gd = (ga[0] + gb[0] - gc[0] + gx, ga[1] + gb[1] - gc[1] + gy)
                    if gd in idx:
                        v += A[ia] * B[ib] * C[ic] * D[idx[gd]]
    return v


@pytest.mark.parametrize("dim_x,dim_y,cut", [(3, 3, 1), (2, 3, 2)])
def test_oracle_matches_reciprocal_lattice_sum(oracle, dim_x, dim_y, cut):
    K = dim_x * dim_y
    rng = np.random.default_rng(dim_x + 7 * cut)
    b = _bloch(rng, K, cut)
    got = oracle.ofl_vtable(dim_x, dim_y, cut, cut, b)
    assert len(got) == K * K * (K + 1) // 2
    picks = list(_entries(K))
    sel = rng.choice(len(picks), size=40, replace=False)
    for e in sel:
        k4, k3, k2 = picks[e]
        k1, gx, gy = oracle.ofl_find_k1(dim_x, dim_y, k2, k3, k4)
        assert (k1 // dim_y - (k3 // dim_y + k4 // dim_y - k2 // dim_y)) == gx * dim_x
        assert (k1 % dim_y - (k3 % dim_y + k4 % dim_y - k2 % dim_y)) == gy * dim_y
        want = _brute_v(b, dim_x, dim_y, cut, k1, k2, k3, k4, gx, gy)
        assert abs(got[e] - want) <= 1e-13 * max(1.0, abs(want)), (k4, k3, k2)


def test_oracle_plane_wave_closed_form(oracle):
    """Bloch vectors with only G = 0: V = sum_s conj(a1_s a2_-s) a3_-s a4_s when
    k1 + k2 = k3 + k4 without a reciprocal lattice vector, 0 otherwise."""
    dim_x, dim_y, cut = 3, 4, 2
    K = dim_x * dim_y
    rng = np.random.default_rng(5)
    a = rng.normal(size=(K, 2)) + 1j * rng.normal(size=(K, 2))
    b = np.zeros((K, (2 * cut + 1) ** 2, 2), dtype=complex)
    b[:, (2 * cut + 1) ** 2 // 2, :] = a
    got = oracle.ofl_vtable(dim_x, dim_y, cut, cut, b)
    nz = 0
    for e, (k4, k3, k2) in enumerate(_entries(K)):
        k1, gx, gy = oracle.ofl_find_k1(dim_x, dim_y, k2, k3, k4)
        if gx or gy:
            assert got[e] == 0
        else:
            nz += 1
            want = sum(np.conj(a[k1, s] * a[k2, 1 - s]) * a[k3, 1 - s] * a[k4, s] for s in (0, 1))
            assert abs(got[e] - want) <= 1e-15 * max(1, abs(want))
    assert nz > 0


# ------------------------------------------------------------------ device parity

@pytest.mark.gpu
@pytest.mark.parametrize("dim_x,dim_y,cut", [(3, 3, 1), (4, 3, 2), (6, 6, 3), (5, 4, 5), (3, 3, 12), (6, 5, 7)])
def test_gpu_ofl_vtable_bit_exact(oracle, dim_x, dim_y, cut):
    say("ofl vtable", dim_x, dim_y, "cut", cut)
    K = dim_x * dim_y
    b = _bloch(np.random.default_rng(K + cut), K, cut)
    n = K * K * (K + 1) // 2
    if n * (2 * cut + 1) ** 4 > 3e8:
        e0, e1 = n // 3, n // 3 + 200  # a bounded oracle sample of a large table
        want = oracle.ofl_vtable(dim_x, dim_y, cut, cut, b, e0, e1)
    else:
        e0, e1 = 0, n
        want = oracle.ofl_vtable(dim_x, dim_y, cut, cut, b)
    got = hd.ofl_vtable(dim_x, dim_y, cut, b)
    assert got.shape == (n,)
    assert np.array_equal(got[e0:e1].view(np.float64), want.view(np.float64))


@pytest.mark.gpu
def test_gpu_ofl_vtable_builds_a_hamiltonian(oracle):
    """The generated table feeds the flux-lattice build: the pair structure is
    accepted and the CRS matches the oracle's bit for bit."""
    dim_x, dim_y, cut, n = 3, 4, 2, 4
    K = dim_x * dim_y
    say("ofl vtable -> build", dim_x, dim_y, "N", n)
    b = _bloch(np.random.default_rng(3), K, cut)
    vt = hd.ofl_vtable(dim_x, dim_y, cut, b) * 0.5  # SetCoefficient(0.5 * strength), strength 1
    kt = hd.ofl_ktable(dim_x, dim_y)
    eps = np.linspace(-1, 1, K).astype(np.complex128)
    basis = hd.FockBasis(hd.Sector.k2d(1, 2, dim_x, dim_y), n, K)
    t = hd.TermTables(hd.LAYOUT_OFL, K, kt, vt, eps)
    ref = oracle.build(n, K, basis.states(), oracle.LAYOUT_OFL, K, kt, vt, eps=eps, is_complex=True, full=True)
    h = hd.Hamiltonian(basis, t, storage=hd.STORAGE_FULL)
    rp, col, val = h.crs()
    assert np.array_equal(rp, ref.rowptr) and np.array_equal(col, ref.col)
    assert np.array_equal(val, ref.values())


@pytest.mark.gpu
@pytest.mark.parametrize("cut", [5, 18])
def test_gpu_ofl_vtable_flux_lattice_6x5_timing(oracle, cut):
    """The 6x5 flux-lattice grid (13,950 entries) at the reference's smallest
    and largest default Bloch cutoffs (5 and 18, interacting_ofl_model_options.hpp:62-66):
    device time for the whole table, the oracle's single-core time on 20
    entries, and those 20 entries bit for bit."""
    import time
    dim_x, dim_y = 6, 5
    K = dim_x * dim_y
    b = _bloch(np.random.default_rng(cut), K, cut)
    hd.ofl_vtable(dim_x, dim_y, cut, b)  # warm-up
    v, ms = hd.ofl_vtable(dim_x, dim_y, cut, b, timing=True)
    e0 = 6000
    t0 = time.perf_counter()
    want = oracle.ofl_vtable(dim_x, dim_y, cut, cut, b, e0, e0 + 20)
    cpu = (time.perf_counter() - t0) / 20 * len(v)
    say(f"ofl vtable 6x5 cut {cut}: {len(v)} entries, device {ms:.2f} ms, "
        f"oracle 1 core {cpu:.1f} s for the table (from 20 entries), ratio {cpu * 1e3 / ms:.0f}")
    assert np.array_equal(v[e0:e0 + 20].view(np.float64), want.view(np.float64))
