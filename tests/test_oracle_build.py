"""Oracle fermion signs and the Hamiltonian build against Jordan-Wigner
brute force (dense operator matrices built from Kronecker products)."""
import itertools

import numpy as np
import pytest

TOL = 1e-14


def _bits(s):
    return [b for b in range(64) if (s >> b) & 1]


def test_cdcdcc_cdc_vs_jw(oracle):
    L = oracle.lib()
    import ctypes
    norb = 6
    ops = oracle.jw_operators(norb)
    sign = ctypes.c_int(0)
    hi = 1 << (norb - 1)
    for k1, k2, k3, k4 in itertools.product(range(norb), repeat=4):
        op = ops[k1].T @ ops[k2].T @ ops[k3] @ ops[k4]
        for s in range(1 << norb):
            out = L.oc_cdcdcc(ctypes.byref(sign), 1 << k1, 1 << k2, 1 << k3, 1 << k4, s, hi)
            col = op[:, s]
            nz = np.nonzero(col)[0]
            if sign.value == 0:
                assert len(nz) == 0
            else:
                assert list(nz) == [out] and col[out] == sign.value
    for k1, k2 in itertools.product(range(norb), repeat=2):
        op = ops[k1].T @ ops[k2]
        for s in range(1 << norb):
            out = L.oc_cdc(ctypes.byref(sign), 1 << k1, 1 << k2, s, hi)
            nz = np.nonzero(op[:, s])[0]
            if sign.value == 0:
                assert len(nz) == 0
            else:
                assert list(nz) == [out] and op[out, s] == sign.value


def _sphere_terms(K, ktab, vtab):
    """All (k1,k2,k3,k4) terms the reference loops can reach: k3 != k4,
    k1 = kTable(k2,k3,k4) < kMax, k1 != k2, |V| > tol."""
    terms = {}
    for k4, k3, k2 in itertools.product(range(K), repeat=3):
        if k3 == k4:
            continue
        i = (k4 * K + k3) * K + k2
        k1 = int(ktab[i])
        if k1 >= K or k1 == k2 or abs(vtab[i]) <= TOL:
            continue
        terms[(k1, k2, k3, k4)] = terms.get((k1, k2, k3, k4), 0.0) + vtab[i]
    return terms


def _ofl_terms(X, Y, ktab, vtab):
    K = X * Y
    terms = {}
    for k4, k3, k2 in itertools.product(range(K), repeat=3):
        if k3 == k4:
            continue
        t1, t2, t3 = (k4, k3, k2)
        if k4 > k3:
            t1, t2 = k3, k4
        k1 = int(ktab[(t1 * K + t2 - t1 * (t1 + 1) // 2) * K + k2])
        if k4 > k3:
            t3 = k1
        v = vtab[(t1 * K + t2 - t1 * (t1 + 1) // 2) * K + t3]
        if k1 == k2 or abs(v) <= TOL:
            continue
        terms[(k1, k2, k3, k4)] = v
    return terms


def random_sphere_table(maxlz, rng, symmetric):
    K = maxlz + 1
    ktab = oracle_ktab(maxlz)
    if symmetric:
        # V_{1234} = V_{4321} real (Hermitian) and antisymmetric pairs
        vt = np.zeros(K ** 3)
        base = {}
        for k4, k3, k2 in itertools.product(range(K), repeat=3):
            k1 = int(ktab[(k4 * K + k3) * K + k2])
            if k1 >= K:
                continue
            key = tuple(sorted([(k1, k2, k3, k4), (k4, k3, k2, k1)]))
            if key not in base:
                base[key] = rng.normal()
            vt[(k4 * K + k3) * K + k2] = base[key]
        return ktab, vt
    return ktab, rng.normal(size=K ** 3)


_KT = {}


def oracle_ktab(maxlz):
    from oracle import oracle as o
    if maxlz not in _KT:
        _KT[maxlz] = o.sphere_ktable(maxlz)
    return _KT[maxlz]


def test_sphere_ktable_definition(oracle):
    maxlz = 7
    K = maxlz + 1
    kt = oracle.sphere_ktable(maxlz)
    for k4, k3, k2 in itertools.product(range(K), repeat=3):
        k1 = k3 + k4 - k2  # m = maxLz - 2k, so m-conservation is k-conservation
        want = k1 if 0 <= k1 < K else 65535
        assert kt[(k4 * K + k3) * K + k2] == want


@pytest.mark.parametrize("norb,n,lz,seed", [(8, 3, 1, 0), (9, 4, 0, 1), (10, 3, 3, 2)])
def test_build_full_nonsymmetric_vs_jw(oracle, norb, n, lz, seed):
    """full=1 keeps every generated element: the stored matrix must equal
    <out|H|in> (row = in) for an arbitrary, non-Hermitian V table."""
    rng = np.random.default_rng(seed)
    maxlz = norb - 1
    ktab, vtab = random_sphere_table(maxlz, rng, symmetric=False)
    bas = oracle.basis(oracle.sector(oracle.SECTOR_LZ, lz, maxlz), n, norb)
    crs = oracle.build(n, norb, bas, oracle.LAYOUT_SPHERE, norb, ktab, vtab, full=True)
    got = crs.to_dense(len(bas), mirror=False)
    H = oracle.dense_hamiltonian_jw(norb, bas, _sphere_terms(norb, ktab, vtab))
    np.testing.assert_allclose(got, H.T, rtol=1e-13, atol=1e-13)
    # columns ascending within each row
    for i in range(len(bas)):
        c = crs.col[crs.rowptr[i]:crs.rowptr[i + 1]]
        assert np.all(np.diff(c) > 0)


@pytest.mark.parametrize("norb,n,lz,seed", [(8, 3, 1, 3), (9, 4, 0, 4)])
def test_build_upper_symmetric_vs_jw(oracle, norb, n, lz, seed):
    rng = np.random.default_rng(seed)
    maxlz = norb - 1
    ktab, vtab = random_sphere_table(maxlz, rng, symmetric=True)
    bas = oracle.basis(oracle.sector(oracle.SECTOR_LZ, lz, maxlz), n, norb)
    up = oracle.build(n, norb, bas, oracle.LAYOUT_SPHERE, norb, ktab, vtab, full=False)
    for i in range(len(bas)):
        c = up.col[up.rowptr[i]:up.rowptr[i + 1]]
        assert np.all(c >= i) and np.all(np.diff(c) > 0)
    H = oracle.dense_hamiltonian_jw(norb, bas, _sphere_terms(norb, ktab, vtab))
    assert np.allclose(H, H.T)
    np.testing.assert_allclose(up.to_dense(len(bas)), H.T, rtol=1e-13, atol=1e-13)
    # SpMV (naive U + (U-D)^H) against the dense product
    x = rng.normal(size=len(bas))
    np.testing.assert_allclose(oracle.spmv_sym(up, x), H.T @ x, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(oracle.spmv_sym(up, x, threads=3), H.T @ x, rtol=1e-12,
                               atol=1e-12)


def ofl_table(X, Y, rng):
    """Momentum-conserving complex V (Gaussian, antisymmetrised in (k1,k2) and
    (k3,k4), Hermitised) folded as the flux-lattice tables store it."""
    K = X * Y
    kx = lambda k: k // Y
    ky = lambda k: k % Y
    k1of = lambda k2, k3, k4: ((kx(k3) + kx(k4) - kx(k2)) % X) * Y + (ky(k3) + ky(k4) - ky(k2)) % Y
    G = {}
    for k2, k3, k4 in itertools.product(range(K), repeat=3):
        G[(k1of(k2, k3, k4), k2, k3, k4)] = rng.normal() + 1j * rng.normal()
    A = {t: (G[t] - G[(t[1], t[0], t[2], t[3])] - G[(t[0], t[1], t[3], t[2])]
             + G[(t[1], t[0], t[3], t[2])]) / 4 for t in G}
    V = {t: (A[t] + np.conj(A[(t[3], t[2], t[1], t[0])])) / 2 for t in A}
    n = (K + 1) * K * K // 2
    kt = np.zeros(n, dtype=np.uint16)
    vt = np.zeros(n, dtype=np.complex128)
    for t1 in range(K):
        for t2 in range(t1, K):
            for t3 in range(K):
                i = (t1 * K + t2 - t1 * (t1 + 1) // 2) * K + t3
                k1 = k1of(t3, t2, t1)
                kt[i] = k1
                vt[i] = V[(k1, t3, t2, t1)]
    return kt, vt, V


@pytest.mark.parametrize("X,Y,n,kx,ky,seed", [(3, 3, 3, 0, 0, 5), (2, 4, 3, 1, 2, 6)])
def test_build_ofl_complex_vs_jw(oracle, X, Y, n, kx, ky, seed):
    rng = np.random.default_rng(seed)
    K = X * Y
    kt, vt, V = ofl_table(X, Y, rng)
    eps = rng.uniform(-0.1, 0.1, size=K).astype(np.complex128)
    bas = oracle.basis(oracle.sector(oracle.SECTOR_K2D, kx, ky, X, Y), n, K)
    terms = _ofl_terms(X, Y, kt, vt)
    # the fold reproduces every V entry
    for t, v in terms.items():
        assert abs(v - V[t]) < 1e-15
    H = oracle.dense_hamiltonian_jw(K, bas, terms, eps=eps.real)
    assert np.allclose(H, H.conj().T)
    up = oracle.build(n, K, bas, oracle.LAYOUT_OFL, K, kt, vt, eps=eps, is_complex=True)
    np.testing.assert_allclose(up.to_dense(len(bas)), H.T, rtol=1e-13, atol=1e-13)
    full = oracle.build(n, K, bas, oracle.LAYOUT_OFL, K, kt, vt, eps=eps, is_complex=True,
                        full=True)
    np.testing.assert_allclose(full.to_dense(len(bas), mirror=False), H.T, rtol=1e-13,
                               atol=1e-13)
    x = rng.normal(size=len(bas)) + 1j * rng.normal(size=len(bas))
    np.testing.assert_allclose(oracle.spmv_sym(up, x), H.T @ x, rtol=1e-12, atol=1e-12)


def test_laughlin_zero_mode(oracle):
    """V1 only at 2S = 3(N-1): the Laughlin state is the unique zero mode in
    L_z = 0 (and no zero modes elsewhere)."""
    n = 4
    maxlz = 3 * (n - 1)
    norb = maxlz + 1
    vt = oracle.sphere_vtable(maxlz, [0.0, 1.0])
    kt = oracle.sphere_ktable(maxlz)
    ev = {}
    for lz in (0, 2, 4):
        bas = oracle.basis(oracle.sector(oracle.SECTOR_LZ, lz, maxlz), n, norb)
        up = oracle.build(n, norb, bas, oracle.LAYOUT_SPHERE, norb, kt, vt)
        ev[lz] = np.linalg.eigvalsh(up.to_dense(len(bas)))
    assert abs(ev[0][0]) < 1e-12 and ev[0][1] > 1e-3
    assert ev[2][0] > 1e-3 and ev[4][0] > 1e-3
    # L^2 multiplets: every lz=2 eigenvalue appears in the lz=0 spectrum
    for e in ev[2]:
        assert np.min(np.abs(ev[0] - e)) < 1e-10
    # trace identity: the diagonal of H is sum over occupied pairs of
    # (V_{abba} - V_{abab}) ... checked against the JW dense H instead
    terms = _sphere_terms(norb, kt, vt)
    bas = oracle.basis(oracle.sector(oracle.SECTOR_LZ, 0, maxlz), n, norb)
    H = oracle.dense_hamiltonian_jw(norb, bas, terms)
    up = oracle.build(n, norb, bas, oracle.LAYOUT_SPHERE, norb, kt, vt)
    np.testing.assert_allclose(up.to_dense(len(bas)), H.T, rtol=1e-13, atol=1e-13)
