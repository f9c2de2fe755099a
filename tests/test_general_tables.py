"""General (hash) term tables and the hybrid Wannier basis on the CPU: the
oracle's general build against Jordan-Wigner brute force, against the array
build on converted tables (bit for bit), and the Wannier transform against
numpy and against the invariant that fixes it (the Wannier Hamiltonian of a
k_y sector has the spectrum of the Bloch sectors (k_x, k_y) for all k_x)."""
import itertools

import numpy as np
import pytest

import hamiltonian_diagonalization_amd as hd
from hamiltonian_diagonalization_amd import models


def _terms_from_lists(K, lists):
    """(k1,k2,k3,k4) terms the reference's quartic loop can reach (k3 != k4)
    and (k1,k2) quadratic terms, in the list format of TermTables.hash."""
    quartic, quadratic = [], []
    nk1, k1, v = lists["nk1"], np.asarray(lists["k1"]), np.asarray(lists["v"])
    for k4, k3, k2 in itertools.product(range(K), repeat=3):
        if k3 == k4:
            continue
        e = (k4 * K + k3) * K + k2
        for j in range(nk1[e]):
            quartic.append(((int(k1[e, j]), k2, k3, k4), v[e, j]))
    if lists.get("nq1") is not None:
        for k2 in range(K):
            for j in range(lists["nq1"][k2]):
                quadratic.append(((int(lists["q1"][k2, j]), k2), lists["e"][k2, j]))
    return quartic, quadratic


@pytest.mark.parametrize("K,n,cplx,sector,seed", [
    (6, 3, True, None, 0), (6, 2, False, None, 1), (7, 3, True, None, 2),
    (7, 3, False, ("lz", 2), 3), (8, 4, True, ("k1d", 1, 2), 4)])
def test_general_oracle_matches_jordan_wigner(oracle, K, n, cplx, sector, seed):
    """Full-storage rows of the general build are <out|H|in> of the operator
    sum V c+c+cc + E c+c over the listed terms (repeats add), restricted to the
    sector: out states outside it are skipped, as FindFockStateIndex's miss."""
    lists = models.random_hash_tables(K, seed=seed, complex_values=cplx)
    if sector is None:
        sec = oracle.sector(oracle.SECTOR_NONE)
    elif sector[0] == "lz":
        sec = oracle.sector(oracle.SECTOR_LZ, sector[1], K - 1)
    else:
        sec = oracle.sector(oracle.SECTOR_K1D, sector[1], sector[2])
    bas = oracle.basis(sec, n, K)
    crs = oracle.build_general(K, bas, K, full=True, **lists)
    quartic, quadratic = _terms_from_lists(K, lists)
    H = oracle.dense_hamiltonian_jw_general(K, bas, quartic, quadratic)
    M = crs.to_dense(len(bas), mirror=False)
    assert np.abs(M - H.T).max() < 1e-12
    # the upper storage keeps the out >= in part of the same rows
    up = oracle.build_general(K, bas, K, full=False, **lists)
    U = up.to_dense(len(bas), mirror=False)
    assert np.abs(U - np.triu(H.T)).max() < 1e-12
    assert crs.nnz > 0


def order_sensitive_lists(oracle):
    """Tables under which one entry of the N=2, K=4 sector (row |0,1>, column
    |0,2>) is reached by three terms whose rounded sum depends on the order:
    the hop E(2,1) c+_2 c_1, then c+_2 c+_0 c_0 c_1 (k4=1, k3=0, k2=0) and
    c+_0 c+_2 c_0 c_1 (k4=1, k3=0, k2=2), with signed elements 1, 2^-53,
    2^-53.  In the reference's order the sum is (1 + 2^-53) + 2^-53 = 1;
    any order that adds the two small terms first gives 1 + 2^-52."""
    import ctypes
    K, L = 4, oracle.lib()
    hi = 1 << (K - 1)
    s = ctypes.c_int(0)
    state = 0b0011

    def sign_cdc(a, b):
        L.oc_cdc(ctypes.byref(s), 1 << a, 1 << b, state, hi)
        return s.value

    def sign_cdcdcc(a, b, c, d):
        L.oc_cdcdcc(ctypes.byref(s), 1 << a, 1 << b, 1 << c, 1 << d, state, hi)
        return s.value

    big, small = 1.0, 2.0 ** -53
    nk1 = np.zeros(K ** 3, dtype=np.uint8)
    k1 = np.zeros((K ** 3, 1), dtype=np.uint16)
    v = np.zeros((K ** 3, 1))
    for (a, b, c, d) in ((2, 0, 0, 1), (0, 2, 0, 1)):
        e = (d * K + c) * K + b
        nk1[e], k1[e, 0], v[e, 0] = 1, a, sign_cdcdcc(a, b, c, d) * small
    nq1 = np.zeros(K, dtype=np.uint8)
    q1 = np.zeros((K, 1), dtype=np.uint16)
    ee = np.zeros((K, 1))
    nq1[1], q1[1, 0], ee[1, 0] = 1, 2, sign_cdc(2, 1) * big
    return {"nk1": nk1, "k1": k1, "v": v, "nq1": nq1, "q1": q1, "e": ee}


def test_general_oracle_sums_duplicates_in_term_order(oracle):
    """A column reached by several terms holds their sum in the reference's
    loop order (quadratic terms first, then the quartic terms by k4, k3, k2)."""
    K, n = 4, 2
    lists = order_sensitive_lists(oracle)
    bas = oracle.basis(oracle.sector(oracle.SECTOR_NONE), n, K)
    crs = oracle.build_general(K, bas, K, full=True, **lists)
    row = int(np.searchsorted(bas, 0b0011))
    col = int(np.searchsorted(bas, 0b0101))
    lo, hi = crs.rowptr[row], crs.rowptr[row + 1]
    got = dict(zip(crs.col[lo:hi].tolist(), crs.values()[lo:hi].tolist()))
    assert got[col] == 1.0
    assert (1.0 + (2.0 ** -53 + 2.0 ** -53)) != 1.0


@pytest.mark.parametrize("layout", ["ofl", "sphere"])
def test_hash_conversion_reproduces_array_build_bits(oracle, layout):
    """ConvertTableFormat: the general build on the hash lists of array tables
    is the array build, bit for bit (same terms in the same order)."""
    if layout == "ofl":
        X, Y, n = 3, 3, 4
        kt, vt, eps = models.ofl_tables(X, Y, seed=5)
        K = X * Y
        sec = oracle.sector(oracle.SECTOR_K2D, 1, 2, X, Y)
        lists = hd.hash_lists_from_array(hd.LAYOUT_OFL, K, kt, vt, eps)
        lay, cplx = oracle.LAYOUT_OFL, True
    else:
        K, n = 13, 4
        kt = oracle.sphere_ktable(K - 1)
        vt = oracle.sphere_vtable(K - 1, [0.0, 1.0, 0.0, 0.4])
        eps = None
        sec = oracle.sector(oracle.SECTOR_LZ, 0, K - 1)
        lists = hd.hash_lists_from_array(hd.LAYOUT_SPHERE, K, kt, vt)
        lay, cplx = oracle.LAYOUT_SPHERE, False
    bas = oracle.basis(sec, n, K)
    for full in (False, True):
        ref = oracle.build(n, K, bas, lay, K, kt, vt, eps=eps, is_complex=cplx, full=full)
        gen = oracle.build_general(K, bas, K, full=full, **lists)
        assert np.array_equal(ref.rowptr, gen.rowptr)
        assert np.array_equal(ref.col, gen.col)
        assert np.array_equal(ref.val.view(np.uint64), gen.val.view(np.uint64))
        assert ref.n_trimmed == gen.n_trimmed


def _unfolded_v(X, Y, kt, vt):
    """Full V[q1,q2,q3,q4] from the folded array (GetK1 / GetVkkkk)."""
    K = X * Y
    V = np.zeros((K, K, K, K), dtype=np.complex128)
    for q4, q3, q2 in itertools.product(range(K), repeat=3):
        t1, t2 = min(q3, q4), max(q3, q4)
        row = (t1 * K + t2 - t1 * (t1 + 1) // 2) * K
        q1 = int(kt[row + q2])
        V[q1, q2, q3, q4] = vt[row + (q1 if q4 > q3 else q2)]
    return V


@pytest.mark.parametrize("X,Y,seed", [(3, 2, 0), (2, 3, 1), (4, 2, 2)])
def test_wannier_oracle_matches_tensor_contraction(oracle, X, Y, seed):
    """The oracle's quartic transform is V' = sum V conj(U U) U U (numpy
    einsum over the unfolded table), within rounding."""
    K = X * Y
    kt, vt, _ = models.ofl_tables(X, Y, seed=seed)
    rng = np.random.default_rng(seed)
    U = models.wannier_matrix(X, Y, 0.25, np.exp(2j * np.pi * rng.random((X, Y))))
    k1, v = oracle.ofl_wannier_quartic(X, Y, kt, vt, U, 0.0)
    V = _unfolded_v(X, Y, kt, vt)
    W = np.einsum("abcd,ia,jb,kc,ld->ijkl", V, U.conj(), U.conj(), U, U)
    slots = np.arange(K ** 3 * X)
    e, x1 = slots // X, slots % X
    k2, k3, k4 = e % K, (e // K) % K, e // (K * K)
    ky1 = (k3 % Y + k4 % Y - k2 % Y) % Y
    assert np.array_equal(k1, (x1 * Y + ky1).astype(np.uint16))
    ref = W[k1.astype(int), k2, k3, k4]
    assert np.abs(v - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())
    # entries that do not conserve k_y vanish in the contraction
    mask = np.ones(W.shape, dtype=bool)
    mask[k1.astype(int), k2, k3, k4] = False
    assert np.abs(W[mask]).max() < 1e-12


def wannier_lists_cpu(oracle, X, Y, offset, kt, vt, eps, U, tol):
    """The Wannier hash lists from the oracle's quartic transform and the
    quadratic formula of ChangeToWannierBasis (numpy's inverse FFT for the
    backward DFT)."""
    K = X * Y
    k1, v = oracle.ofl_wannier_quartic(X, Y, kt, vt, U, tol)
    E = np.asarray(eps, dtype=np.complex128).reshape(X, Y)
    out = np.fft.ifft(E, axis=0)  # out[j, ky] = sum_kx E/X exp(+2 pi i kx j / X)
    nq1 = np.full(K, X, dtype=np.uint8)
    q1 = np.zeros((K, X), dtype=np.uint16)
    e = np.zeros((K, X), dtype=np.complex128)
    for x2, ky, x1 in itertools.product(range(X), range(Y), range(X)):
        d = abs(x2 - x1)
        q1[x2 * Y + ky, x1] = x1 * Y + ky
        e[x2 * Y + ky, x1] = np.exp(2j * np.pi * d * offset / X) * out[d, ky]
    return {"nk1": np.full(K ** 3, X, dtype=np.uint8), "k1": k1.reshape(-1, X), "v": v.reshape(-1, X),
            "nq1": nq1, "q1": q1, "e": e}


@pytest.mark.parametrize("X,Y,n,ky,seed", [(3, 2, 3, 0, 0), (3, 2, 3, 1, 1), (4, 2, 3, 1, 2)])
def test_wannier_spectrum_is_union_of_bloch_sectors(oracle, X, Y, n, ky, seed):
    """A unitary change of single-particle basis within each k_y keeps the
    many-body spectrum: the Wannier Hamiltonian of the k_y sector has the
    eigenvalues of the Bloch sectors (k_x, k_y), k_x = 0..X-1, together.  The
    reference's quadratic formula takes |x2 - x1| for the offset phase and
    the DFT index, exact when offset_x = 0 and E(k_x) = E(-k_x); the test
    draws such a band."""
    K = X * Y
    kt, vt, _ = models.ofl_tables(X, Y, seed=seed)
    rng = np.random.default_rng(100 + seed)
    E = rng.uniform(-0.3, 0.3, size=(X, Y))
    E = (E + E[(-np.arange(X)) % X]) / 2
    eps = E.reshape(K).astype(np.complex128)
    U = models.wannier_matrix(X, Y, 0.0, np.exp(2j * np.pi * rng.random((X, Y))))
    bloch = []
    for kx in range(X):
        bas = oracle.basis(oracle.sector(oracle.SECTOR_K2D, kx, ky, X, Y), n, K)
        if len(bas) == 0:
            continue
        crs = oracle.build(n, K, bas, oracle.LAYOUT_OFL, K, kt, vt, eps=eps, is_complex=True, full=True)
        bloch.append(np.linalg.eigvalsh(crs.to_dense(len(bas), mirror=False)))
    bloch = np.sort(np.concatenate(bloch))
    lists = wannier_lists_cpu(oracle, X, Y, 0.0, kt, vt, eps, U, 1e-14)
    bas = oracle.basis(oracle.sector(oracle.SECTOR_K1D, ky, Y), n, K)
    crs = oracle.build_general(K, bas, K, full=True, **lists)
    M = crs.to_dense(len(bas), mirror=False)
    assert np.abs(M - M.conj().T).max() < 1e-12
    w = np.linalg.eigvalsh(M)
    assert len(w) == len(bloch)
    assert np.abs(w - bloch).max() < 1e-11


def test_mirror_upper_is_the_reference_product(oracle):
    """oracle.mirror_upper lays out U + (U - D)^H: its rows applied to x give
    the reference's symmetric product on the upper CRS, for a non-Hermitian
    table too."""
    K, n = 7, 3
    lists = models.random_hash_tables(K, seed=8, complex_values=True)
    bas = oracle.basis(oracle.sector(oracle.SECTOR_NONE), n, K)
    up = oracle.build_general(K, bas, K, full=False, **lists)
    full = oracle.mirror_upper(up)
    assert np.all(np.diff(full.rowptr) >= 0)
    for r in range(len(bas)):
        c = full.col[full.rowptr[r]:full.rowptr[r + 1]]
        assert np.all(np.diff(c) > 0)
    rng = np.random.default_rng(1)
    x = rng.normal(size=len(bas)) + 1j * rng.normal(size=len(bas))
    M = full.to_dense(len(bas), mirror=False)
    assert np.abs(M @ x - oracle.spmv_sym(up, x)).max() < 1e-12
    assert np.abs(M - up.to_dense(len(bas), mirror=True)).max() == 0.0
