"""The general (hash-table) device build and the hybrid Wannier transform
against the oracle: bit-exact CRS (row pointers, columns, FP64 value bits,
trimmed count) for synthetic general tables, for the hash lists of array
tables (which must also equal the pair-structure build), for the Wannier
tables of the reference's -x 5 -y 5 --nbr 5 --basis 1 run shape, and the
Wannier spectrum against the Bloch sectors it unites."""
import sys

import numpy as np
import pytest

import hamiltonian_diagonalization_amd as hd
from hamiltonian_diagonalization_amd import models

from test_general_tables import order_sensitive_lists, wannier_lists_cpu

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _sector(kind):
    if kind[0] == "none":
        return hd.Sector.none(), None
    if kind[0] == "lz":
        return hd.Sector.lz(kind[1], kind[2]), ("lz", kind[1], kind[2])
    return hd.Sector.k1d(kind[1], kind[2]), ("k1d", kind[1], kind[2])


def _oracle_sector(oracle, kind):
    if kind[0] == "none":
        return oracle.sector(oracle.SECTOR_NONE)
    if kind[0] == "lz":
        return oracle.sector(oracle.SECTOR_LZ, kind[1], kind[2])
    return oracle.sector(oracle.SECTOR_K1D, kind[1], kind[2])


def _same_bits(ham, ref, f32=True):
    rp, col, val = ham.crs()
    assert np.array_equal(rp, ref.rowptr)
    assert np.array_equal(col, ref.col)
    assert np.array_equal(val.view(np.uint64), ref.val.view(np.uint64))
    assert ham.info()["n_trimmed"] == ref.n_trimmed
    if f32:
        _, _, v32 = ham.crs(precision=hd.F32)
        ref32 = ref.values().astype(np.complex64 if ref.is_complex else np.float32)
        assert np.array_equal(v32.view(np.uint32), ref32.view(np.uint32))


CASES = [
    # K, N, complex, sector, max_k1, fill, seed
    (8, 4, True, ("none",), 3, 0.5, 0),
    (10, 5, False, ("lz", 1, 9), 3, 0.5, 1),
    (12, 4, True, ("k1d", 1, 3), 2, 0.6, 2),
    (18, 6, False, ("none",), 4, 0.7, 3),    # rows of > 1024 terms: the 4096-slot merge
    (16, 6, False, ("none",), 16, 0.9, 5),   # rows of > 4096 terms: the 8192-slot merge
    (20, 4, True, ("k1d", 0, 4), 3, 0.5, 4),
    (20, 8, False, ("k1d", 0, 10), 16, 0.9, 6),  # > 8192 terms per full row: the 16384-slot merge
]


@pytest.mark.parametrize("K,n,cplx,kind,max_k1,fill,seed", CASES)
@pytest.mark.parametrize("storage", [hd.STORAGE_FULL, hd.STORAGE_UPPER_REF])
def test_general_build_bit_exact(oracle, K, n, cplx, kind, max_k1, fill, seed, storage):
    say("general build", K, n, cplx, kind, max_k1, fill, seed, "storage", storage)
    lists = models.random_hash_tables(K, seed=seed, complex_values=cplx, max_k1=max_k1, fill=fill)
    sec, _ = _sector(kind)
    basis = hd.FockBasis(sec, n, K)
    tables = hd.TermTables.hash(K, **lists)
    ham = hd.Hamiltonian(basis, tables, storage=storage)
    bas = oracle.basis(_oracle_sector(oracle, kind), n, K)
    assert np.array_equal(basis.states(), bas)
    up = oracle.build_general(K, bas, K, full=False, **lists)
    ref = oracle.mirror_upper(up) if storage == hd.STORAGE_FULL else up
    _same_bits(ham, ref, f32=storage == hd.STORAGE_FULL)
    assert ham.info()["max_row"] == int(np.diff(ref.rowptr).max())
    if storage == hd.STORAGE_FULL:
        # the product is the reference's on its upper CRS, U x + (U - D)^H x
        rng = np.random.default_rng(seed)
        x = rng.normal(size=len(bas)) + (1j * rng.normal(size=len(bas)) if cplx else 0)
        y = ham.spmv(x, precision=hd.F64)
        want = oracle.spmv_sym(up, x)
        assert np.abs(y - want).max() <= 1e-12 * max(1.0, np.abs(want).max())


def test_general_build_sums_duplicates_in_term_order(oracle):
    say("general build: order-sensitive duplicate sum")
    lists = order_sensitive_lists(oracle)
    basis = hd.FockBasis(hd.Sector.none(), 2, 4)
    ham = hd.Hamiltonian(basis, hd.TermTables.hash(4, **lists))
    rp, col, val = ham.crs()
    states = basis.states()
    row = int(np.searchsorted(states, 0b0011))
    c = int(np.searchsorted(states, 0b0101))
    got = dict(zip(col[rp[row]:rp[row + 1]].tolist(), val[rp[row]:rp[row + 1]].tolist()))
    assert got[c] == 1.0


def test_general_build_spmv_and_lanczos_match_dense(oracle):
    """The general operator through the product and the solver: y = M x
    against the dense matrix, lowest eigenvalues against numpy."""
    K, n = 10, 4
    lists = models.random_hash_tables(K, seed=7, complex_values=True, max_k1=2)
    # Hermitise: every listed term plus its conjugate-transposed term
    slots = {}
    for e in range(K ** 3):
        k2, k3, k4 = e % K, (e // K) % K, e // (K * K)
        for j in range(lists["nk1"][e]):
            k1, v = int(lists["k1"][e, j]), lists["v"][e, j]
            for (a, b, c, d, w) in ((k1, k2, k3, k4, v), (k4, k3, k2, k1, np.conj(v))):
                slots.setdefault((d * K + c) * K + b, []).append((a, w))
    width = max(len(x) for x in slots.values())
    herm = {"nk1": np.zeros(K ** 3, dtype=np.uint8), "k1": np.zeros((K ** 3, width), dtype=np.uint16),
            "v": np.zeros((K ** 3, width), dtype=np.complex128)}
    for e, x in slots.items():
        herm["nk1"][e] = len(x)
        for j, (a, w) in enumerate(x):
            herm["k1"][e, j], herm["v"][e, j] = a, w
    basis = hd.FockBasis(hd.Sector.none(), n, K)
    ham = hd.Hamiltonian(basis, hd.TermTables.hash(K, **herm))
    say("general spmv/lanczos dim", basis.dim)
    bas = basis.states()
    M = oracle.build_general(K, bas, K, full=False, **herm).to_dense(len(bas), mirror=True)
    assert np.abs(M - M.conj().T).max() < 1e-12
    rng = np.random.default_rng(0)
    x = rng.normal(size=basis.dim) + 1j * rng.normal(size=basis.dim)
    y = ham.spmv(x, precision=hd.F64)
    assert np.abs(y - M @ x).max() < 1e-11 * np.abs(M @ x).max()
    r = ham.lanczos(nev=4, tol=1e-10)
    w = np.linalg.eigvalsh(M)[:4]
    assert np.abs(np.asarray(r.evals) - w).max() < 1e-9


@pytest.mark.parametrize("case", ["ofl", "sphere"])
def test_hash_lists_of_array_tables_equal_pair_build(oracle, case):
    """ConvertTableFormat (--use-hash 1): the general build on the hash lists
    of array tables equals the pair-structure build bit for bit, and the
    oracle's array build."""
    if case == "ofl":
        X, Y, n = 4, 3, 6
        kt, vt, eps = models.ofl_tables(X, Y, seed=0)
        K = X * Y
        basis = hd.FockBasis(hd.Sector.k2d(1, 2, X, Y), n, K)
        arr = hd.TermTables(hd.LAYOUT_OFL, K, kt, vt, eps)
        lists = hd.hash_lists_from_array(hd.LAYOUT_OFL, K, kt, vt, eps)
        lay, cplx = oracle.LAYOUT_OFL, True
    else:
        n, two_s = 8, 21
        K = two_s + 1
        kt = hd.sphere_ktable(two_s)
        vt = hd.sphere_vtable(two_s, models.coulomb_list(two_s)[:6])
        eps = None
        basis = hd.FockBasis(hd.Sector.lz(0, two_s), n, K)
        arr = hd.TermTables(hd.LAYOUT_SPHERE, K, kt, vt)
        lists = hd.hash_lists_from_array(hd.LAYOUT_SPHERE, K, kt, vt)
        lay, cplx = oracle.LAYOUT_SPHERE, False
    say("hash lists of", case, "tables: dim", basis.dim)
    gen = hd.Hamiltonian(basis, hd.TermTables.hash(K, **lists), storage=hd.STORAGE_UPPER_REF)
    pair = hd.Hamiltonian(basis, arr, storage=hd.STORAGE_UPPER_REF)
    a, b = gen.crs(), pair.crs()
    for u, v in zip(a, b):
        assert np.array_equal(u.view(np.uint8), v.view(np.uint8))
    bas = basis.states()
    up = oracle.build(n, K, bas, lay, K, kt, vt, eps=eps, is_complex=cplx, full=False)
    _same_bits(gen, up, f32=False)
    # full storage: the Hermitian completion of that upper triangle, and the
    # same spectrum as the pair build's full storage
    genf = hd.Hamiltonian(basis, hd.TermTables.hash(K, **lists))
    _same_bits(genf, oracle.mirror_upper(up))
    pairf = hd.Hamiltonian(basis, arr, values=(hd.F64,))
    e1 = genf.lanczos(nev=4, tol=1e-10).evals
    e2 = pairf.lanczos(nev=4, tol=1e-10).evals
    assert np.abs(np.asarray(e1) - np.asarray(e2)).max() < 1e-9


@pytest.mark.parametrize("X,Y,offset", [(3, 2, 0.0), (4, 3, 0.37), (5, 5, 0.0)])
def test_wannier_transform_bit_exact(oracle, X, Y, offset):
    """hd_ofl_wannier: the quartic lists are the oracle's bit for bit; the
    quadratic lists (the DFT the reference takes from FFTW) agree with numpy's
    inverse FFT to a few ulp."""
    say("wannier transform", X, Y, offset)
    K = X * Y
    kt, vt, eps = models.ofl_tables(X, Y, seed=3)
    rng = np.random.default_rng(9)
    U = models.wannier_matrix(X, Y, offset, np.exp(2j * np.pi * rng.random((X, Y))))
    dev, ms = hd.ofl_wannier(X, Y, offset, kt, vt, eps, U, 1e-14, timing=True)
    say("wannier kernel ms", ms)
    cpu = wannier_lists_cpu(oracle, X, Y, offset, kt, vt, eps, U, 1e-14)
    assert np.array_equal(dev["nk1"], cpu["nk1"])
    assert np.array_equal(dev["k1"], cpu["k1"])
    assert np.array_equal(dev["v"].view(np.uint64), cpu["v"].view(np.uint64))
    assert np.array_equal(dev["nq1"], cpu["nq1"])
    assert np.array_equal(dev["q1"], cpu["q1"])
    scale = np.abs(eps).max()
    assert np.abs(dev["e"] - cpu["e"]).max() <= 8 * np.finfo(float).eps * scale


def test_wannier_basis_run_5x5(oracle):
    """The reference's test run -x 5 -y 5 --nbr 5 --basis 1 (run_test.sh:19),
    here on synthetic tables: the Wannier build of the k_y = 0 sector is the
    oracle's bit for bit, and its lowest eigenvalues are the lowest of the
    five Bloch sectors (k_x, 0) together (offset 0 and a band with
    E(k_x) = E(-k_x), where the reference's quadratic formula is exact)."""
    X = Y = 5
    n, K = 5, 25
    kt, vt, _ = models.ofl_tables(X, Y, seed=11)
    rng = np.random.default_rng(12)
    E = rng.uniform(-0.3, 0.3, size=(X, Y))
    E = (E + E[(-np.arange(X)) % X]) / 2
    eps = E.reshape(K).astype(np.complex128)
    U = models.wannier_matrix(X, Y, 0.0, np.exp(2j * np.pi * rng.random((X, Y))))
    lists = hd.ofl_wannier(X, Y, 0.0, kt, vt, eps, U, 1e-14)
    basis = hd.FockBasis(hd.Sector.k1d(0, Y), n, K)
    say("wannier 5x5 N=5 ky=0 dim", basis.dim)
    ham = hd.Hamiltonian(basis, hd.TermTables.hash(K, **lists))
    info = ham.info()
    say("general build ms", {k: round(info[k], 2) for k in ("t_count_ms", "t_scan_ms", "t_fill_ms", "t_sort_ms")},
        "nnz", info["nnz"], "max_row", info["max_row"])
    bas = basis.states()
    ref = oracle.mirror_upper(oracle.build_general(K, bas, K, full=False, **lists))
    _same_bits(ham, ref)
    nev = 4
    w = ham.lanczos(nev=nev, tol=1e-10, ncv=40).evals
    bloch = []
    for kx in range(X):
        b = hd.FockBasis(hd.Sector.k2d(kx, 0, X, Y), n, K)
        h = hd.Hamiltonian(b, hd.TermTables(hd.LAYOUT_OFL, K, kt, vt, eps), values=(hd.F64,))
        bloch.extend(h.lanczos(nev=nev, tol=1e-10, ncv=40).evals)
    bloch = np.sort(bloch)[:nev]
    say("wannier evals", np.asarray(w), "bloch", bloch)
    assert np.abs(np.asarray(w) - bloch).max() < 1e-8


@pytest.mark.parametrize("case", ["sphere", "ofl"])
def test_state_list_basis(oracle, case):
    """hd_basis_create_from_states: a basis no sector descriptor reproduces
    (the L_z or momentum sector without the states occupying two orbitals);
    ranking by binary search, the upper CRS bit-exact against the oracle's
    build over that list (out states outside it skipped), full storage its
    Hermitian completion, the product and the spectrum against the oracle."""
    if case == "sphere":
        n, two_s = 7, 17
        K = two_s + 1
        full = oracle.basis(oracle.sector(oracle.SECTOR_LZ, 1, two_s), n, K)
        kt, vt, eps, lay = hd.sphere_ktable(two_s), hd.sphere_vtable(two_s, [0.0, 1.0, 0.0, 0.3]), None, hd.LAYOUT_SPHERE
        olay, cplx = oracle.LAYOUT_SPHERE, False
    else:
        X, Y, n = 4, 3, 5
        K = X * Y
        full = oracle.basis(oracle.sector(oracle.SECTOR_K2D, 1, 1, X, Y), n, K)
        kt, vt, eps = models.ofl_tables(X, Y, seed=2)
        lay, olay, cplx = hd.LAYOUT_OFL, oracle.LAYOUT_OFL, True
    mask = (1 << 2) | (1 << 5)
    bas = full[(full & mask) == 0]
    say("state-list basis", case, "dim", len(bas), "of", len(full))
    basis = hd.FockBasis.from_states(bas, n, K)
    assert basis.dim == len(bas) and np.array_equal(basis.states(), bas)
    probe = np.concatenate([bas[::5], full[(full & mask) != 0][:7]])
    idx = basis.find_index(probe)
    assert np.array_equal(idx[:len(bas[::5])], np.arange(0, len(bas), 5))
    assert np.all(idx[len(bas[::5]):] == np.iinfo(np.uint64).max)
    tables = hd.TermTables(lay, K, kt, vt, eps)
    up = oracle.build(n, K, bas, olay, K, kt, vt, eps=eps, is_complex=cplx, full=False)
    hu = hd.Hamiltonian(basis, tables, storage=hd.STORAGE_UPPER_REF)
    _same_bits(hu, up, f32=False)
    hf = hd.Hamiltonian(basis, tables)
    _same_bits(hf, oracle.mirror_upper(up))
    rng = np.random.default_rng(3)
    x = rng.normal(size=len(bas)) + (1j * rng.normal(size=len(bas)) if cplx else 0)
    want = oracle.spmv_sym(up, x)
    assert np.abs(hf.spmv(x) - want).max() <= 1e-12 * np.abs(want).max()
    w = np.linalg.eigvalsh(up.to_dense(len(bas)))[:4]
    assert np.abs(np.asarray(hf.lanczos(nev=4, tol=1e-10).evals) - w).max() < 1e-9
