"""Two Landau levels (SphereTwoLevelPseudopotentialModel): the oracle's basis
and Hamiltonian pinned on the CPU by closed forms, Jordan-Wigner brute force,
sector block structure and reduction to the one-level build; the device path
(HD_SECTOR_LZ2 basis, hd_tables_create_two_level + hd_ham_build) against the
oracle bit-exactly."""
import sys
from math import comb

import numpy as np
import pytest

import hamiltonian_diagonalization_amd as hd
from hamiltonian_diagonalization_amd import models


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _combined(n1, first, second):
    return (first.astype(np.uint64) << np.uint64(n1 + 2)) | second.astype(np.uint64)


def _lz(n1, first, second):
    mx = n1 + 1
    return (sum(2 * i - (mx - 2) for i in range(n1) if first >> i & 1) +
            sum(2 * j - mx for j in range(n1 + 2) if second >> j & 1))


def _tables(oracle, n1, seed):
    """Level tables with random symmetric-in-pairs V, so that H is Hermitian:
    V(k1,k2,k3,k4) = W(k1,k2 | k4,k3) with W real symmetric under exchange of
    the created and annihilated pairs."""
    rng = np.random.default_rng(seed)
    out = []
    for level, L in ((0, n1), (1, n1 + 2)):
        kt = oracle.two_level_ktable(n1, level)
        W = rng.normal(size=(L, L, L, L))
        W = W + W.transpose(2, 3, 0, 1)  # <12|V|34> = <34|V|12>
        vt = np.zeros(L ** 3)
        for k4 in range(L):
            for k3 in range(L):
                for k2 in range(L):
                    k1 = int(kt[(k4 * L + k3) * L + k2])
                    if k1 < L:
                        vt[(k4 * L + k3) * L + k2] = W[k1, k2, k3, k4]
        out.append((kt, vt))
    eps1 = rng.normal(size=n1)
    eps2 = rng.normal(size=n1 + 2)
    return out[0], out[1], eps1, eps2


# ------------------------------------------------------------------ oracle pins (CPU)

@pytest.mark.parametrize("n1,N", [(2, 2), (3, 3), (4, 3), (5, 4)])
def test_basis_counts_and_sectors(oracle, n1, N):
    f, g = oracle.two_level_basis(N, n1)
    # GetFullFockSpaceDimension (two_level_fermion_fock_basis.cpp:140-155)
    assert len(f) == sum(comb(n1, k) * comb(n1 + 2, N - k) for k in range(N + 1))
    assert np.array_equal(np.sort(_combined(n1, f, g)), _combined(n1, f, g))
    assert len(set(_combined(n1, f, g).tolist())) == len(f)
    total = 0
    for lz in range(-N * (n1 + 1), N * (n1 + 1) + 1, 2):
        ff, gg = oracle.two_level_basis(N, n1, lz)
        total += len(ff)
        assert all(_lz(n1, a, b) == lz for a, b in zip(ff.tolist(), gg.tolist()))
        # the sector is symmetric under m -> -m: mirror every orbital in its level
        fm = [sum(1 << (n1 - 1 - i) for i in range(n1) if a >> i & 1) for a in ff.tolist()]
        gm = [sum(1 << (n1 + 1 - j) for j in range(n1 + 2) if b >> j & 1) for b in gg.tolist()]
        f2, g2 = oracle.two_level_basis(N, n1, -lz)
        assert sorted(zip(fm, gm)) == list(zip(f2.tolist(), g2.tolist()))
    assert total == len(f)


def _jw_dense(oracle, n1, tabs, eps, states_comb):
    """<out|H|in> from explicit Jordan-Wigner matrices over the 2 n1 + 2
    combined orbitals (level 0 at n1+2+k, level 1 at k)."""
    K2 = n1 + 2
    quartic = {}
    for level, (kt, vt) in enumerate(tabs):
        L = n1 if level == 0 else K2
        c = (lambda k: K2 + k) if level == 0 else (lambda k: k)
        for k4 in range(L):
            for k3 in range(L):
                for k2 in range(L):
                    k1 = int(kt[(k4 * L + k3) * L + k2])
                    if k1 < L:
                        key = (c(k1), c(k2), c(k3), c(k4))
                        quartic[key] = quartic.get(key, 0.0) + vt[(k4 * L + k3) * L + k2]
    e = None
    if eps is not None:
        e = np.concatenate([eps[1], eps[0]])
    return oracle.dense_hamiltonian_jw(2 * n1 + 2, states_comb, quartic, e)


@pytest.mark.parametrize("n1,N,lz,with_eps", [(2, 3, None, True), (3, 3, None, False), (3, 3, 1, True),
                                               (4, 2, 0, True)])
def test_oracle_matches_jordan_wigner(oracle, n1, N, lz, with_eps):
    (kt1, vt1), (kt2, vt2), e1, e2 = _tables(oracle, n1, 11 * n1 + N)
    f, g = oracle.two_level_basis(N, n1, lz)
    comb_states = _combined(n1, f, g)
    want = _jw_dense(oracle, n1, [(kt1, vt1), (kt2, vt2)], (e1, e2) if with_eps else None, comb_states)
    kw = dict(eps1=e1, eps2=e2) if with_eps else {}
    full = oracle.two_level_build(n1, f, g, kt1, vt1, kt2, vt2, full=True, **kw)
    up = oracle.two_level_build(n1, f, g, kt1, vt1, kt2, vt2, full=False, **kw)
    dim = len(f)
    assert np.allclose(full.to_dense(dim, mirror=False), want, atol=1e-12, rtol=1e-12)
    assert np.allclose(up.to_dense(dim, mirror=True), want, atol=1e-12, rtol=1e-12)
    # the stored upper triangle really is upper
    for i in range(dim):
        assert np.all(up.col[up.rowptr[i]:up.rowptr[i + 1]] >= i)


def test_oracle_sector_blocks(oracle):
    """The full-basis matrix is block diagonal in 2 L_z, and each sector build
    is bit-identical to its block."""
    n1, N = 3, 3
    (kt1, vt1), (kt2, vt2), e1, e2 = _tables(oracle, n1, 5)
    f, g = oracle.two_level_basis(N, n1)
    full = oracle.two_level_build(n1, f, g, kt1, vt1, kt2, vt2, e1, e2, full=True)
    lzs = np.array([_lz(n1, a, b) for a, b in zip(f.tolist(), g.tolist())])
    M = full.to_dense(len(f), mirror=False)
    assert np.all(M[lzs[:, None] != lzs[None, :]] == 0)
    for lz in sorted(set(lzs.tolist())):
        ff, gg = oracle.two_level_basis(N, n1, lz)
        sec = oracle.two_level_build(n1, ff, gg, kt1, vt1, kt2, vt2, e1, e2, full=True)
        idx = np.nonzero(lzs == lz)[0]
        assert np.array_equal(sec.to_dense(len(ff), mirror=False), M[np.ix_(idx, idx)])


def test_oracle_reduces_to_one_level(oracle):
    """Rows with the second level empty (first level empty) are the one-level
    sphere build of the level-0 (level-1) tables, value for value."""
    n1, N = 4, 3
    (kt1, vt1), (kt2, vt2), e1, e2 = _tables(oracle, n1, 9)
    f, g = oracle.two_level_basis(N, n1)
    full = oracle.two_level_build(n1, f, g, kt1, vt1, kt2, vt2, e1, e2, full=True)
    M = full.to_dense(len(f), mirror=False)
    for level, sel, L, kt, vt, e, states in ((0, g == 0, n1, kt1, vt1, e1, f), (1, f == 0, n1 + 2, kt2, vt2, e2, g)):
        idx = np.nonzero(sel)[0]
        bas = np.ascontiguousarray(states[idx])
        one = oracle.build(N, L, bas, oracle.LAYOUT_SPHERE, L, kt, vt, eps=e, full=True)
        assert np.array_equal(one.to_dense(len(idx), mirror=False), M[np.ix_(idx, idx)]), level


def test_oracle_mid_trim(oracle):
    """Add_CdCdCC_Terms for level 0 ends with Trim: a diagonal that cancels to
    below 1e-14 after level 0 restarts from 0 before the level-1 terms, so it
    ends bit-identical to one whose E terms cancel exactly."""
    n1, N = 3, 3
    (kt1, vt1), (kt2, vt2), _, _ = _tables(oracle, n1, 3)
    f, g = oracle.two_level_basis(N, n1)
    i = next(r for r in range(len(f)) if bin(int(f[r])).count("1") == 1 and bin(int(g[r])).count("1") == 2)
    k = int(f[i]).bit_length() - 1
    j1 = (int(g[i]) & -int(g[i])).bit_length() - 1

    def diag(e_j1):
        e1 = np.zeros(n1)
        e1[k] = 1.0
        e2 = np.zeros(n1 + 2)
        e2[j1] = e_j1
        a = oracle.two_level_build(n1, f, g, kt1, vt1, kt2, vt2, e1, e2, full=True)
        lo, hi = a.rowptr[i], a.rowptr[i + 1]
        return a.values()[lo:hi][a.col[lo:hi] == i][0]

    tiny, exact = diag(1e-15 - 1.0), diag(-1.0)
    assert 0 < abs(1.0 + (1e-15 - 1.0)) < 1e-14
    assert tiny == exact  # bitwise: the 1e-15 residue was trimmed before level 1
    assert (1.0 + (1e-15 - 1.0)) + exact != exact  # the case is sensitive to the trim


# ------------------------------------------------------------------ device parity

TL_CASES = [(3, 3, None), (3, 4, 0), (4, 4, 2), (5, 5, -1), (6, 6, 0), (8, 6, 0)]


@pytest.mark.gpu
@pytest.mark.parametrize("n1,N,lz", TL_CASES)
def test_gpu_two_level_basis(oracle, n1, N, lz):
    say("two-level basis n1", n1, "N", N, "lz", lz)
    f, g = oracle.two_level_basis(N, n1, lz)
    if len(f) == 0:
        with pytest.raises(hd.HDError):
            hd.FockBasis(hd.Sector.lz2(lz, n1) if lz is not None else hd.Sector.none(), N, 2 * n1 + 2)
        return
    b = hd.FockBasis(hd.Sector.lz2(lz, n1) if lz is not None else hd.Sector.none(), N, 2 * n1 + 2)
    assert b.dim == len(f)
    got = b.states()
    assert np.array_equal(got, _combined(n1, f, g))
    assert np.array_equal(b.find_index(got), np.arange(len(got), dtype=np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("n1,N,lz", TL_CASES)
def test_gpu_two_level_build_bit_exact(oracle, n1, N, lz):
    say("two-level build n1", n1, "N", N, "lz", lz)
    (kt1, vt1), (kt2, vt2), e1, e2 = _tables(oracle, n1, 100 + n1)
    f, g = oracle.two_level_basis(N, n1, lz)
    if len(f) == 0:
        pytest.skip("empty sector")
    b = hd.FockBasis(hd.Sector.lz2(lz, n1) if lz is not None else hd.Sector.none(), N, 2 * n1 + 2)
    for eps in ((e1, e2), (None, None)):
        t = hd.TermTables.two_level(n1, kt1, vt1, kt2, vt2, *eps)
        for storage, full in ((hd.STORAGE_UPPER_REF, False), (hd.STORAGE_FULL, True)):
            ref = oracle.two_level_build(n1, f, g, kt1, vt1, kt2, vt2, eps[0], eps[1], full=full)
            h = hd.Hamiltonian(b, t, storage=storage)
            rp, col, val = h.crs()
            assert np.array_equal(rp, ref.rowptr)
            assert np.array_equal(col, ref.col)
            assert np.array_equal(val, ref.values())
            assert h.info()["n_trimmed"] == ref.n_trimmed


@pytest.mark.gpu
def test_gpu_two_level_coulomb_lanczos(oracle):
    """Coulomb in both levels, N=6, n1=7 (LLL 2S=6, 2LL 2S=8), L_z=0: Lanczos
    against dense eigenvalues of the oracle's matrix."""
    n1, N = 7, 6
    say("two-level Coulomb n1", n1, "N", N)
    b, t = models.two_level_problem(N, n1, 0, eps1=np.zeros(n1), eps2=np.full(n1 + 2, 0.5))
    f, g = oracle.two_level_basis(N, n1, 0)
    kt1, kt2 = t.ktable
    vt1, vt2 = t.vtable
    ref = oracle.two_level_build(n1, f, g, kt1, vt1, kt2, vt2, np.zeros(n1), np.full(n1 + 2, 0.5), full=True)
    h = hd.Hamiltonian(b, t, values=(hd.F64,))
    rp, col, val = h.crs()
    assert np.array_equal(col, ref.col) and np.array_equal(val, ref.values())
    w = np.linalg.eigvalsh(ref.to_dense(len(f), mirror=False))
    r = h.lanczos(nev=4, tol=1e-12, seed=1)
    # with no coupling between the levels the spectrum holds sums E_LLL + E_2LL
    # with exact degeneracies, of which one start vector need not find every
    # copy: the lowest value and membership of every Ritz value are checked
    assert abs(r.evals[0] - w[0]) < 1e-9
    assert all(np.min(np.abs(w - e)) < 1e-9 for e in r.evals)
    # random level tables: no degeneracy, the four lowest in order
    (kt1, vt1), (kt2, vt2), e1, e2 = _tables(oracle, n1, 7)
    t2 = hd.TermTables.two_level(n1, kt1, vt1, kt2, vt2, e1, e2)
    ref2 = oracle.two_level_build(n1, f, g, kt1, vt1, kt2, vt2, e1, e2, full=True)
    w2 = np.linalg.eigvalsh(ref2.to_dense(len(f), mirror=False))
    assert np.min(np.diff(w2[:5])) > 1e-6
    h2 = hd.Hamiltonian(b, t2, values=(hd.F64,))
    r2 = h2.lanczos(nev=4, tol=1e-12, seed=1)
    assert np.allclose(r2.evals, w2[:4], atol=1e-9, rtol=0)


@pytest.mark.gpu
def test_gpu_two_level_mid_trim(oracle):
    say("two-level mid-build trim of the diagonal")
    n1, N = 3, 3
    (kt1, vt1), (kt2, vt2), _, _ = _tables(oracle, n1, 3)
    f, g = oracle.two_level_basis(N, n1)
    b = hd.FockBasis(hd.Sector.none(), N, 2 * n1 + 2)
    i = next(r for r in range(len(f)) if bin(int(f[r])).count("1") == 1 and bin(int(g[r])).count("1") == 2)
    e1 = np.zeros(n1)
    e1[int(f[i]).bit_length() - 1] = 1.0
    e2 = np.zeros(n1 + 2)
    e2[(int(g[i]) & -int(g[i])).bit_length() - 1] = 1e-15 - 1.0
    ref = oracle.two_level_build(n1, f, g, kt1, vt1, kt2, vt2, e1, e2, full=True)
    h = hd.Hamiltonian(b, hd.TermTables.two_level(n1, kt1, vt1, kt2, vt2, e1, e2), storage=hd.STORAGE_FULL)
    rp, col, val = h.crs()
    assert np.array_equal(rp, ref.rowptr) and np.array_equal(col, ref.col)
    assert np.array_equal(val, ref.values())
