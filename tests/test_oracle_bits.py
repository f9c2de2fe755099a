"""Oracle bit tools and basis against definitions (math.comb, itertools)."""
import itertools
import math

import numpy as np
import pytest


def test_binomial(oracle):
    L = oracle.lib()
    for n in range(0, 65):
        for k in range(0, n + 1):
            assert L.oc_binomial(n, k) == math.comb(n, k)


@pytest.mark.parametrize("norb,w", [(6, 3), (9, 4), (12, 5), (10, 1), (10, 10)])
def test_hamming_order_unrank_rank(oracle, norb, w):
    L = oracle.lib()
    ref = sorted(sum(1 << b for b in c) for c in itertools.combinations(range(norb), w))
    x = L.oc_generate_hamming(w, 0)
    for i, r in enumerate(ref):
        assert x == r
        assert L.oc_generate_hamming(w, i) == r
        assert L.oc_index_hamming(w, r) == i
        x = L.oc_next_hamming(x)


def _brute_sector(norb, n, pred):
    return np.array(sorted(s for s in (sum(1 << b for b in c)
                                       for c in itertools.combinations(range(norb), n))
                           if pred(s)), dtype=np.uint64)


def _bits(s):
    return [b for b in range(64) if (s >> b) & 1]


@pytest.mark.parametrize("norb,n,lz", [(10, 4, 0), (11, 4, 2), (9, 3, -4), (12, 5, 1), (8, 3, 0)])
def test_basis_lz(oracle, norb, n, lz):
    maxlz = norb - 1
    got = oracle.basis(oracle.sector(oracle.SECTOR_LZ, lz, maxlz), n, norb)
    want = _brute_sector(norb, n, lambda s: sum(2 * b - maxlz for b in _bits(s)) == lz)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("X,Y,n,kx,ky", [(3, 3, 3, 0, 0), (3, 4, 4, 1, 2), (2, 5, 3, 1, 4)])
def test_basis_k2d(oracle, X, Y, n, kx, ky):
    got = oracle.basis(oracle.sector(oracle.SECTOR_K2D, kx, ky, X, Y), n, X * Y)
    want = _brute_sector(X * Y, n, lambda s: (sum(b // Y for b in _bits(s)) % X == kx and
                                              sum(b % Y for b in _bits(s)) % Y == ky))
    assert np.array_equal(got, want)


def test_basis_k1d_none(oracle):
    got = oracle.basis(oracle.sector(oracle.SECTOR_K1D, 2, 4), 4, 12)
    want = _brute_sector(12, 4, lambda s: sum(b % 4 for b in _bits(s)) % 4 == 2)
    assert np.array_equal(got, want)
    got = oracle.basis(oracle.sector(oracle.SECTOR_NONE), 3, 9)
    assert np.array_equal(got, _brute_sector(9, 3, lambda s: True))


def _lz_dim(norb, n, lz):
    """Sector dimension by subset-sum DP over orbital indices (independent of
    any enumeration): states with sum of k = (lz + n*(norb-1))/2."""
    tot = lz + n * (norb - 1)
    if tot % 2:
        return 0
    target = tot // 2
    dp = [[0] * (target + 1) for _ in range(n + 1)]
    dp[0][0] = 1
    for k in range(norb):
        for c in range(n, 0, -1):
            for t in range(target, k - 1, -1):
                dp[c][t] += dp[c - 1][t - k]
    return dp[n][target]


def test_sector_dimensions(oracle):
    # configuration sizes used by the benchmarks
    assert _lz_dim(22, 8, 0) == 8512
    assert _lz_dim(34, 12, 0) == 7764392
    assert sum(_lz_dim(28, 10, lz) for lz in range(0, 10 * 18 + 1, 2)) == 6684779
    bas = oracle.basis(oracle.sector(oracle.SECTOR_LZ, 0, 21), 8, 22)
    assert len(bas) == 8512


def test_find_index(oracle):
    L = oracle.lib()
    bas = oracle.basis(oracle.sector(oracle.SECTOR_LZ, 0, 11), 5, 12)
    p = bas.ctypes.data_as(__import__("ctypes").c_void_p)
    for i in range(0, len(bas), 7):
        assert L.oc_find_index(p, len(bas), int(bas[i])) == i
    assert L.oc_find_index(p, len(bas), 1) == -1
