"""GPU parity: basis, CRS build, SpMV and Lanczos through the C ABI against the
CPU oracle (bit-exact for integer data and FP64 matrix values)."""
import sys

import numpy as np
import pytest

import hamiltonian_diagonalization_amd as hd
from hamiltonian_diagonalization_amd import models

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


SECTORS = [
    ("lz N=8 2S=21 lz=0", hd.SECTOR_LZ, (0, 21), 8, 22),
    ("lz N=4 2S=9 lz=2", hd.SECTOR_LZ, (2, 9), 4, 10),
    ("lz N=5 2S=11 lz=-3", hd.SECTOR_LZ, (-3, 11), 5, 12),
    ("k2d 3x4 N=4 (1,2)", hd.SECTOR_K2D, (1, 2, 3, 4), 4, 12),
    ("k2d 6x5 N=10 (0,0)", hd.SECTOR_K2D, (0, 0, 6, 5), 10, 30),
    ("k1d 4 N=4 ky=2", hd.SECTOR_K1D, (2, 4), 4, 12),
    ("none N=3 Norb=9", hd.SECTOR_NONE, (), 3, 9),
]


@pytest.mark.parametrize("name,kind,args,n,norb", SECTORS, ids=[s[0] for s in SECTORS])
def test_basis_matches_oracle(oracle, name, kind, args, n, norb):
    say("basis", name)
    want = oracle.basis(oracle.sector(kind, *args), n, norb)
    b = hd.FockBasis(hd.Sector(kind, *args), n, norb)
    assert b.dim == len(want)
    got = b.states()
    assert np.array_equal(got, want)
    idx = b.find_index(got, offset=7)
    assert np.array_equal(idx, np.arange(len(got), dtype=np.uint64) + 7)
    # states outside the sector are not found
    bad = np.array([0, (1 << n) - 1 if kind == hd.SECTOR_NONE else 1, 1 << 63], dtype=np.uint64)
    res = b.find_index(bad)
    for s, r in zip(bad, res):
        inside = len(want) and np.searchsorted(want, s) < len(want) and want[np.searchsorted(want, s)] == s
        assert (r != np.uint64(2**64 - 1)) == bool(inside)


def test_basis_config2(oracle):
    say("basis N=12 2S=33 lz=0 (7,764,392 states)")
    b = hd.FockBasis(hd.Sector.lz(0, 33), 12, 34)
    assert b.dim == 7764392
    want = oracle.basis(oracle.sector(oracle.SECTOR_LZ, 0, 33), 12, 34)
    assert np.array_equal(b.states(), want)


def test_build_config2_rows_and_symmetry(oracle):
    """N=12, 2S=33 (dim 7,764,392, 2.59e9 full entries): three row blocks of
    the upper (reference) and full CRS bit-exact against the oracle, and the
    size-independent check that the full operator is symmetric,
    y.(Hx) = x.(Hy), through the device product."""
    say("build N=12 2S=33: sampled rows bit-exact, symmetry")
    b, t = models.sphere_problem(12, 33)
    bas = b.states()
    dim = b.dim
    blocks = [(0, 1500), (dim // 2, dim // 2 + 1500), (dim - 1500, dim)]
    up = hd.Hamiltonian(b, t, storage=hd.STORAGE_UPPER_REF, values=(hd.F64,))
    for r0, r1 in blocks:
        ref = oracle.build(12, 34, bas, oracle.LAYOUT_SPHERE, 34, t.ktable, t.vtable, full=False, row0=r0, row1=r1)
        _assert_crs_equal(up, ref, r0, r1)
    nnz_up = up.info()["nnz"]
    del up
    full = hd.Hamiltonian(b, t, storage=hd.STORAGE_FULL, values=(hd.F64,))
    for r0, r1 in blocks:
        ref = oracle.build(12, 34, bas, oracle.LAYOUT_SPHERE, 34, t.ktable, t.vtable, full=True, row0=r0, row1=r1)
        _assert_crs_equal(full, ref, r0, r1)
    info = full.info()
    assert info["nnz"] == 2590040752 and nnz_up < info["nnz"] < 2 * nnz_up
    rng = np.random.default_rng(7)
    x, y = rng.normal(size=dim), rng.normal(size=dim)
    a, c = y @ full.spmv(x), x @ full.spmv(y)
    assert abs(a - c) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(y) * 100


def test_empty_sector():
    say("empty sector")
    with pytest.raises(hd.HDError) as e:
        hd.FockBasis(hd.Sector.lz(1, 21), 8, 22)  # wrong parity
    assert e.value.code == hd.HD_ERR_EMPTY


def _random_sphere(two_s, seed):
    rng = np.random.default_rng(seed)
    K = two_s + 1
    return hd.sphere_ktable(two_s), rng.normal(size=K ** 3)


def _assert_crs_equal(h, ref, row0=0, row1=None):
    rp, col, val = h.crs(row0, row1)
    assert np.array_equal(rp, ref.rowptr)
    assert np.array_equal(col, ref.col)
    assert np.array_equal(val, ref.values()), np.max(np.abs(val - ref.values()))


@pytest.mark.parametrize("two_s,n,lz,seed", [(7, 3, 1, 0), (9, 4, 0, 1), (11, 4, 2, 2)])
def test_build_random_sphere_bit_exact(oracle, two_s, n, lz, seed):
    say("build random sphere", two_s, n, lz)
    kt, vt = _random_sphere(two_s, seed)
    K = two_s + 1
    b = hd.FockBasis(hd.Sector.lz(lz, two_s), n, K)
    t = hd.TermTables(hd.LAYOUT_SPHERE, K, kt, vt)
    bas = b.states()
    for storage, full in ((hd.STORAGE_UPPER_REF, False), (hd.STORAGE_FULL, True)):
        ref = oracle.build(n, K, bas, oracle.LAYOUT_SPHERE, K, kt, vt, full=full)
        h = hd.Hamiltonian(b, t, storage=storage)
        _assert_crs_equal(h, ref)
        assert h.info()["n_trimmed"] == ref.n_trimmed


def test_build_config1_bit_exact(oracle):
    say("build N=8 2S=21 Coulomb")
    b, t = models.sphere_problem(8, 21)
    bas = b.states()
    kt, vt = t.ktable, t.vtable
    up = oracle.build(8, 22, bas, oracle.LAYOUT_SPHERE, 22, kt, vt, full=False)
    h = hd.Hamiltonian(b, t, storage=hd.STORAGE_UPPER_REF)
    _assert_crs_equal(h, up)
    full = oracle.build(8, 22, bas, oracle.LAYOUT_SPHERE, 22, kt, vt, full=True)
    hf = hd.Hamiltonian(b, t, storage=hd.STORAGE_FULL)
    _assert_crs_equal(hf, full)
    # FP32 copy is the rounded FP64 copy
    _, _, v32 = hf.crs(precision=hd.F32)
    assert np.array_equal(v32, full.values().astype(np.float32))
    # the full matrix is symmetric to rounding and has the upper pattern mirrored
    import scipy.sparse as sp
    M = sp.csr_matrix((full.values(), full.col, full.rowptr), shape=(len(bas),) * 2)
    assert abs(M - M.T).max() <= 1e-15 * abs(M).max()


@pytest.mark.parametrize("X,Y,n,kx,ky,seed", [(3, 3, 3, 0, 0, 5), (3, 4, 5, 1, 2, 6), (6, 5, 4, 2, 3, 7)])
def test_build_ofl_complex_bit_exact(oracle, X, Y, n, kx, ky, seed):
    say("build ofl", X, Y, n)
    kt, vt, eps = models.ofl_tables(X, Y, seed)
    K = X * Y
    b = hd.FockBasis(hd.Sector.k2d(kx, ky, X, Y), n, K)
    t = hd.TermTables(hd.LAYOUT_OFL, K, kt, vt, eps)
    bas = b.states()
    for storage, full in ((hd.STORAGE_UPPER_REF, False), (hd.STORAGE_FULL, True)):
        ref = oracle.build(n, K, bas, oracle.LAYOUT_OFL, K, kt, vt, eps=eps, is_complex=True, full=full)
        h = hd.Hamiltonian(b, t, storage=storage)
        _assert_crs_equal(h, ref)


def test_build_config3_rows_bit_exact(oracle):
    say("build flux lattice 6x5 N=10 (1,001,402 states), sampled rows")
    b, t = models.ofl_problem(6, 5, 10, 0, 0, seed=0)
    assert b.dim == 1001402
    h = hd.Hamiltonian(b, t, storage=hd.STORAGE_FULL)
    bas = b.states()
    for r0 in (0, 500000, b.dim - 700):
        ref = oracle.build(10, 30, bas, oracle.LAYOUT_OFL, 30, t.ktable, t.vtable, eps=t.eps,
                           is_complex=True, full=True, row0=r0, row1=r0 + 700)
        _assert_crs_equal(h, ref, r0, r0 + 700)


def _dense_eigs(oracle, crs, dim, k):
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    full = crs
    M = sp.csr_matrix((full.values(), full.col, full.rowptr), shape=(dim, dim))
    if dim <= 3000:
        return np.linalg.eigvalsh(M.toarray())[:k]
    return np.sort(sla.eigsh(M, k=k, which="SA", tol=1e-14, ncv=60)[0])


def test_spmv_config1(oracle):
    say("spmv N=8 2S=21")
    b, t = models.sphere_problem(8, 21)
    bas = b.states()
    up = oracle.build(8, 22, bas, oracle.LAYOUT_SPHERE, 22, t.ktable, t.vtable)
    h = hd.Hamiltonian(b, t)
    rng = np.random.default_rng(0)
    x = rng.normal(size=b.dim)
    want = oracle.spmv_sym(up, x)
    for lanes in (4, 8, 16, 32, 64):
        h.set_lanes(lanes)
        got = h.spmv(x)
        assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
        g32 = h.spmv(x.astype(np.float32), precision=hd.F32)
        assert np.max(np.abs(g32 - want)) <= 1e-5 * np.max(np.abs(want))
        gc = h.spmv(x, precision=hd.CODED)
        assert np.max(np.abs(gc - want)) <= 1e-13 * np.max(np.abs(want))
    assert np.allclose(h.simple(x), want, rtol=0, atol=1e-13 * np.max(np.abs(want)))
    # the coded operator is the same FP64 matrix: identical products on unit vectors
    for i in (0, 17, b.dim - 1):
        e = np.zeros(b.dim)
        e[i] = 1.0
        assert np.array_equal(h.spmv(e, precision=hd.CODED), h.spmv(e))


def test_spmv_complex(oracle):
    say("spmv ofl 3x4 N=5 complex")
    kt, vt, eps = models.ofl_tables(3, 4, 3)
    b = hd.FockBasis(hd.Sector.k2d(1, 2, 3, 4), 5, 12)
    t = hd.TermTables(hd.LAYOUT_OFL, 12, kt, vt, eps)
    up = oracle.build(5, 12, b.states(), oracle.LAYOUT_OFL, 12, kt, vt, eps=eps, is_complex=True)
    h = hd.Hamiltonian(b, t)
    rng = np.random.default_rng(1)
    x = rng.normal(size=b.dim) + 1j * rng.normal(size=b.dim)
    want = oracle.spmv_sym(up, x)
    got = h.spmv(x)
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
    g32 = h.spmv(x.astype(np.complex64), precision=hd.F32)
    assert np.max(np.abs(g32 - want)) <= 1e-5 * np.max(np.abs(want))
    gc = h.spmv(x, precision=hd.CODED)
    assert np.max(np.abs(gc - want)) <= 1e-13 * np.max(np.abs(want))


def test_spmv_complex_large_dictionary(oracle):
    # 30 momenta with random complex couplings: far more distinct hop values
    # than fit the LDS copy of the dictionary, so the coded kernel reads it in place
    say("spmv ofl 6x5 N=4 complex, dictionary above the LDS limit")
    kt, vt, eps = models.ofl_tables(6, 5, 9)
    b = hd.FockBasis(hd.Sector.k2d(2, 1, 6, 5), 4, 30)
    t = hd.TermTables(hd.LAYOUT_OFL, 30, kt, vt, eps)
    up = oracle.build(4, 30, b.states(), oracle.LAYOUT_OFL, 30, kt, vt, eps=eps, is_complex=True)
    h = hd.Hamiltonian(b, t, values=(hd.F64, hd.CODED))
    ndict = h.device_coded()[3]
    assert ndict * 16 > 64 * 1024
    rng = np.random.default_rng(2)
    x = rng.normal(size=b.dim) + 1j * rng.normal(size=b.dim)
    want = oracle.spmv_sym(up, x)
    for lanes in (4, 32):
        h.set_lanes(lanes)
        gc = h.spmv(x, precision=hd.CODED)
        assert np.max(np.abs(gc - want)) <= 1e-13 * np.max(np.abs(want))
    for i in (0, b.dim // 2, b.dim - 1):
        e = np.zeros(b.dim, dtype=np.complex128)
        e[i] = 1.0
        assert np.array_equal(h.spmv(e, precision=hd.CODED), h.spmv(e))


@pytest.mark.parametrize("rows,nbytes", [(64, 4), (128, 0), (64, 3), (128, 3), (128, 4)])
@pytest.mark.parametrize("n,two_s,lz", [(4, 11, 0), (5, 13, 3), (6, 15, 0), (6, 15, 24)])
def test_tiled_columns_exact(oracle, n, two_s, lz, rows, nbytes):
    """Every column H e_i through the tiled layout equals the oracle's matrix
    column bit for bit (one nonzero term per output), so the layout holds
    exactly the coded operator's entries, signs and diagonal, with 4-byte and
    packed 3-byte entries."""
    say("tiled columns", n, two_s, lz, rows, nbytes)
    vm = models.coulomb_list(two_s, {1: 1.0, 3: 0.37, 5: 0.11})
    b = hd.FockBasis(hd.Sector.lz(lz, two_s), n, two_s + 1)
    t = models.sphere_tables(two_s, vm)
    K = two_s + 1
    full = oracle.build(n, K, b.states(), oracle.LAYOUT_SPHERE, K, t.ktable, t.vtable, full=True)
    h = hd.Hamiltonian(b, t, values=(hd.CODED,))
    ti = h.tile(rows, nbytes)
    assert ti["rows_per_group"] == rows and ti["groups"] == (b.dim + rows - 1) // rows
    assert ti["entry_bytes"] == (nbytes or 3)
    offdiag = len(full.col) - int(np.sum(full.col == np.repeat(np.arange(b.dim), np.diff(full.rowptr))))
    # 4-byte entries pad each chunk block to 4 entries, 3-byte entries each
    # (chunk, row) block (read as 12-byte quads)
    pad = 3 * ti["chunks"] * (rows if ti["entry_bytes"] == 3 else 1)
    assert offdiag <= ti["entries"] <= offdiag + pad and ti["entries"] % 4 == 0
    M = full.to_dense(b.dim, mirror=False)
    cols = range(b.dim) if b.dim <= 1500 else np.random.default_rng(0).choice(b.dim, 600, replace=False)
    for i in cols:
        e = np.zeros(b.dim)
        e[i] = 1.0
        assert np.array_equal(h.spmv(e, precision=hd.TILED), M[:, i]), i


def test_tiled_spmv_and_lanczos_config1(oracle):
    say("tiled spmv + lanczos N=8 2S=21")
    b, t = models.sphere_problem(8, 21)
    up = oracle.build(8, 22, b.states(), oracle.LAYOUT_SPHERE, 22, t.ktable, t.vtable)
    h = hd.Hamiltonian(b, t, values=(hd.F64, hd.CODED))
    h64 = hd.Hamiltonian(b, t, values=(hd.CODED,))
    h64.tile(64)
    ti = h.tile()
    say(ti)
    assert ti["rows_per_group"] == 128 and ti["entry_bytes"] == 3
    rng = np.random.default_rng(4)
    x = rng.normal(size=b.dim)
    want = oracle.spmv_sym(up, x)
    got = h.spmv(x, precision=hd.TILED)
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
    assert np.max(np.abs(h64.spmv(x, precision=hd.TILED) - want)) <= 1e-13 * np.max(np.abs(want))
    r64 = h.lanczos(nev=5, tol=1e-12, seed=0)
    rt = h.lanczos(nev=5, tol=1e-12, seed=0, precision=hd.TILED)
    assert rt.converged and np.all(rt.residuals <= 1e-12)
    assert np.max(np.abs(rt.evals - r64.evals)) < 1e-11
    # the tiled product is deterministic
    assert np.array_equal(h.spmv(x, precision=hd.TILED), got)


def test_tiled_block_orders_identical():
    """The block order only changes which workgroup runs a row group, so
    every order (launch order, windows with a partial last one, one window
    over the grid, the autotuned choice) gives the same product bit for bit."""
    say("tiled block orders N=8 2S=21")
    b, t = models.sphere_problem(8, 21)
    h = hd.Hamiltonian(b, t, values=(hd.CODED,))
    ti = h.tile()
    groups = ti["groups"]
    assert groups == (b.dim + ti["rows_per_group"] - 1) // ti["rows_per_group"] and ti["block_window"] >= 0
    x = np.random.default_rng(5).normal(size=b.dim)
    h.tile_order(0)
    want = h.spmv(x, precision=hd.TILED)
    for w in (1, 3, 16, (groups + 7) // 8, groups, -1):
        got_w = h.tile_order(w)
        assert got_w == w or (w == -1 and got_w >= 0)
        assert np.array_equal(h.spmv(x, precision=hd.TILED), want), w
    with pytest.raises(hd.HDError):
        h.tile_order(-2)


@pytest.mark.parametrize("rows", [64, 128])
@pytest.mark.parametrize("X,Y,n,kx,ky,seed", [(3, 4, 5, 1, 2, 6), (6, 5, 4, 2, 3, 7), (6, 5, 5, 0, 0, 8)])
def test_tiled_columns_exact_complex(oracle, X, Y, n, kx, ky, seed, rows):
    """Complex flux-lattice matrices on the tiled layout (8 complex values per
    staged line, dictionary read from L2): every column H e_i equals the
    oracle's full-storage column bit for bit, and the Lanczos eigenvalues on
    the tiled operator equal the dense ones."""
    say("tiled columns complex", X, Y, n, kx, ky, rows)
    kt, vt, eps = models.ofl_tables(X, Y, seed)
    K = X * Y
    b = hd.FockBasis(hd.Sector.k2d(kx, ky, X, Y), n, K)
    t = hd.TermTables(hd.LAYOUT_OFL, K, kt, vt, eps)
    full = oracle.build(n, K, b.states(), oracle.LAYOUT_OFL, K, kt, vt, eps=eps, is_complex=True, full=True)
    h = hd.Hamiltonian(b, t, values=(hd.CODED,))
    ti = h.tile(rows)
    assert ti["rows_per_group"] == rows and ti["entry_bytes"] == 4
    M = full.to_dense(b.dim, mirror=False)
    cols = range(b.dim) if b.dim <= 1500 else np.random.default_rng(0).choice(b.dim, 600, replace=False)
    for i in cols:
        e = np.zeros(b.dim, dtype=np.complex128)
        e[i] = 1.0
        assert np.array_equal(h.spmv(e, precision=hd.TILED), M[:, i]), i
    rng = np.random.default_rng(3)
    x = rng.normal(size=b.dim) + 1j * rng.normal(size=b.dim)
    want = M @ x
    got = h.spmv(x, precision=hd.TILED)
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
    nev = min(4, b.dim - 1)
    if 60 < b.dim <= 2000:
        r = h.lanczos(nev=nev, tol=1e-11, seed=0, precision=hd.TILED)
        assert r.converged and np.all(r.residuals <= 1e-11)
        ev = np.linalg.eigvalsh(M)[:nev]
        assert np.max(np.abs(r.evals - ev) / np.maximum(1, np.abs(ev))) < 1e-10


def test_tiled_complex_rejects_packed_entries():
    kt, vt, eps = models.ofl_tables(3, 3, 1)
    b = hd.FockBasis(hd.Sector.k2d(0, 0, 3, 3), 3, 9)
    h = hd.Hamiltonian(b, hd.TermTables(hd.LAYOUT_OFL, 9, kt, vt, eps), values=(hd.CODED,))
    with pytest.raises(hd.HDError):
        h.tile(0, 3)


def test_lanczos_config1(oracle):
    say("lanczos N=8 2S=21 nev=5")
    b, t = models.sphere_problem(8, 21)
    bas = b.states()
    up = oracle.build(8, 22, bas, oracle.LAYOUT_SPHERE, 22, t.ktable, t.vtable)
    full = oracle.build(8, 22, bas, oracle.LAYOUT_SPHERE, 22, t.ktable, t.vtable, full=True)
    want = _dense_eigs(oracle, full, b.dim, 5)
    h = hd.Hamiltonian(b, t)
    r = h.lanczos(nev=5, tol=1e-12, seed=0)
    say(r)
    assert r.converged and np.all(r.residuals <= 1e-12)
    assert np.max(np.abs(r.evals - want) / np.maximum(1, np.abs(want))) < 1e-10
    # eigenvector residual through the oracle's naive SpMV
    x0 = h.eigenvector(0)
    res = oracle.spmv_sym(up, x0) - r.evals[0] * x0
    assert np.linalg.norm(res) < 1e-11 * max(1, abs(r.evals[0]))
    r32 = h.lanczos(nev=5, tol=1e-10, precision=hd.F32, seed=0)
    say(r32)
    assert r32.converged and np.all(r32.residuals <= 1e-10)
    assert np.max(np.abs(r32.evals - want) / np.maximum(1, np.abs(want))) < 1e-10
    rc = h.lanczos(nev=5, tol=1e-12, precision=hd.CODED, seed=0)
    say(rc)
    assert rc.converged and np.all(rc.residuals <= 1e-12)
    assert np.max(np.abs(rc.evals - want) / np.maximum(1, np.abs(want))) < 1e-10


def test_tiled_x32_product_exact_for_fp32_inputs(oracle):
    """HD_TILED_X32 reads x in FP32 and sums the FP64 dictionary values in FP64:
    its product equals the FP64 tiled product of the same (FP32-valued) x bit
    for bit, and the oracle's naive SpMV of it to 1e-13."""
    say("tiled x32 product N=8 2S=21")
    b, t = models.sphere_problem(8, 21)
    up = oracle.build(8, 22, b.states(), oracle.LAYOUT_SPHERE, 22, t.ktable, t.vtable)
    h = hd.Hamiltonian(b, t, values=(hd.CODED,))
    for rows, nbytes in ((128, 4), (64, 3)):
        h.tile(rows, nbytes)
        x32 = np.random.default_rng(rows).normal(size=b.dim).astype(np.float32)
        got = h.spmv(x32, precision=hd.TILED_X32)
        assert got.dtype == np.float64
        assert np.array_equal(got, h.spmv(x32.astype(np.float64), precision=hd.TILED))
        want = oracle.spmv_sym(up, x32.astype(np.float64))
        assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))


@pytest.mark.parametrize("tiled", [True, False])
def test_mixed_precision_lanczos(oracle, tiled):
    """HD_F32: the restart cycles after the coefficients have become small run
    on the FP32-input operator (tiled, or the FP32-value CRS); the eigenpairs
    equal the dense ones to 1e-10 with FP64 residuals below tol."""
    say("mixed-precision Lanczos N=8 2S=21, tiled", tiled)
    b, t = models.sphere_problem(8, 21)
    full = oracle.build(8, 22, b.states(), oracle.LAYOUT_SPHERE, 22, t.ktable, t.vtable, full=True)
    want = _dense_eigs(oracle, full, b.dim, 5)
    h = hd.Hamiltonian(b, t, values=(hd.F64, hd.F32, hd.CODED))
    if tiled:
        h.tile()
    r = h.lanczos(nev=5, tol=1e-10, precision=hd.F32, seed=0)
    say(r)
    assert r.converged and np.all(r.residuals <= 1e-10)
    assert r.matvecs_f32 > 0 and r.matvecs_f64 > 0
    assert np.max(np.abs(r.evals - want) / np.maximum(1, np.abs(want))) < 1e-10


@pytest.mark.parametrize("case", ["sphere", "flux_lattice"])
def test_partial_vs_full_reorthogonalisation(oracle, case):
    """Partial reorthogonalisation (the default) reaches the same eigenpairs as
    full reorthogonalisation, with explicit FP64 residuals below tolerance and
    fewer orthogonalisation passes than steps; no spurious copies of converged
    eigenvalues appear (the eigenvalues equal the dense ones, which are
    non-degenerate here)."""
    if case == "sphere":
        b, t = models.sphere_problem(8, 21)
        full = oracle.build(8, 22, b.states(), oracle.LAYOUT_SPHERE, 22, t.ktable, t.vtable, full=True)
        nev = 5
    else:
        kt, vt, eps = models.ofl_tables(3, 4, 11)
        b = hd.FockBasis(hd.Sector.k2d(0, 0, 3, 4), 5, 12)
        t = hd.TermTables(hd.LAYOUT_OFL, 12, kt, vt, eps)
        full = oracle.build(5, 12, b.states(), oracle.LAYOUT_OFL, 12, kt, vt, eps=eps, is_complex=True, full=True)
        nev = 6
    say("partial vs full reorthogonalisation", case)
    want = _dense_eigs(oracle, full, b.dim, nev)
    h = hd.Hamiltonian(b, t)
    for prec in (hd.F64, hd.F32):
        tol = 1e-12 if prec == hd.F64 else 1e-10
        rp = h.lanczos(nev=nev, tol=tol, precision=prec, seed=0)
        rf = h.lanczos(nev=nev, tol=tol, precision=prec, seed=0, full_reorth=True)
        say(case, prec, "partial", rp)
        say(case, prec, "full", rf)
        for r in (rp, rf):
            assert r.converged and np.all(r.residuals <= tol)
            assert np.max(np.abs(r.evals - want) / np.maximum(1, np.abs(want))) < 1e-10
        # real matrices: Simon's signed estimate saves most passes; complex ones
        # propagate a magnitude bound, which may ask for a pass at every step
        assert rp.reorth_steps <= rf.reorth_steps
        if prec == hd.F64 and case == "sphere":
            assert rp.reorth_steps < rf.reorth_steps / 2


def test_lanczos_laughlin_zero_mode():
    say("lanczos laughlin N=5 2S=12 V1 only")
    b, t = models.sphere_problem(5, 12, 0, pseudopotentials=[0.0, 1.0])
    h = hd.Hamiltonian(b, t)
    r = h.lanczos(nev=3, tol=1e-11, seed=3)
    assert r.converged
    assert abs(r.evals[0]) < 1e-10 and r.evals[1] > 1e-3


def test_lanczos_complex(oracle):
    say("lanczos ofl 3x4 N=5 complex nev=6")
    kt, vt, eps = models.ofl_tables(3, 4, 11)
    b = hd.FockBasis(hd.Sector.k2d(0, 0, 3, 4), 5, 12)
    t = hd.TermTables(hd.LAYOUT_OFL, 12, kt, vt, eps)
    full = oracle.build(5, 12, b.states(), oracle.LAYOUT_OFL, 12, kt, vt, eps=eps, is_complex=True,
                        full=True)
    want = _dense_eigs(oracle, full, b.dim, 6)
    h = hd.Hamiltonian(b, t)
    for prec in (hd.F64, hd.F32, hd.CODED):
        r = h.lanczos(nev=6, tol=1e-10, precision=prec, seed=0)
        say(prec, r)
        assert r.converged
        assert np.max(np.abs(r.evals - want) / np.maximum(1, np.abs(want))) < 1e-10


@pytest.mark.parametrize("block", [4, 8, 16])
def test_block_lanczos_config1(oracle, block):
    say("block lanczos N=8 2S=21 nev=5 block", block)
    b, t = models.sphere_problem(8, 21)
    bas = b.states()
    up = oracle.build(8, 22, bas, oracle.LAYOUT_SPHERE, 22, t.ktable, t.vtable)
    full = oracle.build(8, 22, bas, oracle.LAYOUT_SPHERE, 22, t.ktable, t.vtable, full=True)
    want = _dense_eigs(oracle, full, b.dim, 5)
    h = hd.Hamiltonian(b, t, values=(hd.F64, hd.CODED))
    for prec in (hd.CODED, hd.F64):
        r = h.lanczos(nev=5, tol=1e-11, seed=0, block=block, precision=prec)
        say(prec, r)
        assert r.converged and np.all(r.residuals <= 1e-11)
        assert np.max(np.abs(r.evals - want) / np.maximum(1, np.abs(want))) < 1e-10
    x0 = h.eigenvector(0)
    res = oracle.spmv_sym(up, x0) - r.evals[0] * x0
    assert np.linalg.norm(res) < 1e-10 * max(1, abs(r.evals[0]))


def test_block_lanczos_user_start():
    say("block lanczos with a start vector in column 0 (N=6 2S=15)")
    b, t = models.sphere_problem(6, 15)
    h = hd.Hamiltonian(b, t, values=(hd.F64, hd.CODED))
    r0 = h.lanczos(nev=1, tol=1e-11, seed=0)
    r1 = h.lanczos(nev=1, tol=1e-11, seed=0, block=4, start=h.eigenvector(0))
    say(r0, r1)
    assert abs(r1.evals[0] - r0.evals[0]) < 1e-10 * max(1, abs(r0.evals[0]))
    assert r1.matvecs <= 4 * 8
