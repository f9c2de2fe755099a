"""Every row-walk and row-sort kernel of the build and the matrix-free product
kernels on the same matrices.  The default build fills and sorts rows in one
pass (one wavefront per row, split ranking, rows of up to 512 entries);
HD_BUILD_FILL=split stages the wavefront walk's entries for the register row
sort with 32-bit keys, =rank uses the wavefront walk with the rank table and
the 64-bit keys, and =thread one thread per row and the LDS row sort of up to
1,024 entries (the kernels that run for longer rows, for dimensions from 2^23
or when the tables do not fit in LDS).  Each is compared with the oracle's
CRS bit for bit; the matrix-free products with the stored CRS to 1e-13."""
import sys

import numpy as np
import pytest

import hamiltonian_diagonalization_amd as hd
from hamiltonian_diagonalization_amd import models

pytestmark = pytest.mark.gpu

PATHS = ["auto", "split", "rank", "thread"]


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _set_path(monkeypatch, path):
    if path == "auto":
        monkeypatch.delenv("HD_BUILD_FILL", raising=False)
    else:
        monkeypatch.setenv("HD_BUILD_FILL", path)


def _assert_crs_equal(h, ref):
    rp, col, val = h.crs()
    assert np.array_equal(rp, ref.rowptr)
    assert np.array_equal(col, ref.col)
    assert np.array_equal(val, ref.values())


@pytest.mark.parametrize("path", PATHS)
def test_fill_paths_sphere_bit_exact(oracle, monkeypatch, path):
    say("fill path", path, "sphere N=8 2S=21: upper and full CRS, FP32 and coded copies")
    _set_path(monkeypatch, path)
    b, t = models.sphere_problem(8, 21)
    bas = b.states()
    for storage, full in ((hd.STORAGE_UPPER_REF, False), (hd.STORAGE_FULL, True)):
        ref = oracle.build(8, 22, bas, oracle.LAYOUT_SPHERE, 22, t.ktable, t.vtable, full=full)
        h = hd.Hamiltonian(b, t, storage=storage)
        _assert_crs_equal(h, ref)
        assert h.info()["n_trimmed"] == ref.n_trimmed
    _, _, v32 = h.crs(precision=hd.F32)
    assert np.array_equal(v32, ref.values().astype(np.float32))
    # the coded copy holds the same operator (a wrong code is an O(1) error)
    x = np.random.default_rng(0).normal(size=b.dim)
    y64 = h.spmv(x, precision=hd.F64)
    assert np.max(np.abs(h.spmv(x, precision=hd.CODED) - y64)) <= 1e-13 * np.max(np.abs(y64))


@pytest.mark.parametrize("path", PATHS)
def test_fill_paths_flux_lattice_complex_bit_exact(oracle, monkeypatch, path):
    say("fill path", path, "flux lattice 3x4 N=5 (1,2) complex: upper and full CRS")
    _set_path(monkeypatch, path)
    kt, vt, eps = models.ofl_tables(3, 4, 6)
    b = hd.FockBasis(hd.Sector.k2d(1, 2, 3, 4), 5, 12)
    t = hd.TermTables(hd.LAYOUT_OFL, 12, kt, vt, eps)
    bas = b.states()
    for storage, full in ((hd.STORAGE_UPPER_REF, False), (hd.STORAGE_FULL, True)):
        ref = oracle.build(5, 12, bas, oracle.LAYOUT_OFL, 12, kt, vt, eps=eps, is_complex=True, full=full)
        _assert_crs_equal(hd.Hamiltonian(b, t, storage=storage), ref)


@pytest.mark.parametrize("path", PATHS)
def test_fill_paths_unconstrained_sector(oracle, monkeypatch, path):
    say("fill path", path, "no sector constraint N=3 2S=7")
    _set_path(monkeypatch, path)
    t = models.sphere_tables(7, models.coulomb_list(7))
    bn = hd.FockBasis(hd.Sector.none(), 3, 8)
    ref = oracle.build(3, 8, bn.states(), oracle.LAYOUT_SPHERE, 8, t.ktable, t.vtable, full=True)
    _assert_crs_equal(hd.Hamiltonian(bn, t, storage=hd.STORAGE_FULL), ref)


@pytest.mark.parametrize("path", PATHS)
def test_matrix_free_paths(monkeypatch, path):
    say("matrix-free path", path, "sphere N=8 2S=21 and flux lattice 3x4 N=4 complex")
    _set_path(monkeypatch, path)
    rng = np.random.default_rng(1)
    cases = [(models.sphere_problem(8, 21), False),
             (models.ofl_problem(3, 4, 4, kx=1, ky=2, seed=2), True)]
    for (b, t), cplx in cases:
        x = rng.normal(size=b.dim) + (1j * rng.normal(size=b.dim) if cplx else 0)
        y_crs = hd.Hamiltonian(b, t, storage=hd.STORAGE_FULL, values=(hd.F64,)).spmv(x)
        y_free = hd.Hamiltonian(b, t, storage=hd.STORAGE_MATRIX_FREE).spmv(x)
        assert np.max(np.abs(y_free - y_crs)) <= 1e-13 * np.max(np.abs(y_crs))

