"""hd_spmv_host_noalloc (Simple() staged through the library's fixed device
array) against hd_spmv_host (the allocating Simple()), bit for bit, through
the C ABI: two calls in a row with the second on a smaller sector, complex and
coded handles, and the size limit."""
import sys

import numpy as np
import pytest

import hamiltonian_diagonalization_amd as hd
from hamiltonian_diagonalization_amd import models

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def test_noalloc_simple_matches_simple_two_calls():
    say("no-alloc Simple: N=8 2S=21 then N=6 2S=15")
    b1, t1 = models.sphere_problem(8, 21)
    h1 = hd.Hamiltonian(b1, t1, values=(hd.F64,))
    b2, t2 = models.sphere_problem(6, 15)
    h2 = hd.Hamiltonian(b2, t2, values=(hd.F64,))
    assert b2.dim < b1.dim
    rng = np.random.default_rng(0)
    x1, x2 = rng.normal(size=b1.dim), rng.normal(size=b2.dim)
    want1, want2 = h1.simple(x1), h2.simple(x2)
    got1 = h1.simple_noalloc(x1)
    got2 = h2.simple_noalloc(x2)  # smaller input: the staging array keeps the first call's tail
    assert np.array_equal(got1, want1)
    assert np.array_equal(got2, want2)
    # and back to the larger one
    assert np.array_equal(h1.simple_noalloc(x1), want1)


def test_noalloc_simple_complex_and_coded():
    say("no-alloc Simple: complex flux lattice 3x4 N=5, coded and tiled handles")
    kt, vt, eps = models.ofl_tables(3, 4, 3)
    b = hd.FockBasis(hd.Sector.k2d(1, 2, 3, 4), 5, 12)
    h = hd.Hamiltonian(b, hd.TermTables(hd.LAYOUT_OFL, 12, kt, vt, eps))
    rng = np.random.default_rng(1)
    x = rng.normal(size=b.dim) + 1j * rng.normal(size=b.dim)
    assert np.array_equal(h.simple_noalloc(x), h.simple(x))
    bs, ts = models.sphere_problem(8, 21)
    hc = hd.Hamiltonian(bs, ts, values=(hd.CODED,))
    hc.tile()
    xs = rng.normal(size=bs.dim)
    assert np.array_equal(hc.simple_noalloc(xs), hc.simple(xs))


def test_noalloc_simple_rejects_large_sector():
    say("no-alloc Simple: N=12 2S=33 (7,764,392 states) is beyond the staging limit")
    b, t = models.sphere_problem(12, 33)
    h = hd.Hamiltonian(b, t, storage=hd.STORAGE_MATRIX_FREE)
    with pytest.raises(hd.HDError) as e:
        h.simple_noalloc(np.zeros(b.dim))
    assert e.value.code == hd.HD_ERR_ARG and "32 MiB" in str(e.value)
