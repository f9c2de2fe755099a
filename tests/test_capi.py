"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and the host-only table routines match the oracle."""
import ctypes

import numpy as np
import pytest

import hamiltonian_diagonalization_amd as hd


def test_exports_every_header_symbol():
    names = hd.header_functions()
    assert len(names) >= 30
    L = hd.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the package binds exactly the declared set
    assert sorted(hd.SIGNATURES) == names


@pytest.mark.parametrize("max_lz", [3, 7, 21, 33])
def test_sphere_ktable_matches_oracle(oracle, max_lz):
    assert np.array_equal(hd.sphere_ktable(max_lz), oracle.sphere_ktable(max_lz))


@pytest.mark.parametrize("max_lz,vm", [(9, [0.0, 1.0]), (7, [0.3, 1.0, 0.0, 0.25]),
                                      (12, [0.0, 1.0, 0.0, 0.5, 0.0, 0.1])])
def test_sphere_vtable_matches_exact_rationals(oracle, max_lz, vm):
    # both sides take every CG correctly rounded from its exact square and sum
    # in the reference's order: the tables are bit-identical
    got = hd.sphere_vtable(max_lz, vm)
    want = oracle.sphere_vtable(max_lz, vm)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("two_s", [21, 27, 33, 41])
def test_sphere_vtable_coulomb_bit_identical(oracle, two_s):
    """The benchmark fluxes (2S = 21, 27, 33) and one beyond them.  At 2S = 33
    a long double square root of the 96-bit-truncated CG^2, rounded to double,
    missed the correctly rounded CG in one ulp for entries of 32 table cells;
    the library now checks its square roots against the exact rational."""
    from hamiltonian_diagonalization_amd import models
    vm = models.coulomb_list(two_s)
    got = hd.sphere_vtable(two_s, vm)
    want = oracle.sphere_vtable(two_s, list(vm))
    assert np.array_equal(got, want)


def test_cg_closed_forms(oracle):
    """<j m; j -m | 0 0> = (-1)^(j-m)/sqrt(2j+1) and orthonormality of the
    CG matrix at fixed M (pins the exact-rational oracle CG)."""
    j = 7  # doubled: j = 7/2
    for m in range(-j, j + 1, 2):
        want = (-1) ** ((j - m) // 2) / np.sqrt(j + 1)
        assert abs(oracle.cg_value(j, m, j, -m, 0, 0) - want) < 1e-15
    j1, j2, M = 5, 4, 1
    ms = [m1 for m1 in range(-j1, j1 + 1, 2) if abs(M - m1) <= j2]
    Js = list(range(abs(j1 - j2), j1 + j2 + 1, 2))
    C = np.array([[oracle.cg_value(j1, m1, j2, M - m1, J, M) if abs(M) <= J else 0.0
                   for J in Js if J >= abs(M)] for m1 in ms])
    assert np.allclose(C.T @ C, np.eye(C.shape[1]), atol=1e-14)


def test_coulomb_pseudopotentials_physical():
    from hamiltonian_diagonalization_amd import models
    from math import gamma, factorial
    v = models.coulomb_list(33)
    assert len(v) == 32 and np.all(v[0::2] == 0.0)
    odd = v[1::2]
    assert np.all(np.diff(odd) < 0) and np.all(odd > 0)
    # large-sphere limit: planar V_m = Gamma(m + 1/2) / (2 m!), approached
    # with O(m^2/S) corrections
    for m in (1, 3, 5):
        planar = gamma(m + 0.5) / (2 * factorial(m))
        assert abs(v[m] / planar - 1) < 0.15
        assert abs(models.coulomb_pseudopotential(4001, m) / planar - 1) < 2e-3


def test_tables_reject_bad_arguments():
    kt = np.zeros(8, dtype=np.uint16)
    vt = np.zeros(8)
    h = ctypes.c_void_p()
    code = hd.lib().hd_tables_create(7, 2, kt.ctypes.data_as(ctypes.c_void_p),
                                     vt.ctypes.data_as(ctypes.c_void_p), 0, None,
                                     ctypes.byref(h))
    assert code == hd.HD_ERR_ARG
    assert b"layout" in hd.lib().hd_last_error()
    with pytest.raises(hd.HDError):
        hd.sphere_vtable(5, [1.0] * 6)  # more than maxLz coefficients
