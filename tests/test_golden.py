"""Golden fixtures (tests/golden/golden.json, written by tests/golden/make_golden.py
with the CPU oracle): the oracle must reproduce them (CPU), and the device path
must match them bit for bit in basis and CRS, and to 1e-10 in eigenvalues (GPU)."""
import hashlib
import json
import os

import numpy as np
import pytest

import hamiltonian_diagonalization_amd as hd
from hamiltonian_diagonalization_amd import models

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GOLD = json.load(open(os.path.join(HERE, "golden.json")))
CASES = GOLD["cases"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def crs_digests(rowptr, col, val, n_trimmed=None):
    d = {"nnz": int(len(col)), "rowptr_int32": sha(np.asarray(rowptr).astype(np.int32)),
         "col": sha(np.asarray(col).astype(np.int32)), "val": sha(val)}
    if n_trimmed is not None:
        d["n_trimmed"] = int(n_trimmed)
    return d


def case_inputs(c, oracle=None):
    """(sector args, N, Norb, ktable, vtable, eps) of a fixture case."""
    if c["model"] == "sphere":
        vm = np.array([float(v) for v in c["pseudopotentials"]])
        kt = hd.sphere_ktable(c["two_s"])
        vt = oracle.sphere_vtable(c["two_s"], vm) if oracle else hd.sphere_vtable(c["two_s"], vm)
        return (hd.SECTOR_LZ, (c["lz"], c["two_s"])), c["n"], c["two_s"] + 1, kt, vt, None
    kt, vt, eps = models.ofl_tables(c["dim_x"], c["dim_y"], c["seed"])
    return (hd.SECTOR_K2D, (c["kx"], c["ky"], c["dim_x"], c["dim_y"])), c["n"], c["dim_x"] * c["dim_y"], kt, vt, eps


# ------------------------------------------------------------------------ CPU

def test_coulomb_values_match_fixture():
    for ts, vals in GOLD["coulomb"].items():
        assert ["%.17g" % v for v in models.coulomb_list(int(ts))] == vals


@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_oracle_reproduces_fixture(oracle, c):
    (kind, args), n, K, kt, vt, eps = case_inputs(c, oracle)
    bas = oracle.basis(oracle.sector(kind, *args), n, K)
    assert len(bas) == c["dim"] and sha(bas) == c["basis"]
    layout = oracle.LAYOUT_SPHERE if c["model"] == "sphere" else oracle.LAYOUT_OFL
    cplx = c["model"] == "ofl"
    up = oracle.build(n, K, bas, layout, K, kt, vt, eps=eps, is_complex=cplx)
    assert crs_digests(up.rowptr, up.col, up.values(), up.n_trimmed) == c["upper"]
    if c["model"] == "sphere" and "V1 only" in c["name"]:
        # Laughlin flux with V1 alone: the lowest eigenvalue is exactly zero
        assert abs(float(c["eigenvalues"][0])) < 1e-12


def test_config1_basis_file(oracle):
    bas = np.load(os.path.join(HERE, "basis_n8_2s21_lz0.npy"))
    assert np.array_equal(bas, oracle.basis(oracle.sector(oracle.SECTOR_LZ, 0, 21), 8, 22))


# ------------------------------------------------------------------------ GPU

@pytest.mark.gpu
@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_device_matches_fixture(c):
    print("[gpu-case] golden", c["name"], flush=True)
    (kind, args), n, K, kt, vt, eps = case_inputs(c)
    b = hd.FockBasis(hd.Sector(kind, *args), n, K)
    assert b.dim == c["dim"] and sha(b.states()) == c["basis"]
    layout = hd.LAYOUT_SPHERE if c["model"] == "sphere" else hd.LAYOUT_OFL
    t = hd.TermTables(layout, K, kt, vt, eps)
    hu = hd.Hamiltonian(b, t, storage=hd.STORAGE_UPPER_REF)
    rp, col, val = hu.crs()
    assert crs_digests(rp, col, val, hu.info()["n_trimmed"]) == c["upper"]
    hf = hd.Hamiltonian(b, t, storage=hd.STORAGE_FULL)
    rp, col, val = hf.crs()
    assert crs_digests(rp, col, val, hf.info()["n_trimmed"]) == c["full"]
    want = np.array([float(e) for e in c["eigenvalues"]])
    k = len(want)
    if b.dim >= 60:
        r = hf.lanczos(nev=k, tol=1e-11, seed=0)
        got = r.evals
    else:
        got = hf.full_diagonalize(k) if c["model"] == "sphere" else None
        if got is None:
            r = hf.lanczos(nev=k, tol=1e-11, seed=0, ncv=b.dim)
            got = r.evals
    assert np.max(np.abs(np.asarray(got) - want) / np.maximum(1, np.abs(want))) < 1e-10
