"""Batched sectors x parameter points (hd_batch_*) against the oracle: the shared
coded pattern, the per-point dictionaries and diagonals (bit-exact against
single-point oracle builds), the batched SpMM, and the eigenvalues of every
(sector, point) problem."""
import sys

import numpy as np
import pytest

import hamiltonian_diagonalization_amd as hd
from hamiltonian_diagonalization_amd import models

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _points(n, seed):
    rng = np.random.default_rng(seed)
    return [(0.5 + rng.random(), 0.5 * rng.random(), 0.2 * rng.random()) for _ in range(n)]


def _vtables(two_s, pts):
    return [models.coulomb_list(two_s, {1: v1, 3: v3, 5: v5}) for v1, v3, v5 in pts]


def _batch(n, two_s, lzs, pts):
    bases = [hd.FockBasis(hd.Sector.lz(lz, two_s), n, two_s + 1) for lz in lzs]
    vms = _vtables(two_s, pts)
    tables = [models.sphere_tables(two_s, vm) for vm in vms]
    return hd.BatchProblem(bases, tables), bases, tables


def _oracle_full(oracle, n, two_s, bas, t):
    K = two_s + 1
    return oracle.build(n, K, bas, oracle.LAYOUT_SPHERE, K, t.ktable, t.vtable, full=True)


def test_batch_pattern_dictionary_diagonal_bit_exact(oracle):
    say("batch pattern N=6 2S=15, 5 sectors x 5 points")
    n, two_s, lzs = 6, 15, [0, 10, 30, 58, 60]
    pts = _points(5, 1)
    bp, bases, tables = _batch(n, two_s, lzs, pts)
    rp, col, code, dic, dg = bp.pattern()
    assert bp.rows == sum(b.dim for b in bases)
    for s, b in enumerate(bases):
        r0, dim = int(bp.row0[s]), b.dim
        bas = b.states()
        rows = slice(r0, r0 + dim + 1)
        lrp = rp[rows] - rp[r0]
        e0, e1 = int(rp[r0]), int(rp[r0 + dim])
        lcol = col[e0:e1] - r0
        lcode = code[e0:e1]
        for p, t in enumerate(tables):
            ref = _oracle_full(oracle, n, two_s, bas, t)
            assert ref.n_trimmed == 0
            assert np.array_equal(lrp, ref.rowptr) and np.array_equal(lcol, ref.col)
            off = lcode != 0xFFFF
            vals = dic[(lcode[off] & 0x7FFF).astype(np.int64), p]
            vals = np.where(lcode[off] & 0x8000, -vals, vals)
            assert np.array_equal(vals, ref.values()[off])
            # the diagonal entries of the oracle CRS are the batch's diagonal column
            diag_ref = ref.values()[~off]
            assert np.array_equal(dg[r0:r0 + dim, p], diag_ref)


def test_batch_spmm_matches_oracle(oracle):
    say("batch spmm N=6 2S=15, 4 sectors x 7 points")
    n, two_s, lzs = 6, 15, [0, 4, 20, 44]
    pts = _points(7, 2)
    bp, bases, tables = _batch(n, two_s, lzs, pts)
    x = np.random.default_rng(3).normal(size=(bp.rows, bp.points))
    y = bp.spmm(x)
    K = two_s + 1
    for s, b in enumerate(bases):
        r0, dim = int(bp.row0[s]), b.dim
        bas = b.states()
        for p, t in enumerate(tables):
            up = oracle.build(n, K, bas, oracle.LAYOUT_SPHERE, K, t.ktable, t.vtable)
            want = oracle.spmv_sym(up, x[r0:r0 + dim, p])
            got = y[r0:r0 + dim, p]
            assert np.max(np.abs(got - want)) <= 1e-13 * max(1.0, np.max(np.abs(want)))


def test_batch_solve_all_sectors_vs_dense(oracle):
    say("batch solve N=6 2S=15: 31 sectors x 6 points, nev=3 (dense + Lanczos)")
    n, two_s = 6, 15
    lzs = list(range(0, 61, 2))
    pts = _points(6, 4)
    bp, bases, tables = _batch(n, two_s, lzs, pts)
    ev, rs, st = bp.solve(nev=3, tol=1e-10, seed=0)
    say(st)
    assert st["converged"] == st["problems"] == len(lzs) * len(pts)
    assert st["lanczos_problems"] > 0 and st["dense_problems"] > 0
    assert st["max_residual"] <= 1e-10
    for s, b in enumerate(bases):
        bas = b.states()
        for p, t in enumerate(tables):
            ref = _oracle_full(oracle, n, two_s, bas, t)
            M = np.zeros((b.dim, b.dim))
            for i in range(b.dim):
                M[i, ref.col[ref.rowptr[i]:ref.rowptr[i + 1]]] = ref.values()[ref.rowptr[i]:ref.rowptr[i + 1]]
            want = np.linalg.eigvalsh(M)[:3]
            k = min(3, b.dim)
            got = ev[s, p, :k]
            assert np.max(np.abs(got - want[:k]) / np.maximum(1, np.abs(want[:k]))) < 1e-10, (lzs[s], p)
            assert np.all(np.isnan(ev[s, p, k:]))
    # the captured graph replays exactly the eager launches
    ev2, _, st2 = bp.solve(nev=3, tol=1e-10, seed=0, use_graph=False)
    assert np.array_equal(ev, ev2, equal_nan=True) and st2["steps"] == st["steps"]


def test_batch_solve_config1_sectors_vs_single_lanczos():
    say("batch solve N=8 2S=21: 57 sectors x 4 points, nev=3, vs single-problem Lanczos")
    n, two_s = 8, 21
    lzs = list(range(0, 113, 2))
    pts = _points(4, 5)
    bp, bases, tables = _batch(n, two_s, lzs, pts)
    ev, rs, st = bp.solve(nev=3, tol=1e-10, seed=1)
    say(st)
    assert st["converged"] == st["problems"]
    for s in (0, 3, 20):
        for p in (0, 3):
            h = hd.Hamiltonian(bases[s], tables[p], values=(hd.F64,))
            r = h.lanczos(nev=3, tol=1e-11, seed=0)
            assert np.max(np.abs(ev[s, p] - r.evals) / np.maximum(1, np.abs(r.evals))) < 1e-10, (lzs[s], p)


def test_batch_partial_vs_full_reorthogonalisation():
    """Partial reorthogonalisation (the default) against a full pass at every
    step: the same eigenvalues to 1e-10, true residuals below tol, fewer steps
    with a pass, and the captured graph still replays the eager launches."""
    say("batch N=8 2S=21: partial vs full reorthogonalisation")
    n, two_s = 8, 21
    lzs = list(range(0, 113, 4))
    pts = _points(4, 6)
    bp, bases, tables = _batch(n, two_s, lzs, pts)
    ev, rs, st = bp.solve(nev=3, tol=1e-10, seed=2)
    evf, rsf, stf = bp.solve(nev=3, tol=1e-10, seed=2, full_reorth=True)
    say("partial", st)
    say("full", stf)
    assert st["converged"] == st["problems"] and stf["converged"] == stf["problems"]
    assert st["max_residual"] <= 1e-10
    ok = ~np.isnan(ev)
    assert np.array_equal(ok, ~np.isnan(evf))
    assert np.max(np.abs(ev[ok] - evf[ok]) / np.maximum(1, np.abs(evf[ok]))) < 1e-10
    assert st["reorth_steps"] < stf["reorth_steps"]
    ev2, _, st2 = bp.solve(nev=3, tol=1e-10, seed=2, use_graph=False)
    assert np.array_equal(ev, ev2, equal_nan=True) and st2["steps"] == st["steps"]


@pytest.mark.parametrize("ncv", [40, 64])
def test_batch_solve_large_krylov_basis(oracle, ncv):
    """Krylov bases above 32 vectors take the 64-register form of the restart
    combination (ncv 40: its coefficient block still fits in LDS) or the form
    that reads the coefficients from global memory (ncv 64: 64 x 33 x 16
    doubles exceed the LDS): both give the dense eigenvalues."""
    say(f"batch solve N=6 2S=15 ncv {ncv}: 4 sectors x 5 points, nev=4")
    n, two_s = 6, 15
    lzs = [0, 2, 6, 12]
    pts = _points(5, 7)
    bp, bases, tables = _batch(n, two_s, lzs, pts)
    ev, rs, st = bp.solve(nev=4, tol=1e-10, seed=0, ncv=ncv)
    say(st)
    assert st["converged"] == st["problems"]
    assert st["max_residual"] <= 1e-10
    for s, b in enumerate(bases):
        bas = b.states()
        for p, t in enumerate(tables):
            ref = _oracle_full(oracle, n, two_s, bas, t)
            M = np.zeros((b.dim, b.dim))
            for i in range(b.dim):
                M[i, ref.col[ref.rowptr[i]:ref.rowptr[i + 1]]] = ref.values()[ref.rowptr[i]:ref.rowptr[i + 1]]
            want = np.linalg.eigvalsh(M)[:4]
            assert np.max(np.abs(ev[s, p] - want) / np.maximum(1, np.abs(want))) < 1e-10, (lzs[s], p)
