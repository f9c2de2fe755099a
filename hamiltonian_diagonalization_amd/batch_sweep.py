"""All L_z sectors x a grid of pseudopotentials in one batched solve.

python -m hamiltonian_diagonalization_amd.batch_sweep [--n 10 --two-s 27 --grid 8 --nev 3]

Sphere, N electrons at flux 2S, every non-empty sector lzTot = 0, 2, ...,
N(Norb-N) (the reference program's sector list, odd-parity sectors being empty),
times a grid x grid sweep of V1 in [0.5, 1.5] and V3 in [0, 0.5] with the other
odd V_m from the Coulomb formula.  Prints one JSON line with the build and
solve times, the batched SpMM time, the solver statistics, and a spot check
of a few problems against the single-problem Lanczos.
"""
import argparse
import json
import time

import numpy as np

from . import CODED, F64, BatchProblem, FockBasis, Hamiltonian, Sector, models


def log(*a):
    print("[batch_sweep]", *a, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=10)
    ap.add_argument("--two-s", type=int, default=27)
    ap.add_argument("--grid", type=int, default=8)
    ap.add_argument("--nev", type=int, default=3)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--ncv", type=int, default=0, help="Krylov basis size (0: 2 nev + 21, the reference's NCV)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--check", type=int, default=3, help="sectors spot-checked against single solves")
    ap.add_argument("--spmm-only", action="store_true", help="stop after timing the batched SpMM (for profiling)")
    args = ap.parse_args()
    n, two_s = args.n, args.two_s
    norb = two_s + 1
    lz_max = n * (norb - n)
    lzs = [lz for lz in range(0, lz_max + 1) if (lz + n * two_s) % 2 == 0]
    v1 = np.linspace(0.5, 1.5, args.grid)
    v3 = np.linspace(0.0, 0.5, args.grid)
    pts = [(a, b) for a in v1 for b in v3]
    t0 = time.perf_counter()
    bases = [FockBasis(Sector.lz(lz, two_s), n, norb) for lz in lzs]
    tables = [models.sphere_tables(two_s, models.coulomb_list(two_s, {1: a, 3: b})) for a, b in pts]
    t_tables = time.perf_counter() - t0
    log(f"{len(lzs)} sectors (dims {max(b.dim for b in bases)}..{min(b.dim for b in bases)}, "
        f"total {sum(b.dim for b in bases)}) x {len(pts)} points; bases+tables {t_tables:.1f} s")
    t0 = time.perf_counter()
    bp = BatchProblem(bases, tables)
    t_build = time.perf_counter() - t0
    log(f"batch build {t_build:.2f} s: rows {bp.rows}, nnz {bp.nnz}, {bp.ndict} value tuples")
    spmm_ms = bp.spmm_time_ms(5)
    log(f"batched SpMM over all {len(pts)} points: {spmm_ms:.2f} ms "
        f"({bp.nnz * len(pts) / spmm_ms / 1e6:.1f} G entry-products/s)")
    if args.spmm_only:
        print(json.dumps({"rows": bp.rows, "nnz": bp.nnz, "value_tuples": bp.ndict, "spmm_ms": round(spmm_ms, 3)}),
              flush=True)
        return
    t0 = time.perf_counter()
    ev, rs, st = bp.solve(nev=args.nev, tol=args.tol, ncv=args.ncv, use_graph=not args.no_graph, raise_on_fail=False)
    t_solve = time.perf_counter() - t0
    log(f"solve {t_solve:.2f} s: {st}")
    checks = []
    pick = [0, len(lzs) // 4, len(lzs) // 2][:args.check]
    for s in pick:
        for p in (0, len(pts) - 1):
            if bases[s].dim < 60:
                continue
            h = Hamiltonian(bases[s], tables[p], values=(F64, CODED))
            r = h.lanczos(nev=args.nev, tol=1e-11, seed=0, precision=CODED)
            err = float(np.max(np.abs(ev[s, p] - r.evals) / np.maximum(1, np.abs(r.evals))))
            checks.append({"lz": lzs[s], "point": p, "max_rel_err": err})
            log(f"check lz={lzs[s]} point {p}: max rel err {err:.2e}")
            del h
    out = {
        "workload": f"sphere N={n} 2S={two_s}: {len(lzs)} L_z sectors x {len(pts)} (V1,V3) points, "
                    f"{args.nev} lowest eigenvalues each, residual {args.tol:g}",
        "problems": st["problems"], "rows": bp.rows, "nnz": bp.nnz, "value_tuples": bp.ndict,
        "tables_s": round(t_tables, 3), "build_s": round(t_build, 3), "solve_s": round(t_solve, 3),
        "spmm_ms": round(spmm_ms, 3), "stats": st, "spot_checks": checks,
        "ground_energy_lz0": [float(e) for e in ev[0, :, 0][:4]],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
