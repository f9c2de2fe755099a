"""Flux-lattice-style complex model in one momentum sector: build and solve.

python -m hamiltonian_diagonalization_amd.ofl_solve [--dim-x 6 --dim-y 5 --n 10 --nev 6]

Single band on a dim_x x dim_y k-grid (k = kx*dim_y + ky), N fermions in the
total-momentum sector (kx, ky); band energies U(-0.1, 0.1) and complex Gaussian
momentum-conserving V, antisymmetrised and Hermitised (models.ofl_tables,
seed 0).  The Hermitian CRS is built with complex FP64, complex FP32 and coded
values, plus the tiled layout of the coded values; the lowest nev eigenvalues
are solved with each operator (the FP32 one with FP64 recurrence and FP64
refinement) and compared.  Prints one JSON line.

--use-hash: the tables as hash lists (ConvertTableFormat,
interacting_ofl_model.cpp:597-619) through the general device build (FP64
and FP32 operators).  --basis 1: the hybrid Wannier basis (ChangeToWannierBasis,
:728-923) in the k_y sector --ky, with the change-of-basis matrix of
GenerateWannierCoefficients from synthetic unit-modulus links A_x (seed 0),
offset --kx-shift, matrix-element tolerance --tol-vkkkk.
"""
import argparse
import json
import time

import numpy as np

from . import (CODED, F32, F64, LAYOUT_OFL, TILED, FockBasis, Hamiltonian, Sector, TermTables,
               hash_lists_from_array, models, ofl_wannier)


def log(*a):
    print("[ofl_solve]", *a, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--dim-x", type=int, default=6)
    ap.add_argument("--dim-y", type=int, default=5)
    ap.add_argument("--n", type=int, default=10)
    ap.add_argument("--kx", type=int, default=0)
    ap.add_argument("--ky", type=int, default=0)
    ap.add_argument("--nev", type=int, default=6)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--ncv", type=int, default=0)
    ap.add_argument("--use-hash", type=int, default=0)
    ap.add_argument("--basis", type=int, default=0)
    ap.add_argument("--kx-shift", type=float, default=0.0)
    ap.add_argument("--tol-vkkkk", type=float, default=1e-12)
    args = ap.parse_args()
    X, Y = args.dim_x, args.dim_y
    K = X * Y
    general = bool(args.use_hash or args.basis == 1)
    t0 = time.perf_counter()
    t_tables = 0.0
    if general:
        kt, vt, eps = models.ofl_tables(X, Y, seed=0)
        if args.basis == 1:
            links = np.exp(2j * np.pi * np.random.default_rng(0).random((X, Y)))
            u = models.wannier_matrix(X, Y, args.kx_shift, links)
            lists, t_tables = ofl_wannier(X, Y, args.kx_shift, kt, vt, eps, u, args.tol_vkkkk, timing=True)
            log(f"Wannier tables: quartic transform {t_tables:.2f} ms on the device")
            sector, label = Sector.k1d(args.ky, Y), f"Wannier basis, k_y sector {args.ky}"
        else:
            lists = hash_lists_from_array(LAYOUT_OFL, K, kt, vt, eps)
            sector, label = Sector.k2d(args.kx, args.ky, X, Y), f"hash tables, sector ({args.kx},{args.ky})"
        basis, tables = FockBasis(sector, args.n, K), TermTables.hash(K, **lists)
    else:
        basis, tables = models.ofl_problem(X, Y, args.n, args.kx, args.ky, seed=0)
        label = f"sector ({args.kx},{args.ky})"
    t_basis = time.perf_counter() - t0
    log(f"{X}x{Y} N={args.n} {label}: dim {basis.dim}")
    t0 = time.perf_counter()
    ham = Hamiltonian(basis, tables, values=(F64, F32) if general else (F64, F32, CODED))
    t_build = time.perf_counter() - t0
    info = ham.info()
    log(f"build {t_build:.3f} s, nnz {info['nnz']}, max row {info['max_row']}, trimmed {info['n_trimmed']}")
    ham.set_lanes(0)
    tile, t_tile = None, 0.0
    fmts = {"c64": F64, "c32": F32}
    if not general:
        t0 = time.perf_counter()
        tile = ham.tile()
        t_tile = time.perf_counter() - t0
        log(f"tiled layout {t_tile:.3f} s: {tile['rows_per_group']}-row groups, block order {tile['block_window']}, "
            f"{tile['lines'] / tile['entries']:.4f} staged lines per entry")
        fmts.update(coded=CODED, tiled=TILED)
    sp = {name: ham.spmv_time_ms(p, 10) for name, p in fmts.items()}
    for k, v in sp.items():
        log(f"SpMV {k}: {v:.3f} ms ({ham.spmv_bytes(fmts[k]) / v / 1e6:.0f} GB/s)")
    res = {}
    ref = None
    solves = (("fp64", F64), ("c32_mixed", F32)) if general else \
        (("fp64", F64), ("coded_fp64", CODED), ("tiled_fp64", TILED), ("c32_mixed", F32))
    for name, p in solves:
        r = ham.lanczos(nev=args.nev, tol=args.tol, precision=p, seed=0, ncv=args.ncv, raise_on_fail=False)
        if ref is None:
            ref = r.evals
        res[name] = {"s": round(r.time_ms / 1e3, 4), "matvecs": r.matvecs, "matvecs_f32": r.matvecs_f32,
                     "max_residual": float(r.residuals.max()), "converged": r.converged,
                     "max_abs_dE_vs_fp64": float(np.max(np.abs(r.evals - ref)))}
        log(f"{name}: {r}")
    out = {
        "workload": f"flux lattice {X}x{Y}, N={args.n}, {label}, "
                    f"{args.nev} lowest eigenvalues, residual {args.tol:g}",
        "dim": basis.dim, "nnz_full": info["nnz"], "basis_s": round(t_basis, 3), "build_s": round(t_build, 3),
        "tables_ms": round(t_tables, 2),
        "build_ms": {k: round(info[k], 2) for k in ("t_count_ms", "t_scan_ms", "t_fill_ms", "t_sort_ms")},
        "tile_s": round(t_tile, 3), "tile": tile,
        "spmv_ms": {k: round(v, 4) for k, v in sp.items()},
        "spmv_gbs": {k: round(ham.spmv_bytes(fmts[k]) / v / 1e6, 1) for k, v in sp.items()},
        "solves": res, "eigenvalues": [float(e) for e in ref],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
