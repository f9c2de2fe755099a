"""A short fixed workload that runs every device stage of the hot path once,
for kernel-trace and hardware-counter runs.

python -m hamiltonian_diagonalization_amd.profile_stages [--n 12 --two-s 33 --iters 3]

Stages, in order, on the N=12, 2S=33, L_z=0 Coulomb sphere sector
(dim 7,764,392): the basis (k_unrank), the full-storage build with FP64,
FP32 and coded values (k_count, the scan, k_fill_sorted), the tiled layout
(k_tile_count, k_tile_fill, k_tile_pack for 3-byte entries, then its
block-order autotune), `iters` tiled products with FP64 x and with FP32 x (one
warm-up each), the amplitude network over the whole sector (k_mlp_infer and
the normalisation kernels), and the plain streaming copy over 1 GiB `iters`
times, which calibrates FETCH_SIZE / WRITE_SIZE against a known byte count.
Prints the algorithmic bytes (or flops) of every stage as one JSON line.
"""
import argparse
import ctypes
import json
import os
import time

from . import (CODED, F32, F64, STORAGE_FULL, TILED, TILED_X32, DeviceBuffer, FockBasis, Hamiltonian, Sector,
               _check, lib)
from . import models


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=12)
    ap.add_argument("--two-s", type=int, default=33)
    ap.add_argument("--iters", type=int, default=3)
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--skip-network", action="store_true")
    args = ap.parse_args()
    out = {}
    t0 = time.perf_counter()
    basis = FockBasis(Sector.lz(0, args.two_s), args.n, args.two_s + 1)
    dim = basis.dim
    out["basis"] = {"s": time.perf_counter() - t0, "dim": dim, "bytes_written": 8 * dim}
    vm = models.coulomb_list(args.two_s)
    tables = models.sphere_tables(args.two_s, vm)
    ham = Hamiltonian(basis, tables, storage=STORAGE_FULL, values=(F64, F32, CODED))
    info = ham.info()
    nnz = info["nnz"]
    out["build"] = {k: info[k] for k in ("nnz", "max_row", "t_count_ms", "t_scan_ms", "t_fill_ms", "t_sort_ms")}
    # per-stage algorithmic bytes: count reads the state and writes the row
    # count, diagonal and flag; the one-pass fill reads the state, row bounds,
    # diagonal and flag and writes column, FP64, FP32 and code per entry
    out["build"]["alg_bytes"] = {"k_count": dim * (8 + 8 + 8 + 1),
                                 "k_fill_sorted": nnz * (4 + 8 + 4 + 2) + dim * (8 + 8 + 8 + 1)}
    tile = ham.tile()
    out["tile"] = tile
    # the fill writes 4-byte entries; with 3-byte entries k_tile_pack reads
    # them and writes 3 bytes each
    out["tile"]["alg_bytes"] = {"k_tile_count": nnz * 6 + dim * 8,
                                "k_tile_fill": nnz * 6 + tile["entries"] * 4 + dim * 16 +
                                4 * (tile["lines"] + tile["chunks"] * 130),
                                "k_tile_pack": tile["entries"] * (4 + 3) if tile["entry_bytes"] == 3 else 0}
    out["spmv"] = {}
    for name, p in (("tiled", TILED), ("tiled_x32", TILED_X32)):
        ms = ham.spmv_time_ms(p, args.iters)
        out["spmv"][name] = {"ms": ms, "alg_bytes": ham.spmv_bytes(p)}
        print(f"[profile] spmv {name}: {ms:.3f} ms, {ham.spmv_bytes(p)} algorithmic bytes", flush=True)
    ckdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "checkpoints")
    ckpt = next((os.path.join(ckdir, f) for f in (f"start5_n{args.n}_2s{args.two_s}.npz",
                                                  f"start_n{args.n}_2s{args.two_s}.npz")
                 if os.path.exists(os.path.join(ckdir, f))), "")
    if not args.skip_network and ckpt:
        from . import learned
        net = learned.AmplitudeNet.load(ckpt)
        for rep in range(2):  # the second pass is the measured one
            _check(lib().hd_synchronize())
            t0 = time.perf_counter()
            net.start_vector(basis, [vm[1], vm[3], vm[5]])
            _check(lib().hd_synchronize())
            dt = time.perf_counter() - t0
        S = net.n_states
        flop = 2.0 * (37 * 256 + 2 * 256 * 256 + 256 * 2 * S) * dim
        flop_padded = 2.0 * (64 * 256 + 2 * 256 * 256 + 256 * 32) * dim  # 32-wide output tile
        out["network"] = {"s": dt, "states": dim, "amplitudes": S, "flop": flop, "flop_padded": flop_padded,
                          "checkpoint": os.path.basename(ckpt)}
        print(f"[profile] network over the sector: {dt * 1e3:.1f} ms", flush=True)
    src, dst = DeviceBuffer(args.copy_bytes), DeviceBuffer(args.copy_bytes)
    v = ctypes.c_double()
    _check(lib().hd_stream_copy(src.ptr, dst.ptr, args.copy_bytes, args.iters, ctypes.byref(v)))
    out["copy"] = {"ms": v.value, "bytes_read": args.copy_bytes, "bytes_written": args.copy_bytes}
    print(f"[profile] copy {args.copy_bytes} B: {v.value:.3f} ms", flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
