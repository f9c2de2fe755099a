"""A short fixed workload for hardware-counter runs of the SpMV kernels.

python -m hamiltonian_diagonalization_amd.profile_spmv [--n 12 --two-s 33 --iters 3] [--orders]
python -m hamiltonian_diagonalization_amd.profile_spmv --ofl 6 5 10   (flux lattice dim_x dim_y N, sector (0,0))

Builds the N=12, 2S=33 L_z=0 Coulomb sphere Hamiltonian with FP64, FP32 and
coded values and the tiled layout (its automatic entry width: 3-byte entries
for a real dictionary of at most 2047 values, else 4-byte; then the other
width for comparison), then launches each SpMV format, the tiled one with FP64
and with FP32 input vectors, `iters` times
(after one warm-up launch) and the plain streaming copy kernel over 1 GiB `iters` times, which
calibrates FETCH_SIZE / WRITE_SIZE against a known byte count.  Prints the
algorithmic bytes of every launch so that the counter totals can be compared.
"""
import argparse
import ctypes
import json

from . import CODED, F32, F64, STORAGE_FULL, TILED, TILED_X32, DeviceBuffer, FockBasis, Hamiltonian, Sector, _check, lib
from . import models


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=12)
    ap.add_argument("--two-s", type=int, default=33)
    ap.add_argument("--iters", type=int, default=3)
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--rows", type=int, default=0, help="rows per group of the tiled layout (0 = automatic)")
    ap.add_argument("--ofl", type=int, nargs=3, metavar=("DIM_X", "DIM_Y", "N"),
                    help="the complex flux-lattice model (seed 0, sector (0,0)) instead of the sphere")
    ap.add_argument("--orders", action="store_true",
                    help="also time the tiled product under each block order (groups per XCD window)")
    args = ap.parse_args()
    if args.ofl:
        basis, tables = models.ofl_problem(args.ofl[0], args.ofl[1], args.ofl[2], 0, 0, seed=0)
    else:
        basis = FockBasis(Sector.lz(0, args.two_s), args.n, args.two_s + 1)
        tables = models.sphere_tables(args.two_s, models.coulomb_list(args.two_s))
    ham = Hamiltonian(basis, tables, storage=STORAGE_FULL, values=(F64, F32, CODED))
    tile = ham.tile(args.rows)
    ham.set_lanes(0)
    out = {"dim": basis.dim, "nnz": ham.nnz, "lanes": ham.lanes(), "iters": args.iters, "tile": tile, "spmv": {}}
    for name, p in (("f64", F64), ("f32", F32), ("coded", CODED), ("tiled", TILED), ("tiled_x32", TILED_X32)):
        ms = ham.spmv_time_ms(p, args.iters)
        out["spmv"][name] = {"ms": ms, "alg_bytes": ham.spmv_bytes(p)}
        print(f"[profile] spmv {name}: {ms:.3f} ms, {ham.spmv_bytes(p)} algorithmic bytes", flush=True)
    if args.orders:
        chosen = tile["block_window"]
        out["orders"] = {}
        for w in (0, 16, 64, 256, (tile["groups"] + 7) // 8):
            ham.tile_order(w)
            ms = ham.spmv_time_ms(TILED, 20)
            out["orders"][w] = ms
            print(f"[profile] tiled block order {w} groups per XCD: {ms:.3f} ms", flush=True)
        ham.tile_order(chosen)
        print(f"[profile] autotuned order: {chosen}", flush=True)
    # the other entry width of the tiled layout (4-byte entries always build; 3-byte ones need a
    # real dictionary of at most 2047 values)
    if not args.ofl:
        other = 4 if tile["entry_bytes"] == 3 else 3
        try:
            t2 = ham.tile(args.rows, other)
        except Exception as e:
            print(f"[profile] {other}-byte entries not available: {e}", flush=True)
        else:
            ms = ham.spmv_time_ms(TILED, args.iters)
            out["spmv"][f"tiled{other}"] = {"ms": ms, "alg_bytes": t2["spmv_bytes"]}
            print(f"[profile] spmv tiled ({other}-byte entries): {ms:.3f} ms, {t2['spmv_bytes']} algorithmic bytes",
                  flush=True)
    src, dst = DeviceBuffer(args.copy_bytes), DeviceBuffer(args.copy_bytes)
    v = ctypes.c_double()
    _check(lib().hd_stream_copy(src.ptr, dst.ptr, args.copy_bytes, args.iters, ctypes.byref(v)))
    out["copy"] = {"ms": v.value, "bytes_read": args.copy_bytes, "bytes_written": args.copy_bytes}
    print(f"[profile] copy {args.copy_bytes} B: {v.value:.3f} ms "
          f"({2 * args.copy_bytes / v.value / 1e6:.0f} GB/s read+write)", flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
