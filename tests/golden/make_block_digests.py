"""Writes tests/golden/block_digests.json: SHA-256 digests of the whole
reference-layout (upper-triangle) CRS of the two large benchmark sectors,
computed with the CPU oracle in streamed row blocks.

python tests/golden/make_block_digests.py [--workers 7]

Cases:
  sphere N=12, 2S=33, L_z=0, Coulomb pseudopotentials (dim 7,764,392);
  flux lattice 6x5, N=10, K=(0,0), synthetic tables with seed 0 (dim 1,001,402).

The rows are cut into blocks of BLOCK_ROWS rows.  A block's digest is the
SHA-256 of its row lengths (int64), then its columns (int32), then its values
(float64, or complex128 as interleaved float64 pairs).  The case digest is the
SHA-256 of the concatenated hexadecimal block digests, in row order, so a test
can stream the device CRS block by block and compare without holding either
matrix in host memory.  The oracle runs one block per worker process.
"""
import argparse
import hashlib
import json
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

BLOCK_ROWS = 65536
_ctx = {}


def block_digest(rowptr, col, val):
    h = hashlib.sha256()
    h.update(np.diff(np.asarray(rowptr, dtype=np.int64)).astype(np.int64).tobytes())
    h.update(np.ascontiguousarray(col, dtype=np.int32).tobytes())
    h.update(np.ascontiguousarray(val).tobytes())
    return h.hexdigest()


def case_digest(block_hex):
    return hashlib.sha256("".join(block_hex).encode()).hexdigest()


def _inputs(name):
    from oracle import oracle as O
    from hamiltonian_diagonalization_amd import models
    if name == "sphere":
        bas = O.basis(O.sector(O.SECTOR_LZ, 0, 33), 12, 34)
        kt = O.sphere_ktable(33)
        vt = O.sphere_vtable(33, models.coulomb_list(33))
        return dict(N=12, Norb=34, bas=bas, layout=O.LAYOUT_SPHERE, K=34, ktab=kt, vtab=vt, eps=None,
                    is_complex=False)
    kt, vt, eps = models.ofl_tables(6, 5, 0)
    bas = O.basis(O.sector(O.SECTOR_K2D, 0, 0, 6, 5), 10, 30)
    return dict(N=10, Norb=30, bas=bas, layout=O.LAYOUT_OFL, K=30, ktab=kt, vtab=vt, eps=eps, is_complex=True)


def _init(name):
    _ctx.update(_inputs(name))


def _work(b):
    from oracle import oracle as O
    c = _ctx
    dim = len(c["bas"])
    r0, r1 = b * BLOCK_ROWS, min((b + 1) * BLOCK_ROWS, dim)
    crs = O.build(c["N"], c["Norb"], c["bas"], c["layout"], c["K"], c["ktab"], c["vtab"], eps=c["eps"],
                  is_complex=c["is_complex"], row0=r0, row1=r1)
    return b, block_digest(crs.rowptr, crs.col, crs.values()), int(crs.nnz), int(crs.n_trimmed)


def run_case(name, workers):
    inp = _inputs(name)
    dim = len(inp["bas"])
    nb = (dim + BLOCK_ROWS - 1) // BLOCK_ROWS
    t0 = time.time()
    out = [None] * nb
    with Pool(workers, initializer=_init, initargs=(name,)) as pool:
        for b, hx, nnz, ntr in pool.imap_unordered(_work, range(nb)):
            out[b] = (hx, nnz, ntr)
            print(f"[block_digests] {name} block {b + 1}/{nb} ({time.time() - t0:.0f} s)", flush=True)
    return {"dim": dim, "block_rows": BLOCK_ROWS, "blocks": nb, "nnz": sum(o[1] for o in out),
            "n_trimmed": sum(o[2] for o in out), "digest": case_digest([o[0] for o in out]),
            "block_digests": [o[0] for o in out]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=max(1, (os.cpu_count() or 2) - 1))
    ap.add_argument("--cases", default="ofl,sphere")
    args = ap.parse_args()
    path = os.path.join(HERE, "block_digests.json")
    res = json.load(open(path)) if os.path.exists(path) else {}
    res["generator"] = "tests/golden/make_block_digests.py (CPU oracle, upper-triangle reference CRS)"
    names = {"sphere": "sphere N=12 2S=33 L_z=0 Coulomb", "ofl": "flux lattice 6x5 N=10 (0,0) seed 0"}
    for key in args.cases.split(","):
        res[key] = dict(name=names[key], **run_case(key, args.workers))
        with open(path, "w") as f:
            json.dump(res, f, indent=1)
        print(f"[block_digests] {key}: nnz {res[key]['nnz']} digest {res[key]['digest']}", flush=True)


if __name__ == "__main__":
    main()
