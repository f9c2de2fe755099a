"""Train the amplitude network on plain-solver ground states and evaluate the
learned start vector against the random start.

python -m hamiltonian_diagonalization_amd.train_start --n 12 --two-s 33 \
    --train-points 128 --eval-points 4 --out <checkpoint.npz>

Training points (V1, V3, V5) are uniform in [0.5,1.5] x [0,0.5] x [0,0.2]
(seed 0), visited in greedy nearest-neighbour order so each ground-state solve
warm-starts from the previous point's vector; held-out points use seed 1.
For each held-out point the 5 lowest eigenpairs are solved to residual 1e-10
from (a) the random-normal start (seed 0), (b) the network's start vector and
(c) the network's start vector plus a random admixture, with the FP64
operator; (d) the network start with the FP32-value mixed solve.  Results go to
<checkpoint name>_report.json in the working directory (or --report).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import CODED, F32, F64, TILED, DeviceBuffer, FockBasis, Hamiltonian, Sector, lib
from . import learned, models


def log(*a):
    print("[train_start]", *a, flush=True)


def nn_order(points: np.ndarray) -> np.ndarray:
    scale = np.array([1.0, 2.0, 5.0])  # ranges normalised to 1
    left = list(range(len(points)))
    order = [left.pop(0)]
    while left:
        p = points[order[-1]] * scale
        d = [np.sum((points[i] * scale - p) ** 2) for i in left]
        order.append(left.pop(int(np.argmin(d))))
    return np.array(order)


def hamiltonian(basis, two_s, v):
    """Coded values and the tiled layout (the solver's fastest exact operator)."""
    vm = models.coulomb_list(two_s, {1: float(v[0]), 3: float(v[1]), 5: float(v[2])})
    h = Hamiltonian(basis, models.sphere_tables(two_s, vm), values=(CODED,))
    h.tile()
    return h


def evaluate(net, basis, two_s, v, nev, tol, mix=0.3):
    h = hamiltonian(basis, two_s, v)
    out = {"params": [float(x) for x in v]}
    plain = h.lanczos(nev=nev, tol=tol, seed=0, precision=TILED, ncv=48)
    vecs = [h.eigenvector(k) for k in range(nev)]
    psi0 = vecs[0]
    out["plain"] = {"matvecs": plain.matvecs, "ms": plain.time_ms, "evals": plain.evals.tolist(),
                    "max_resid": float(plain.residuals.max())}
    t0 = time.perf_counter()
    start = net.start_vector(basis, v)
    lib().hd_synchronize()
    t_net = (time.perf_counter() - t0) * 1e3
    vec = start.to_array(np.float64, basis.dim)
    out["overlap"] = float(abs(vec @ psi0) / (np.linalg.norm(vec) * np.linalg.norm(psi0)))
    # overlap of the start vector with each wanted eigenvector, and of each
    # predicted amplitude with its own eigenvector
    out["overlaps_wanted"] = [float(abs(vec @ x) / np.linalg.norm(vec)) for x in vecs]
    out["amplitude_overlaps"] = learned.amplitude_overlaps(net, basis, v, vecs[:net.n_states])
    out["net_eval_ms"] = t_net
    # diagnostics: |psi0|^2-weighted sign agreement (up to the global sign) and
    # the overlap of the magnitudes alone
    w = psi0 ** 2 / np.sum(psi0 ** 2)
    agree = float(np.sum(w[np.sign(vec) == np.sign(psi0)]))
    out["sign_agreement"] = max(agree, 1.0 - agree)
    out["magnitude_overlap"] = float(np.abs(vec) @ np.abs(psi0) / (np.linalg.norm(vec) * np.linalg.norm(psi0)))
    runs = {"learned": (vec, TILED), "learned_mix": (None, TILED), "learned_f32": (vec, F32), "mixed": ("rnd", F32)}
    rnd = np.random.default_rng(0).normal(size=basis.dim)
    for name, (s, prec) in runs.items():
        if s is None:
            s = vec / np.linalg.norm(vec) + mix * rnd / np.linalg.norm(rnd)
        if isinstance(s, str):
            r = h.lanczos(nev=nev, tol=tol, precision=prec, seed=0, ncv=48)
            extra = 0.0
        else:
            r = h.lanczos(nev=nev, tol=tol, precision=prec, start=s, ncv=48)
            extra = t_net
        out[name] = {"matvecs": r.matvecs, "ms": r.time_ms + extra, "evals": r.evals.tolist(),
                     "max_resid": float(r.residuals.max()),
                     "max_abs_dE_vs_plain": float(np.max(np.abs(r.evals - plain.evals)))}
    # the ground state alone, where a start close to psi0 can pay off
    p1 = h.lanczos(nev=1, tol=tol, seed=0, precision=TILED)
    l1 = h.lanczos(nev=1, tol=tol, start=vec, precision=TILED)
    out["ground_only"] = {"plain_matvecs": p1.matvecs, "plain_ms": p1.time_ms, "learned_matvecs": l1.matvecs,
                          "learned_ms": l1.time_ms + t_net, "abs_dE": float(abs(l1.evals[0] - p1.evals[0]))}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=12)
    ap.add_argument("--two-s", type=int, default=33)
    ap.add_argument("--train-points", type=int, default=128)
    ap.add_argument("--per-point", type=int, default=32768)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-points", type=int, default=4)
    ap.add_argument("--states", type=int, default=1,
                    help="amplitudes the network predicts (trained on that many lowest eigenvectors)")
    ap.add_argument("--eval-only", action="store_true")
    ap.add_argument("--nev", type=int, default=5)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--out", required=True, help="checkpoint to write (or read with --eval-only)")
    ap.add_argument("--report", default=None,
                    help="JSON report path (default: <checkpoint name>_report.json in the working directory)")
    args = ap.parse_args()

    basis = FockBasis(Sector.lz(0, args.two_s), args.n, args.two_s + 1)
    log(f"N={args.n} 2S={args.two_s} L_z=0 dim={basis.dim}")
    report = {"n": args.n, "two_s": args.two_s, "dim": basis.dim}
    if args.eval_only:
        net = learned.AmplitudeNet.load(args.out)
    else:
        pts = learned.sample_params(np.random.default_rng(0), args.train_points)
        pts = pts[nn_order(pts)]
        t0 = time.perf_counter()
        data = learned.make_dataset(basis, args.two_s, pts, args.per_point, seed=0, tol=1e-8, log=log,
                                    n_states=args.states, tiled=True)
        t_data = time.perf_counter() - t0
        log(f"dataset: {len(data.states)} samples in {t_data:.1f} s")
        net = learned.AmplitudeNet(args.two_s + 1, 3, seed=0, n_states=args.states)
        t0 = time.perf_counter()
        losses = learned.train(net, data, steps=args.steps, batch=args.batch, lr=args.lr, log=log,
                               log_every=max(200, args.steps // 20))
        t_train = time.perf_counter() - t0
        log(f"trained {args.steps} steps in {t_train:.1f} s; loss {losses[0]:.4f} -> {losses[-1]:.4f}")
        net.save(args.out)
        report.update({"train_points": args.train_points, "per_point": args.per_point, "steps": args.steps,
                       "batch": args.batch, "dataset_s": t_data, "train_s": t_train,
                       "loss_first": losses[0], "loss_last": losses[-1],
                       "loss_curve": [float(x) for x in losses[:: max(1, len(losses) // 50)]]})
    ev = learned.sample_params(np.random.default_rng(1), args.eval_points)
    report["eval"] = []
    for i, v in enumerate(ev):
        r = evaluate(net, basis, args.two_s, v, args.nev, args.tol)
        g = r["ground_only"]
        log(f"held-out {i}: ground state only: matvecs plain {g['plain_matvecs']} learned {g['learned_matvecs']}")
        log(f"held-out {i}: overlap {r['overlap']:.4f} (signs {r['sign_agreement']:.4f}, magnitudes "
            f"{r['magnitude_overlap']:.4f}); matvecs plain {r['plain']['matvecs']} "
            f"learned {r['learned']['matvecs']} mix {r['learned_mix']['matvecs']} "
            f"f32 {r['learned_f32']['matvecs']}; ms plain {r['plain']['ms']:.0f} learned "
            f"{r['learned']['ms']:.0f} mix {r['learned_mix']['ms']:.0f} f32 {r['learned_f32']['ms']:.0f}")
        log(f"held-out {i}: start-vector overlap with the {args.nev} wanted eigenvectors "
            f"{np.round(r['overlaps_wanted'], 3).tolist()}, amplitude overlaps {np.round(r['amplitude_overlaps'], 3).tolist()}; "
            f"mixed (random start) {r['mixed']['matvecs']} matvecs {r['mixed']['ms']:.0f} ms")
        report["eval"].append(r)
    ev = report["eval"]
    report["summary"] = {
        "held_out_points": len(ev),
        "mean_overlap": float(np.mean([r["overlap"] for r in ev])),
        "nev_matvecs_plain_mean": float(np.mean([r["plain"]["matvecs"] for r in ev])),
        "nev_matvecs_learned_mean": float(np.mean([r["learned"]["matvecs"] for r in ev])),
        "nev_ms_plain_mean": float(np.mean([r["plain"]["ms"] for r in ev])),
        "nev_ms_learned_mean": float(np.mean([r["learned"]["ms"] for r in ev])),
        "ground_matvecs_plain_mean": float(np.mean([r["ground_only"]["plain_matvecs"] for r in ev])),
        "ground_matvecs_learned_mean": float(np.mean([r["ground_only"]["learned_matvecs"] for r in ev])),
        "ground_ms_plain_mean": float(np.mean([r["ground_only"]["plain_ms"] for r in ev])),
        "ground_ms_learned_mean": float(np.mean([r["ground_only"]["learned_ms"] for r in ev])),
        "nev_matvecs_learned_f32_mean": float(np.mean([r["learned_f32"]["matvecs"] for r in ev])),
        "nev_ms_learned_f32_mean": float(np.mean([r["learned_f32"]["ms"] for r in ev])),
        "nev_ms_mixed_mean": float(np.mean([r["mixed"]["ms"] for r in ev])),
        "mean_overlaps_wanted": np.mean([r["overlaps_wanted"] for r in ev], axis=0).tolist(),
        "mean_amplitude_overlaps": np.mean([r["amplitude_overlaps"] for r in ev], axis=0).tolist(),
        "fewer_matvecs_points": int(sum(r["learned"]["matvecs"] < r["plain"]["matvecs"] for r in ev)),
        "max_abs_dE": float(max(max(r["learned"]["max_abs_dE_vs_plain"], r["learned_f32"]["max_abs_dE_vs_plain"],
                                    r["ground_only"]["abs_dE"]) for r in ev)),
    }
    log("summary", report["summary"])
    stem = os.path.basename(args.out)
    path = args.report or ((stem[:-4] if stem.endswith(".npz") else stem) + "_report.json")
    with open(path, "w") as f:
        json.dump(report, f, indent=1)
    log(f"wrote {path}")


if __name__ == "__main__":
    sys.exit(main())
