"""Learned Lanczos start vector: the amplitude network, its training data
from plain-solver eigenvectors, training, and the start vector.

The network (HIP kernels in csrc/hd_mlp.hip) maps the occupation bitstring of
a sector state plus the Hamiltonian parameters (V1, V3, V5) to log|psi_s| and a
sign logit for each of its amplitudes (the lowest eigenvectors it is trained
on); evaluated over the whole sector it gives the start vector that replaces
the reference's ARPACK initial-vector file
(src/utilities/wrappers/arpack_wrapper.hpp:753-774).
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import (CODED, F64, TILED, DeviceBuffer, FockBasis, Hamiltonian, Sector, _check, _ptr, lib)
from . import models

LAYER_SHAPES = [(256, 64), (256, 256), (256, 256), (64, 256)]  # padded (rows, cols)


def unflatten(flat: np.ndarray):
    """[(W, b)] per layer from the flat FP32 layout of hd_mlp_get_weights."""
    out, off = [], 0
    for r, c in LAYER_SHAPES:
        W = flat[off:off + r * c].reshape(r, c)
        off += r * c
        b = flat[off:off + r]
        off += r
        out.append((W, b))
    return out


def flatten(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers]).astype(np.float32)


class AmplitudeNet:
    """Handle on the device network (hd_mlp_*): n_states amplitudes (the
    lowest eigenvectors it is trained on), two outputs each."""

    def __init__(self, n_bits: int, n_params: int = 3, seed: int = 0, n_states: int = 1):
        h = ctypes.c_void_p()
        _check(lib().hd_mlp_create_states(n_bits, n_params, n_states, seed, ctypes.byref(h)))
        self.h = h
        self.n_bits, self.n_params, self.n_states = n_bits, n_params, n_states
        c = ctypes.c_int64()
        _check(lib().hd_mlp_num_weights(h, ctypes.byref(c)))
        self.n_weights = c.value

    def get_weights(self) -> np.ndarray:
        w = np.zeros(self.n_weights, dtype=np.float32)
        _check(lib().hd_mlp_get_weights(self.h, _ptr(w)))
        return w

    def set_weights(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.n_weights
        _check(lib().hd_mlp_set_weights(self.h, _ptr(w)))

    def forward(self, states: np.ndarray, params: np.ndarray) -> np.ndarray:
        st = np.ascontiguousarray(states, dtype=np.uint64)
        pr = np.ascontiguousarray(params, dtype=np.float32).reshape(len(st), self.n_params)
        out = np.zeros((len(st), 2 * self.n_states), dtype=np.float32)
        _check(lib().hd_mlp_forward(self.h, _ptr(st), _ptr(pr), len(st), _ptr(out)))
        return out

    def gradients(self, states, params, targets, lam=1.0):
        st = np.ascontiguousarray(states, dtype=np.uint64)
        pr = np.ascontiguousarray(params, dtype=np.float32).reshape(len(st), self.n_params)
        tg = np.ascontiguousarray(targets, dtype=np.float32).reshape(len(st), 3 * self.n_states)
        g = np.zeros(self.n_weights, dtype=np.float32)
        loss = ctypes.c_float()
        _check(lib().hd_mlp_gradients(self.h, _ptr(st), _ptr(pr), _ptr(tg), len(st), lam,
                                      ctypes.byref(loss), _ptr(g)))
        return loss.value, g

    def train_step(self, d_states: DeviceBuffer, d_params: DeviceBuffer, d_targets: DeviceBuffer,
                   n: int, lam: float, lr: float) -> float:
        loss = ctypes.c_float()
        _check(lib().hd_mlp_train_step(self.h, d_states.ptr, d_params.ptr, d_targets.ptr, n, lam, lr,
                                       ctypes.byref(loss)))
        return loss.value

    def start_vector(self, basis: FockBasis, params: Sequence[float]) -> DeviceBuffer:
        pr = np.ascontiguousarray(params, dtype=np.float32)
        buf = DeviceBuffer(basis.dim * 8)
        _check(lib().hd_mlp_start_vector(self.h, basis.h, _ptr(pr), buf.ptr))
        return buf

    def write_initial_vector(self, basis: FockBasis, params: Sequence[float], path: str) -> np.ndarray:
        """The start vector for `params` as the raw float64[dim] file the
        reference's ArpackWrapper reads with --use-initial
        (arpack_wrapper.hpp:753-774); returns the vector."""
        from . import fileio
        v = self.start_vector(basis, params).to_array(np.float64, basis.dim)
        fileio.write_initial_vector(path, v)
        return v

    def save(self, path: str):
        np.savez(path, weights=self.get_weights(), n_bits=self.n_bits, n_params=self.n_params,
                 n_states=self.n_states)

    @classmethod
    def load(cls, path: str) -> "AmplitudeNet":
        z = np.load(path)
        n_states = int(z["n_states"]) if "n_states" in z.files else 1
        net = cls(int(z["n_bits"]), int(z["n_params"]), n_states=n_states)
        net.set_weights(z["weights"])
        return net

    def __del__(self):
        if getattr(self, "h", None) is not None:
            from . import _lib
            if _lib is not None:
                _lib.hd_mlp_destroy(self.h)
            self.h = None


def gemm_check(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """C = A . B^T through the MFMA kernel (A: M x K, B: N x K, BF16 bit patterns)."""
    M, K = A.shape
    N = B.shape[0]
    a = np.ascontiguousarray(A, dtype=np.uint16)
    b = np.ascontiguousarray(B, dtype=np.uint16)
    c = np.zeros((M, N), dtype=np.float32)
    _check(lib().hd_mlp_gemm_check(M, N, K, _ptr(a), _ptr(b), _ptr(c)))
    return c


# ------------------------------------------------------------- training data

def sample_params(rng: np.random.Generator, n: int) -> np.ndarray:
    """(V1, V3, V5) uniform in [0.5,1.5] x [0,0.5] x [0,0.2]."""
    lo = np.array([0.5, 0.0, 0.0])
    hi = np.array([1.5, 0.5, 0.2])
    return lo + (hi - lo) * rng.random((n, 3))


@dataclass
class Dataset:
    states: np.ndarray   # uint64
    params: np.ndarray   # float32 (n, n_params)
    targets: np.ndarray  # float32 (n, 3 n_states): per amplitude log|psi| - max, sign, weight
    solve_ms: float


def sphere_hamiltonian(basis: FockBasis, two_s: int, v135, tiled: bool = False) -> Hamiltonian:
    """The sphere Hamiltonian at (V1, V3, V5): FP64 CRS, or with tiled=True the
    coded values and the tiled layout (solve with precision TILED)."""
    vm = models.coulomb_list(two_s, {1: float(v135[0]), 3: float(v135[1]), 5: float(v135[2])})
    if not tiled:
        return Hamiltonian(basis, models.sphere_tables(two_s, vm), values=(F64,))
    h = Hamiltonian(basis, models.sphere_tables(two_s, vm), values=(CODED,))
    h.tile()
    return h


def ground_state_targets(psi: np.ndarray, floor: float = -30.0):
    """(log|psi| - max log|psi|, sign) with the global sign fixed so that the
    largest-magnitude component is positive.  Log-amplitudes are clamped at
    `floor` (e^-30 ~ 1e-13 of the largest component, irrelevant to a start
    vector) so that vanishing components do not dominate the squared error."""
    i = int(np.argmax(np.abs(psi)))
    if psi[i] < 0:
        psi = -psi
    a = np.log(np.maximum(np.abs(psi), 1e-300))
    a = np.maximum(a - a[i], floor)
    return a.astype(np.float32), np.where(psi >= 0, 1.0, -1.0).astype(np.float32)


def make_dataset(basis: FockBasis, two_s: int, points: np.ndarray, per_point: int, seed: int = 0,
                 tol: float = 1e-8, log=print, hamiltonian=None, n_states: int = 1, tiled: bool = False) -> Dataset:
    """Plain-solver eigenvectors (the n_states lowest) at each parameter point;
    per point, half the samples drawn by the mean of their |psi_s|^2 and half
    uniformly over the sector.  The model is the sphere at flux two_s with
    parameters (V1, V3, V5) (on the tiled operator with tiled=True), or
    hamiltonian(basis, p) when given.  Each point's solve starts from the sum
    of the previous point's eigenvectors."""
    rng = np.random.default_rng(seed)
    states_all = basis.states()
    S, P, T = [], [], []
    t_solve = 0.0
    start = None
    for i, p in enumerate(points):
        h = hamiltonian(basis, p) if hamiltonian else sphere_hamiltonian(basis, two_s, p, tiled=tiled)
        t0 = time.perf_counter()
        r = h.lanczos(nev=n_states, tol=tol, seed=i, start=start, precision=TILED if tiled and not hamiltonian else F64)
        t_solve += time.perf_counter() - t0
        vecs = [h.eigenvector(k) for k in range(n_states)]
        start = np.sum(vecs, axis=0)  # warm start for the next point of the sweep
        cols, prob = [], np.zeros(basis.dim)
        for psi in vecs:
            a, s = ground_state_targets(psi)
            cols.append((a, s))
            prob += psi ** 2
        prob /= prob.sum()
        k = min(per_point // 2, basis.dim)
        idx = np.concatenate([rng.choice(basis.dim, size=k, p=prob),
                              rng.choice(basis.dim, size=per_point - k)])
        S.append(states_all[idx])
        P.append(np.repeat(np.asarray(p, dtype=np.float32)[None, :], len(idx), axis=0))
        T.append(np.concatenate([np.stack([a[idx], s[idx], np.ones(len(idx), dtype=np.float32)], axis=1)
                                 for a, s in cols], axis=1))
        if log:
            log(f"  point {i}: p=({','.join('%.3f' % x for x in p)}) E={np.round(r.evals[:n_states], 8).tolist()} "
                f"matvecs={r.matvecs}")
        del h
    return Dataset(np.concatenate(S), np.concatenate(P).astype(np.float32),
                   np.concatenate(T).astype(np.float32), t_solve * 1e3)


def train(net: AmplitudeNet, data: Dataset, steps: int = 2000, batch: int = 4096, lr: float = 1e-3,
          lam: float = 0.5, seed: int = 0, log=print, log_every: int = 200) -> List[float]:
    """Adam on minibatches drawn uniformly from the dataset, uploaded per step
    into fixed device buffers."""
    rng = np.random.default_rng(seed)
    n = len(data.states)
    b_st, b_pr = DeviceBuffer(batch * 8), DeviceBuffer(batch * 4 * net.n_params)
    b_tg = DeviceBuffer(batch * 12 * net.n_states)
    losses = []
    for it in range(steps):
        idx = rng.integers(0, n, size=batch)
        b_st.upload(data.states[idx])
        b_pr.upload(data.params[idx])
        b_tg.upload(data.targets[idx])
        lr_t = lr * (0.5 * (1 + np.cos(np.pi * it / steps)))  # cosine decay
        losses.append(net.train_step(b_st, b_pr, b_tg, batch, lam, float(lr_t)))
        if log and (it % log_every == 0 or it == steps - 1):
            log(f"  step {it}: loss {losses[-1]:.5f}")
    return losses


def overlap(net: AmplitudeNet, basis: FockBasis, params, psi: np.ndarray) -> float:
    """|<psi_net|psi>| / (|psi_net| |psi|)."""
    v = net.start_vector(basis, params).to_array(np.float64, basis.dim)
    return float(abs(v @ psi) / (np.linalg.norm(v) * np.linalg.norm(psi)))


def amplitude_overlaps(net: AmplitudeNet, basis: FockBasis, params, vecs) -> List[float]:
    """Overlap of each predicted amplitude s with vecs[s] (unit vectors)."""
    st = basis.states()
    out = net.forward(st, np.repeat(np.asarray(params, dtype=np.float32)[None, :], len(st), axis=0))
    res = []
    for k, psi in enumerate(vecs):
        a = np.exp((out[:, 2 * k] - out[:, 2 * k].max()).astype(np.float64)) * np.where(out[:, 2 * k + 1] >= 0, 1.0, -1.0)
        res.append(float(abs(a @ psi) / (np.linalg.norm(a) * np.linalg.norm(psi))))
    return res
