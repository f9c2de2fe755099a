"""Model inputs for the benchmark configurations (host side, no GPU).

* Sphere (FQHE in the lowest Landau level): Coulomb pseudopotentials and the
  term tables of SpherePseudopotentialModel::BuildTermTables
  (src/fqhe_sphere/pseudopotential_model/pseudopotential_model.cpp:64-95).
* Interacting flux-lattice-style model in one band: a synthetic,
  momentum-conserving complex V in the folded table layout of
  src/optical_flux_lattice/interacting_ofl_model/term_tables.cpp:95-175.
"""
from __future__ import annotations

from decimal import Decimal, getcontext
from fractions import Fraction
from math import comb
from typing import Dict, Optional, Sequence

import numpy as np

from . import (LAYOUT_OFL, LAYOUT_SPHERE, FockBasis, Sector, TermTables, sphere_ktable,
               sphere_vtable)


def coulomb_pseudopotential(two_s: int, m: int) -> float:
    """LLL Coulomb pseudopotential V_m on the sphere of flux 2S (radius sqrt(S)):
    V_m = (2/sqrt(S)) C(4S-2J, 2S-J) C(4S+2J+2, 2S+J+1) / C(4S+2, 2S+1)^2 with
    J = 2S - m (Fano, Ortolani, Colombo, PRB 34 (1986) 2670), evaluated with
    exact integers and rounded once to double.  Same formula as
    src/fqhe_sphere/interaction_pseudopotentials/coulomb_pseudopotentials.cpp:61-67."""
    J = two_s - m
    r = Fraction(comb(2 * two_s - 2 * J, two_s - J) * comb(2 * two_s + 2 * J + 2, two_s + J + 1),
                 comb(2 * two_s + 2, two_s + 1) ** 2)
    getcontext().prec = 60
    # 2/sqrt(S) = 2 sqrt(2/two_s)
    v = 2 * (Decimal(r.numerator) / Decimal(r.denominator)) * (Decimal(2) / Decimal(two_s)).sqrt()
    return float(v)


def coulomb_list(two_s: int, overrides: Optional[Dict[int, float]] = None) -> np.ndarray:
    """The V_m list for the sphere benchmarks: indices m = 0 .. 2S-2 (length
    2S-1, within the reference's limit of 2S coefficients), Coulomb values on
    odd m, zero on even m (spinless fermions see only odd m); `overrides`
    replaces selected odd m (e.g. {1: V1, 3: V3})."""
    v = np.zeros(two_s - 1, dtype=np.float64)
    for m in range(1, two_s - 1, 2):
        v[m] = coulomb_pseudopotential(two_s, m)
    for m, val in (overrides or {}).items():
        v[m] = val
    return v


def sphere_tables(two_s: int, pseudopotentials: Sequence[float]) -> TermTables:
    """Quartic tables of the sphere model (kMax = 2S + 1, no quadratic terms)."""
    kt = sphere_ktable(two_s)
    vt = sphere_vtable(two_s, pseudopotentials)
    return TermTables(LAYOUT_SPHERE, two_s + 1, kt, vt)


def sphere_problem(n: int, two_s: int, lz: int = 0, pseudopotentials=None):
    """(basis, tables) for N electrons at flux 2S in the 2*L_z = lz sector.
    The reference labels sectors by lzTot = sum(2k - 2S) = -2 L_z."""
    if pseudopotentials is None:
        pseudopotentials = coulomb_list(two_s)
    basis = FockBasis(Sector.lz(lz, two_s), n, two_s + 1)
    return basis, sphere_tables(two_s, pseudopotentials)


def two_level_tables(n1: int, pseudopotentials1=None, pseudopotentials2=None, eps1=None, eps2=None):
    """Tables of SphereTwoLevelPseudopotentialModel::BuildTermTables
    (two_level_pseudopotential_model.cpp:61-87): level 0 over n1 orbitals at
    maxLz = n1 - 1, level 1 over n1 + 2 orbitals at maxLz = n1 + 1, each from
    its own pseudopotential list (default: Coulomb at that level's flux)."""
    if pseudopotentials1 is None:
        pseudopotentials1 = coulomb_list(n1 - 1) if n1 > 2 else [0.0, 1.0]
    if pseudopotentials2 is None:
        pseudopotentials2 = coulomb_list(n1 + 1)
    kt1, vt1 = sphere_ktable(n1 - 1), sphere_vtable(n1 - 1, pseudopotentials1)
    kt2, vt2 = sphere_ktable(n1 + 1), sphere_vtable(n1 + 1, pseudopotentials2)
    return TermTables.two_level(n1, kt1, vt1, kt2, vt2, eps1, eps2)


def two_level_problem(n: int, n1: int, lz: Optional[int] = 0, **kw):
    """(basis, tables) for N electrons in the LLL (n1 orbitals) plus the second
    Landau level (n1 + 2 orbitals) in the 2 L_z = lz sector (None: every
    state), SphereTwoLevelPseudopotentialModel::BuildFockBasis (:114-165)."""
    sector = Sector.none() if lz is None else Sector.lz2(lz, n1)
    basis = FockBasis(sector, n, 2 * n1 + 2)
    return basis, two_level_tables(n1, **kw)


def ofl_tables(dim_x: int, dim_y: int, seed: int = 0, band_width: float = 0.1):
    """Synthetic single-band flux-lattice tables on a dim_x x dim_y k grid
    (k = kx*dim_y + ky):
      eps_k ~ U(-band_width, band_width);
      G(k1,k2,k3,k4) complex Gaussian (sigma = 1 per component) on every
      momentum-conserving tuple k1 = k3 + k4 - k2 (mod grid);
      V = antisymmetrised in (k1,k2) and (k3,k4), then Hermitised,
      V_1234 = conj(V_4321).
    Drawn from numpy's PCG64 generator with `seed`.  Returns
    (ktable, vtable, eps) in the folded layout."""
    rng = np.random.default_rng(seed)
    X, Y = dim_x, dim_y
    K = X * Y
    eps = rng.uniform(-band_width, band_width, size=K)
    k = np.arange(K)
    kx, ky = k // Y, k % Y
    # k1[k2, k3, k4]
    k1 = (((kx[None, :, None] + kx[None, None, :] - kx[:, None, None]) % X) * Y
          + (ky[None, :, None] + ky[None, None, :] - ky[:, None, None]) % Y)
    G = rng.normal(size=(K, K, K)) + 1j * rng.normal(size=(K, K, K))  # indexed [k2, k3, k4]
    i2, i3, i4 = np.meshgrid(k, k, k, indexing="ij")
    i1 = k1
    # A(1,2,3,4) = [G(1,2,3,4) - G(2,1,3,4) - G(1,2,4,3) + G(2,1,4,3)] / 4
    A = (G[i2, i3, i4] - G[i1, i3, i4] - G[i2, i4, i3] + G[i1, i4, i3]) / 4
    # V(1,2,3,4) = [A(1,2,3,4) + conj(A(4,3,2,1))] / 2, A(4,3,2,1) sits at [k3, k2, k1]
    V = (A + np.conj(A[i3, i2, i1])) / 2
    n = (K + 1) * K * K // 2
    kt = np.zeros(n, dtype=np.uint16)
    vt = np.zeros(n, dtype=np.complex128)
    for t1 in range(K):
        base = (t1 * K - t1 * (t1 + 1) // 2)
        for t2 in range(t1, K):
            idx = (base + t2) * K + np.arange(K)
            kt[idx] = k1[:, t2, t1]
            vt[idx] = V[:, t2, t1]
    return kt, vt, eps.astype(np.complex128)


def ofl_problem(dim_x: int, dim_y: int, n: int, kx: int = 0, ky: int = 0, seed: int = 0):
    kt, vt, eps = ofl_tables(dim_x, dim_y, seed)
    K = dim_x * dim_y
    basis = FockBasis(Sector.k2d(kx, ky, dim_x, dim_y), n, K)
    return basis, TermTables(LAYOUT_OFL, K, kt, vt, eps)


def lattice_tables(dim_x: int, dim_y: int):
    """StateAddition, StateSubtraction and InverseR60 of the flux-lattice k grid
    (src/optical_flux_lattice/interacting_ofl_model/linear_momentum_constraints_2D.cpp:128-264,
    one unit cell) as tables: add[k, q] = k + q, sub[k, q] = k - q (component-wise
    mod the grid, k = kx*dim_y + ky) and inv_r60[k, alpha] = InverseR60(k, alpha)."""
    K = dim_x * dim_y
    kx, ky = np.divmod(np.arange(K), dim_y)
    add = ((kx[:, None] + kx[None, :]) % dim_x) * dim_y + (ky[:, None] + ky[None, :]) % dim_y
    sub = ((kx[:, None] - kx[None, :]) % dim_x) * dim_y + (ky[:, None] - ky[None, :]) % dim_y
    rot = [(kx, ky), (-ky, kx + ky), (-kx - ky, kx), (-kx, -ky), (ky, -kx - ky), (kx + ky, -kx)]
    inv = np.stack([(rx % dim_x) * dim_y + ry % dim_y for rx, ry in rot], axis=1)
    return add.astype(np.uint16), sub.astype(np.uint16), inv.astype(np.uint16)



def random_hash_tables(kmax: int, seed: int = 0, complex_values: bool = True, max_k1: int = 3,
                       quadratic: bool = True, fill: float = 0.5) -> dict:
    """Synthetic general term tables in the list format of TermTables.hash,
    drawn from numpy's PCG64 generator with `seed`: each (k2,k3,k4) entry
    lists 0..max_k1 labels k1 (with probability `fill` per slot, repeats and
    k1 in {k2,k3,k4} allowed) with normal (complex normal) values; with
    `quadratic`, each k2 lists 0..max_k1 labels (k1 = k2 among them half the
    time).  No symmetry: a stress input for the general build."""
    rng = np.random.default_rng(seed)
    K = kmax
    n = K ** 3

    def values(shape):
        v = rng.normal(size=shape)
        return v + 1j * rng.normal(size=shape) if complex_values else v

    nk1 = (rng.random((n, max_k1)) < fill).sum(axis=1).astype(np.uint8)
    k1 = rng.integers(0, K, size=(n, max_k1)).astype(np.uint16)
    v = values((n, max_k1))
    out = {"nk1": nk1, "k1": k1, "v": v}
    if quadratic:
        nq1 = (rng.random((K, max_k1)) < fill).sum(axis=1).astype(np.uint8)
        q1 = rng.integers(0, K, size=(K, max_k1)).astype(np.uint16)
        diag = rng.random(K) < 0.5
        q1[diag, 0] = np.arange(K)[diag]
        nq1[diag] = np.maximum(nq1[diag], 1)
        out.update(nq1=nq1, q1=q1, e=values((K, max_k1)))
    return out


def berry_links(dim_x: int, dim_y: int, bloch) -> np.ndarray:
    """A_x(kx, ky) of NonInteractingOflModelGrid::GenerateWannierCoefficients
    (noninteracting_ofl_model_grid.cpp:799-814): the normalised overlap
    <u(kx,ky)|u(kx+1,ky)> (zdotc) of the band-0 Bloch vectors, kx+1 periodic;
    bloch has one vector per k = kx*dim_y + ky.  Returns a (dim_x, dim_y)
    array of unit-modulus values."""
    b = np.asarray(bloch, dtype=np.complex128).reshape(dim_x * dim_y, -1)
    A = np.zeros((dim_x, dim_y), dtype=np.complex128)
    for kx in range(dim_x):
        for ky in range(dim_y):
            a = np.vdot(b[kx * dim_y + ky], b[((kx + 1) % dim_x) * dim_y + ky])
            A[kx, ky] = a / abs(a)
    return A


def wannier_matrix(dim_x: int, dim_y: int, offset_x: float, links) -> np.ndarray:
    """The change-of-basis matrix of GenerateWannierCoefficients
    (noninteracting_ofl_model_grid.cpp:815-925) from the links A_x(kx, ky):
    lambda(ky) = (prod_kx A_x)^(1/dim_x) (principal branch),
    aProd(kx, ky) = prod_{kappa < kx} A_x(kappa, ky), and
    M[(x,ky), (kx,ky)] = exp(-2 pi i (kx + offset_x) x / dim_x) lambda(ky)^kx
    / (sqrt(dim_x) aProd(kx, ky)), zero between different ky.  Unitary."""
    A = np.asarray(links, dtype=np.complex128).reshape(dim_x, dim_y)
    lam = np.prod(A, axis=0) ** (1.0 / dim_x)
    aprod = np.ones((dim_x, dim_y), dtype=np.complex128)
    for kx in range(1, dim_x):
        aprod[kx] = A[kx - 1] * aprod[kx - 1]
    K = dim_x * dim_y
    M = np.zeros((K, K), dtype=np.complex128)
    for x in range(dim_x):
        for ky in range(dim_y):
            for kx in range(dim_x):
                M[x * dim_y + ky, kx * dim_y + ky] = (np.exp(-2j * np.pi * (kx + offset_x) * x / dim_x)
                                                      * lam[ky] ** kx / (np.sqrt(dim_x) * aprod[kx, ky]))
    return M
