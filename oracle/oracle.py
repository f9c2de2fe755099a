"""ctypes wrapper around the CPU oracle (oracle/hd_oracle.c) plus the exact
restatements that are easier in Python (Clebsch-Gordan products with exact
rationals, the sphere V table, Jordan-Wigner brute force).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

Reference files restated here (paths relative to the reference repository):
  src/fqhe_sphere/pseudopotential_model/pseudopotential_model_base.hpp:150-229
      BaseTermsFromPseudopotentials (V table from pseudopotentials)
  src/utilities/wrappers/clebsch_gordan_wrapper.hpp:93-173
      CG lookup; coefficients with |CG| <= 1e-12 are not stored (read as 0)
  src/utilities/mathematics/fermions.hpp:52-227  (Jordan-Wigner brute force
      is the textbook definition the reference's sign rule must reproduce)

Pinning: the exact-rational CG values are pinned by the orthonormality of the
CG matrix and by closed forms (<j j; j -j | 0 0> = (-1)^(j-m)/sqrt(2j+1));
see tests/test_oracle_tables.py.  The reference's own CG routine
(clebsch_gordan_coefficients.cpp:357-784) is a numerical SU(N) algorithm that
agrees with the exact values only to ~1e-13, so the V table is "parity
unpinned" at the bit level against the reference and pinned to 1e-12 relative.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from decimal import Decimal, getcontext
from fractions import Fraction
from functools import lru_cache
from math import factorial

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libhd_oracle.so")

SECTOR_NONE, SECTOR_LZ, SECTOR_K2D, SECTOR_K1D = 0, 1, 2, 3
LAYOUT_SPHERE, LAYOUT_OFL = 0, 1
ZERO_TOL = 1e-14  # _SPARSE_ZERO_TOL_ (hamiltonian_data.hpp:54)


class _Sector(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("a", ctypes.c_int), ("b", ctypes.c_int),
                ("c", ctypes.c_int), ("d", ctypes.c_int)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u64, i64, i32, dbl = ctypes.c_uint64, ctypes.c_int64, ctypes.c_int, ctypes.c_double
        P = ctypes.c_void_p
        L.oc_binomial.restype = u64
        L.oc_binomial.argtypes = [i32, i32]
        L.oc_next_hamming.restype = u64
        L.oc_next_hamming.argtypes = [u64]
        L.oc_generate_hamming.restype = u64
        L.oc_generate_hamming.argtypes = [i32, u64]
        L.oc_index_hamming.restype = u64
        L.oc_index_hamming.argtypes = [i32, u64]
        L.oc_sector_test.restype = i32
        L.oc_sector_test.argtypes = [ctypes.POINTER(_Sector), u64]
        L.oc_basis.restype = i64
        L.oc_basis.argtypes = [ctypes.POINTER(_Sector), i32, i32, P, i64]
        L.oc_find_index.restype = i64
        L.oc_find_index.argtypes = [P, i64, u64]
        L.oc_cdcdcc.restype = u64
        L.oc_cdcdcc.argtypes = [ctypes.POINTER(i32), u64, u64, u64, u64, u64, u64]
        L.oc_cdc.restype = u64
        L.oc_cdc.argtypes = [ctypes.POINTER(i32), u64, u64, u64, u64]
        L.oc_sphere_ktable.restype = None
        L.oc_sphere_ktable.argtypes = [i32, P]
        L.oc_build.restype = i64
        L.oc_build.argtypes = [i32, i32, P, i64, i64, i64, i32, i32, P, P, P, i32, dbl, i32,
                               P, P, P, i64, ctypes.POINTER(i64)]
        L.oc_build_mt.restype = i64
        L.oc_build_mt.argtypes = [i32, i32, P, i64, i32, i32, P, P, P, i32, dbl, i32,
                                  P, P, P, i64, i32]
        L.oc_spmv_sym.restype = None
        L.oc_spmv_sym.argtypes = [i64, P, P, P, i32, P, P]
        L.oc_spmv_sym_block.restype = None
        L.oc_spmv_sym_block.argtypes = [i64, i64, P, P, P, i32, P, P]
        L.oc_spmv_sym_mt.restype = None
        L.oc_spmv_sym_mt.argtypes = [i64, P, P, P, i32, P, P, i32]
        L.oc_ofl_find_k1.restype = i32
        L.oc_ofl_find_k1.argtypes = [i32, i32, i32, i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
        L.oc_conv2d.restype = None
        L.oc_conv2d.argtypes = [P, P, i32, i32, P, i32, i32, i32, i32]
        L.oc_ofl_vtable.restype = None
        L.oc_ofl_vtable.argtypes = [i32, i32, i32, i32, P, i64, i64, P]
        L.oc_test_lz2.restype = i32
        L.oc_test_lz2.argtypes = [u64, u64, i32, i32]
        L.oc_two_level_basis.restype = i64
        L.oc_two_level_basis.argtypes = [i32, i32, i32, i32, i32, i32, P, P, i64]
        L.oc_build_general.restype = i64
        L.oc_build_general.argtypes = [i32, P, i64, i64, i64, i32, i32, i32, P, P, P, i32, P, P, P, dbl, i32,
                                       P, P, P, i64, ctypes.POINTER(i64)]
        L.oc_ofl_wannier.restype = None
        L.oc_ofl_wannier.argtypes = [i32, i32, P, P, P, dbl, i64, i64, P, P]
        L.oc_two_level_build.restype = i64
        L.oc_two_level_build.argtypes = [i32, i32, P, P, i64, i64, i64, P, P, P, P, P, P, dbl, i32,
                                         P, P, P, i64, ctypes.POINTER(i64)]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def sector(kind, a=0, b=0, c=0, d=0):
    return _Sector(kind, a, b, c, d)


def basis(sec, N, Norb):
    L = lib()
    dim = L.oc_basis(ctypes.byref(sec), N, Norb, None, 0)
    out = np.zeros(max(dim, 1), dtype=np.uint64)
    L.oc_basis(ctypes.byref(sec), N, Norb, _p(out), dim)
    return out[:dim]


def sphere_ktable(maxLz):
    K = maxLz + 1
    kt = np.zeros(K ** 3, dtype=np.uint16)
    lib().oc_sphere_ktable(maxLz, _p(kt))
    return kt


class CRS:
    """Upper (reference) or full CRS as produced by the oracle."""

    def __init__(self, rowptr, col, val, is_complex, n_trimmed=0):
        self.rowptr, self.col, self.val = rowptr, col, val
        self.is_complex = is_complex
        self.n_trimmed = n_trimmed

    @property
    def nnz(self):
        return int(self.rowptr[-1])

    def values(self):
        if self.is_complex:
            return self.val.view(np.complex128)
        return self.val

    def to_dense(self, dim, mirror=True):
        """Dense matrix; with mirror=True the stored upper triangle U becomes
        U + (U - D)^H, the operator the reference's SpMV applies."""
        dt = np.complex128 if self.is_complex else np.float64
        M = np.zeros((dim, dim), dtype=dt)
        v = self.values()
        rp = self.rowptr
        n = len(rp) - 1
        for i in range(n):
            for j in range(rp[i], rp[i + 1]):
                M[i, self.col[j]] += v[j]
        if mirror:
            D = np.diag(np.diag(M))
            M = M + (M - D).conj().T
        return M


def build(N, Norb, bas, layout, K, ktab, vtab, eps=None, is_complex=False,
          tol=ZERO_TOL, full=False, row0=0, row1=None, threads=1):
    """Add_CdC_Terms + Add_CdCdCC_Terms + Trim + CRS for rows [row0, row1)."""
    L = lib()
    dim = len(bas)
    row1 = dim if row1 is None else row1
    vt = np.ascontiguousarray(vtab.view(np.float64) if is_complex else vtab, dtype=np.float64)
    ep = None
    if eps is not None:
        ep = np.ascontiguousarray(eps.view(np.float64) if is_complex else eps, dtype=np.float64)
    kt = np.ascontiguousarray(ktab, dtype=np.uint16)
    bas = np.ascontiguousarray(bas, dtype=np.uint64)
    ntr = ctypes.c_int64(0)
    if threads > 1 and row0 == 0 and row1 == dim:
        rp = np.zeros(dim + 1, dtype=np.int64)
        nnz = L.oc_build_mt(N, Norb, _p(bas), dim, layout, K, _p(kt), _p(vt), _p(ep),
                            int(is_complex), tol, int(full), _p(rp), None, None, 0, threads)
        if nnz < 0:
            raise RuntimeError("oracle build failed")
        col = np.zeros(max(nnz, 1), dtype=np.int32)
        val = np.zeros(max(nnz, 1) * (2 if is_complex else 1), dtype=np.float64)
        nnz2 = L.oc_build_mt(N, Norb, _p(bas), dim, layout, K, _p(kt), _p(vt), _p(ep),
                             int(is_complex), tol, int(full), _p(rp), _p(col), _p(val), nnz,
                             threads)
        assert nnz2 == nnz
        return CRS(rp, col[:nnz], val[:nnz * (2 if is_complex else 1)], is_complex)
    nnz = L.oc_build(N, Norb, _p(bas), dim, row0, row1, layout, K, _p(kt), _p(vt), _p(ep),
                     int(is_complex), tol, int(full), None, None, None, 0, ctypes.byref(ntr))
    if nnz < 0:
        raise RuntimeError("oracle build failed")
    rp = np.zeros(row1 - row0 + 1, dtype=np.int64)
    col = np.zeros(max(nnz, 1), dtype=np.int32)
    val = np.zeros(max(nnz, 1) * (2 if is_complex else 1), dtype=np.float64)
    nnz2 = L.oc_build(N, Norb, _p(bas), dim, row0, row1, layout, K, _p(kt), _p(vt), _p(ep),
                      int(is_complex), tol, int(full), _p(rp), _p(col), _p(val), nnz,
                      ctypes.byref(ntr))
    assert nnz2 == nnz
    return CRS(rp, col[:nnz], val[:nnz * (2 if is_complex else 1)], is_complex, ntr.value)


def build_general(Norb, bas, K, nk1, k1, v, nq1=None, q1=None, e=None, tol=ZERO_TOL, full=False,
                  row0=0, row1=None):
    """The build with general (hash) term tables, in the list format of
    TermTables.hash: Add_CdC_Terms + Add_CdCdCC_Terms with k1 lists, map
    accumulation in term order, Trim, CRS for rows [row0, row1)."""
    L = lib()
    dim = len(bas)
    row1 = dim if row1 is None else row1
    v = np.asarray(v)
    cplx = np.iscomplexobj(v) or (e is not None and np.iscomplexobj(e))
    vdt = np.complex128 if cplx else np.float64
    n = K ** 3
    nk = np.ascontiguousarray(nk1, dtype=np.uint8).reshape(n)
    kk = np.ascontiguousarray(k1, dtype=np.uint16).reshape(n, -1)
    vv = np.ascontiguousarray(v, dtype=vdt).reshape(n, -1).view(np.float64)
    qn = qq = ee = None
    mq = 0
    if nq1 is not None:
        qn = np.ascontiguousarray(nq1, dtype=np.uint8).reshape(K)
        qq = np.ascontiguousarray(q1, dtype=np.uint16).reshape(K, -1)
        ee = np.ascontiguousarray(e, dtype=vdt).reshape(K, -1).view(np.float64)
        mq = qq.shape[1]
    bas = np.ascontiguousarray(bas, dtype=np.uint64)
    ntr = ctypes.c_int64(0)
    args = (Norb, _p(bas), dim, row0, row1, K, int(cplx), kk.shape[1], _p(nk), _p(kk), _p(vv), mq,
            _p(qn), _p(qq), _p(ee), tol, int(full))
    nnz = L.oc_build_general(*args, None, None, None, 0, ctypes.byref(ntr))
    if nnz < 0:
        raise RuntimeError("oracle general build failed")
    rp = np.zeros(row1 - row0 + 1, dtype=np.int64)
    col = np.zeros(max(nnz, 1), dtype=np.int32)
    val = np.zeros(max(nnz, 1) * (2 if cplx else 1), dtype=np.float64)
    nnz2 = L.oc_build_general(*args, _p(rp), _p(col), _p(val), nnz, ctypes.byref(ntr))
    assert nnz2 == nnz
    return CRS(rp, col[:nnz], val[:nnz * (2 if cplx else 1)], cplx, ntr.value)


def mirror_upper(crs):
    """Full storage of the operator U + (U - D)^H that the reference's product
    applies to its upper CRS (sparse_linear_algebra.hpp:142-165): row r holds
    conj(U[i][r]) for every stored U[i][r] with i < r, ascending in i, then the
    stored entries of row r."""
    dim = len(crs.rowptr) - 1
    cnt = np.diff(crs.rowptr)
    rows = np.repeat(np.arange(dim), cnt)
    v = crs.values()
    low = crs.col > rows
    # lower entries (row = col of the upper entry, col = its row), sorted by (row, col)
    lr, lc, lv = crs.col[low].astype(np.int64), rows[low], np.conj(v[low]) if crs.is_complex else v[low]
    order = np.lexsort((lc, lr))
    lr, lc, lv = lr[order], lc[order], lv[order]
    # merge: for each row, its lower entries then its upper entries
    key_l = lr * 2
    key_u = rows.astype(np.int64) * 2 + 1
    allkey = np.concatenate([key_l, key_u])
    allcol = np.concatenate([lc, crs.col.astype(np.int64)])
    allval = np.concatenate([lv, v])
    order = np.argsort(allkey, kind="stable")
    col = allcol[order].astype(np.int32)
    val = allval[order]
    lcnt = np.bincount(lr, minlength=dim) if len(lr) else np.zeros(dim, dtype=np.int64)
    rp = np.zeros(dim + 1, dtype=np.int64)
    rp[1:] = np.cumsum(cnt + lcnt)
    raw = val.view(np.float64) if crs.is_complex else val
    return CRS(rp, col, np.ascontiguousarray(raw), crs.is_complex, crs.n_trimmed)


def ofl_wannier_quartic(dim_x, dim_y, ktable, vtable, u, tol, s0=0, s1=None):
    """ChangeToWannierBasis quartic terms for list slots [s0, s1) of the
    (k4*K + k3)*K + k2, x1 enumeration: (k1 labels, complex values)."""
    K = dim_x * dim_y
    s1 = K ** 3 * dim_x if s1 is None else s1
    kt = np.ascontiguousarray(ktable, dtype=np.uint16)
    vt = np.ascontiguousarray(vtable, dtype=np.complex128)
    uu = np.ascontiguousarray(u, dtype=np.complex128)
    k1 = np.zeros(s1 - s0, dtype=np.uint16)
    v = np.zeros(s1 - s0, dtype=np.complex128)
    lib().oc_ofl_wannier(dim_x, dim_y, _p(kt), _p(vt), _p(uu), tol, s0, s1, _p(k1), _p(v))
    return k1, v


# ------------------------------------------------------------------ two levels

def two_level_basis(N, n1, lz=None):
    """TwoLevelFermionFockBasis::GenerateFockSpace with n2 = n1 + 2 second-level
    orbitals; lz = 2 L_z total (None: no sector).  Returns (first, second)
    sorted by (first, second)."""
    L = lib()
    n2 = n1 + 2
    use, lzt = (0, 0) if lz is None else (1, lz)
    dim = L.oc_two_level_basis(N, n1, n2, use, lzt, n1 + 1, None, None, 0)
    f = np.zeros(max(dim, 1), dtype=np.uint64)
    g = np.zeros(max(dim, 1), dtype=np.uint64)
    L.oc_two_level_basis(N, n1, n2, use, lzt, n1 + 1, _p(f), _p(g), dim)
    return f[:dim], g[:dim]


def two_level_ktable(n1, level):
    """BaseGenerate4KTable for Landau level `level` of the two-level model
    (pseudopotential_model_base.hpp:96-139): maxLz = n1-1 (level 0) or n1+1
    (level 1); FindM1's error value m_maxLz2+2 = n1+3 converts to the label
    (maxLz - n1 - 3)/2 stored as uint16."""
    maxLz = n1 - 1 if level == 0 else n1 + 1
    K = maxLz + 1
    kt = np.zeros(K ** 3, dtype=np.uint16)
    err = n1 + 3
    it = 0
    for m4 in range(maxLz, -maxLz - 1, -2):
        for m3 in range(maxLz, -maxLz - 1, -2):
            for m2 in range(maxLz, -maxLz - 1, -2):
                m1 = m3 + m4 - m2
                if not (-maxLz <= m1 <= maxLz):
                    m1 = err
                kt[it] = ((-m1 + maxLz) // 2) & 0xFFFF
                it += 1
    return kt


def two_level_build(n1, first, second, kt1, vt1, kt2, vt2, eps1=None, eps2=None, tol=ZERO_TOL,
                    full=False, row0=0, row1=None):
    """The two-level Hamiltonian's upper (or full) CRS for rows [row0, row1):
    Add_CdC level 0, level 1, Add_CdCdCC level 0 (+Trim), level 1 (+Trim)."""
    L = lib()
    n2 = n1 + 2
    dim = len(first)
    row1 = dim if row1 is None else row1
    f = np.ascontiguousarray(first, dtype=np.uint64)
    g = np.ascontiguousarray(second, dtype=np.uint64)
    k1 = np.ascontiguousarray(kt1, dtype=np.uint16)
    k2 = np.ascontiguousarray(kt2, dtype=np.uint16)
    v1 = np.ascontiguousarray(vt1, dtype=np.float64)
    v2 = np.ascontiguousarray(vt2, dtype=np.float64)
    e1 = None if eps1 is None else np.ascontiguousarray(eps1, dtype=np.float64)
    e2 = None if eps2 is None else np.ascontiguousarray(eps2, dtype=np.float64)
    assert len(k1) == n1 ** 3 and len(k2) == n2 ** 3 and len(v1) == n1 ** 3 and len(v2) == n2 ** 3
    ntr = ctypes.c_int64(0)
    args = (n1, n2, _p(f), _p(g), dim, row0, row1, _p(k1), _p(v1), _p(e1), _p(k2), _p(v2), _p(e2), tol, int(full))
    nnz = L.oc_two_level_build(*args, None, None, None, 0, ctypes.byref(ntr))
    if nnz < 0:
        raise RuntimeError("oracle two-level build failed")
    rp = np.zeros(row1 - row0 + 1, dtype=np.int64)
    col = np.zeros(max(nnz, 1), dtype=np.int32)
    val = np.zeros(max(nnz, 1), dtype=np.float64)
    nnz2 = L.oc_two_level_build(*args, _p(rp), _p(col), _p(val), nnz, ctypes.byref(ntr))
    assert nnz2 == nnz
    return CRS(rp, col[:nnz], val[:nnz], False, ntr.value)


# ------------------------------------------------ flux-lattice V from Bloch

def ofl_find_k1(dim_x, dim_y, k2, k3, k4):
    gx, gy = ctypes.c_int(), ctypes.c_int()
    k1 = lib().oc_ofl_find_k1(dim_x, dim_y, k2, k3, k4, ctypes.byref(gx), ctypes.byref(gy))
    return k1, gx.value, gy.value


def conv2d(right, left, X, Y, top=0, left_off=0, bottom=0, right_off=0):
    """ArrayConvolution2D on (X*Y) complex arrays; returns the (2X-1)x(2Y-1) output."""
    r = np.ascontiguousarray(right, dtype=np.complex128)
    lf = np.ascontiguousarray(left, dtype=np.complex128)
    out = np.zeros((2 * X - 1) * (2 * Y - 1), dtype=np.complex128)
    lib().oc_conv2d(_p(r), _p(lf), X, Y, _p(out), top, left_off, bottom, right_off)
    return out.reshape(2 * X - 1, 2 * Y - 1)


def ofl_vtable(dim_x, dim_y, cut_x, cut_y, bloch, e0=0, e1=None):
    """GenerateQuarticTerms for folded-table entries [e0, e1); bloch is
    complex [kMax][(2cx+1)(2cy+1)][2] (spin fastest)."""
    K = dim_x * dim_y
    n = K * K * (K + 1) // 2
    e1 = n if e1 is None else e1
    b = np.ascontiguousarray(bloch, dtype=np.complex128)
    assert b.size == K * 2 * (2 * cut_x + 1) * (2 * cut_y + 1)
    out = np.zeros(e1 - e0, dtype=np.complex128)
    lib().oc_ofl_vtable(dim_x, dim_y, cut_x, cut_y, _p(b), e0, e1, _p(out))
    return out


def spmv_sym(crs, x, threads=1):
    dim = len(crs.rowptr) - 1
    if crs.is_complex:
        xx = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(dim, dtype=np.complex128)
    else:
        xx = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(dim, dtype=np.float64)
    vv = np.ascontiguousarray(crs.val)
    lib().oc_spmv_sym_mt(dim, _p(crs.rowptr), _p(crs.col), _p(vv), int(crs.is_complex),
                         _p(xx), _p(y), threads)
    return y


# ------------------------------------------------------------ exact CG / V table

@lru_cache(maxsize=None)
def _cg2_exact(j1, m1, j2, m2, J, M):
    """Exact CG coefficient as (sign, CG^2 as Fraction); arguments doubled.
    Racah's closed form (Racah 1942), Condon-Shortley phase."""
    if m1 + m2 != M or abs(m1) > j1 or abs(m2) > j2 or abs(M) > J:
        return 0, Fraction(0)
    if J > j1 + j2 or J < abs(j1 - j2) or (j1 + j2 + J) % 2:
        return 0, Fraction(0)
    h = lambda x: x // 2  # all combinations below are even
    A = h(j1 + j2 - J); B = h(j1 - m1); C = h(j2 + m2)
    D = h(J - j2 + m1); E = h(J - j1 - m2)
    S = Fraction(0)
    for k in range(max(0, -D, -E), min(A, B, C) + 1):
        S += Fraction((-1) ** k, factorial(k) * factorial(A - k) * factorial(B - k) *
                      factorial(C - k) * factorial(D + k) * factorial(E + k))
    pref = Fraction((J + 1) * factorial(h(J + j1 - j2)) * factorial(h(J - j1 + j2)) *
                    factorial(h(j1 + j2 - J)), factorial(h(j1 + j2 + J) + 1))
    pref *= (factorial(h(J + M)) * factorial(h(J - M)) * factorial(h(j1 - m1)) *
             factorial(h(j1 + m1)) * factorial(h(j2 - m2)) * factorial(h(j2 + m2)))
    sgn = (S > 0) - (S < 0)
    return sgn, pref * S * S


def _sqrt_frac(f: Fraction) -> float:
    getcontext().prec = 60
    return float((Decimal(f.numerator) / Decimal(f.denominator)).sqrt())


@lru_cache(maxsize=None)
def cg_value(j1, m1, j2, m2, J, M):
    """CG rounded once to double, zeroed when |CG| <= 1e-12 (the wrapper
    stores only |x| > EPS, clebsch_gordan_wrapper.hpp:121-126)."""
    s, f = _cg2_exact(j1, m1, j2, m2, J, M)
    if s == 0:
        return 0.0
    v = s * _sqrt_frac(f)
    return v if abs(v) > 1e-12 else 0.0


def sphere_vtable(maxLz, pseudopotentials):
    """BaseTermsFromPseudopotentials (pseudopotential_model_base.hpp:150-229),
    in the same floating-point operation order: for each (m4, m3, m2)
    descending, total = 0; for each V_m: temp = 0 (+ CG*CG when |M| <= 2J);
    total -= V_m * temp."""
    K = maxLz + 1
    out = np.zeros(K ** 3, dtype=np.float64)
    it = 0
    for m4 in range(maxLz, -maxLz - 1, -2):
        for m3 in range(maxLz, -maxLz - 1, -2):
            for m2 in range(maxLz, -maxLz - 1, -2):
                total = 0.0
                m1 = m3 + m4 - m2
                if -maxLz <= m1 <= maxLz:
                    for L in range(0, 2 * len(pseudopotentials), 2):
                        temp = 0.0
                        Lact = 2 * maxLz - L
                        M = m1 + m2
                        if -Lact <= M <= Lact:
                            temp += (cg_value(maxLz, m1, maxLz, m2, Lact, M) *
                                     cg_value(maxLz, m3, maxLz, m4, Lact, M))
                        total -= pseudopotentials[L // 2] * temp
                out[it] = total
                it += 1
    return out


# ------------------------------------------------------ Jordan-Wigner brute force

def jw_operators(norb):
    """Annihilation operators c_k on the 2^norb Fock space in the occupation
    basis |n_{norb-1} ... n_0>, with the Jordan-Wigner string over orbitals
    below k (orbital 0 is the rightmost operator, as in fermions.hpp)."""
    I2 = np.eye(2)
    Z = np.diag([1.0, -1.0])
    a = np.array([[0.0, 1.0], [0.0, 0.0]])  # |0><1|
    ops = []
    for k in range(norb):
        mats = []
        for q in range(norb - 1, -1, -1):  # kron order: most significant first
            if q == k:
                mats.append(a)
            elif q < k:
                mats.append(Z)
            else:
                mats.append(I2)
        M = mats[0]
        for m in mats[1:]:
            M = np.kron(M, m)
        ops.append(M)
    return ops


def dense_hamiltonian_jw_general(norb, states, quartic, quadratic=None):
    """<out|H|in> restricted to `states` for H = sum V c+_k1 c+_k2 c_k3 c_k4
    (quartic: iterable of ((k1,k2,k3,k4), V), repeats added) + sum E c+_k1 c_k2
    (quadratic: iterable of ((k1,k2), E))."""
    ops = jw_operators(norb)
    terms = list(quartic) + list(quadratic or [])
    dt = np.complex128 if any(np.iscomplexobj(np.asarray(v)) for _, v in terms) else np.float64
    H = np.zeros((2 ** norb, 2 ** norb), dtype=dt)
    for (k1, k2, k3, k4), v in quartic:
        H += v * (ops[k1].T @ ops[k2].T @ ops[k3] @ ops[k4])
    for (k1, k2), e in quadratic or []:
        H += e * (ops[k1].T @ ops[k2])
    idx = np.asarray(states, dtype=np.int64)
    return H[np.ix_(idx, idx)]


def dense_hamiltonian_jw(norb, states, quartic, eps=None):
    """<out|H|in> restricted to `states` from explicit operator matrices.
    quartic: dict {(k1,k2,k3,k4): V} for H = sum V c+_k1 c+_k2 c_k3 c_k4;
    eps: optional diagonal energies for sum eps_k c+_k c_k."""
    ops = jw_operators(norb)
    dt = np.complex128 if any(isinstance(v, complex) for v in quartic.values()) else np.float64
    H = np.zeros((2 ** norb, 2 ** norb), dtype=dt)
    for (k1, k2, k3, k4), v in quartic.items():
        H += v * (ops[k1].T @ ops[k2].T @ ops[k3] @ ops[k4])
    if eps is not None:
        for k, e in enumerate(eps):
            H += e * (ops[k].T @ ops[k])
    idx = np.asarray(states, dtype=np.int64)
    return H[np.ix_(idx, idx)]
