"""hamiltonian_diagonalization_amd -- MI355X exact-diagonalization hot path.

Python plumbing over the C ABI of ``lib/libhdiag_hip.so`` (declared in
``include/hdiag.h``).  The classes mirror the reference's interfaces
(paths relative to the reference repository):

  FockBasis        FermionFockBasis         src/basis/fermion_fock_basis.hpp:45-74
  TermTables       Quartic/QuadraticTermTables  src/hamiltonians/term_tables_base.hpp:44-174
  TermTables.hash  Quartic/QuadraticTermHashTables  src/hamiltonians/term_hash_tables_base.hpp:44-276
  Hamiltonian      SpinlessFermionHamiltonian<T> + HamiltonianBase<T,B>
                   src/hamiltonians/spinless_fermion_hamiltonian.hpp:52-436,
                   src/hamiltonians/hamiltonian_base.hpp:70-931
  Hamiltonian.spmv MatrixVectorFunction::Simple  src/hamiltonians/matrix_vector_routines.hpp:94-104
  Hamiltonian.lanczos  ArpackWrapper::DiagonalizeSymmetricMatrix
                   src/utilities/wrappers/arpack_wrapper.hpp:595-937

All compute runs in the HIP library; there is no CPU fallback.  Importing the
package does not touch the GPU; the first call that needs the library loads it
and raises ``HDError`` when it is missing.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libhdiag_hip.so")

HD_OK, HD_ERR_ARG, HD_ERR_HIP, HD_ERR_NOMEM, HD_ERR_UNSUPPORTED = 0, 1, 2, 3, 4
HD_ERR_NOT_CONVERGED, HD_ERR_EMPTY = 5, 6
SECTOR_NONE, SECTOR_LZ, SECTOR_K2D, SECTOR_K1D, SECTOR_LZ2, SECTOR_LIST = 0, 1, 2, 3, 4, 5
LAYOUT_SPHERE, LAYOUT_OFL, LAYOUT_HASH = 0, 1, 2
F64, F32, CODED, TILED, TILED_X32 = 0, 1, 2, 3, 4
STORAGE_FULL, STORAGE_UPPER_REF, STORAGE_MATRIX_FREE = 0, 1, 2
START_RANDOM, START_USER = 0, 1
ZERO_TOL = 1e-14  # _SPARSE_ZERO_TOL_ (src/hamiltonians/hamiltonian_data.hpp:54)


class HDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hdiag error {code}: {msg}")
        self.code = code


class _Sector(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("a", ctypes.c_int32), ("b", ctypes.c_int32),
                ("c", ctypes.c_int32), ("d", ctypes.c_int32)]


class _HamInfo(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int64), ("nnz", ctypes.c_int64), ("storage", ctypes.c_int32),
                ("is_complex", ctypes.c_int32), ("has_f32", ctypes.c_int32),
                ("has_f64", ctypes.c_int32), ("n_trimmed", ctypes.c_int64),
                ("max_row", ctypes.c_int64), ("t_count_ms", ctypes.c_double),
                ("t_scan_ms", ctypes.c_double), ("t_fill_ms", ctypes.c_double),
                ("t_sort_ms", ctypes.c_double), ("device_bytes", ctypes.c_size_t)]


class _TileInfo(ctypes.Structure):
    _fields_ = [("groups", ctypes.c_int64), ("entries", ctypes.c_int64), ("lines", ctypes.c_int64),
                ("chunks", ctypes.c_int64), ("rows_per_group", ctypes.c_int64), ("spmv_bytes", ctypes.c_int64),
                ("build_ms", ctypes.c_double), ("entry_bytes", ctypes.c_int64), ("block_window", ctypes.c_int64)]


class _LanczosOpts(ctypes.Structure):
    _fields_ = [("nev", ctypes.c_int32), ("ncv", ctypes.c_int32),
                ("max_restarts", ctypes.c_int32), ("tol", ctypes.c_double),
                ("start", ctypes.c_int32), ("seed", ctypes.c_uint64),
                ("d_start", ctypes.c_void_p), ("precision", ctypes.c_int32),
                ("f32_budget", ctypes.c_double), ("want_vectors", ctypes.c_int32),
                ("block", ctypes.c_int32), ("full_reorth", ctypes.c_int32)]


class _LanczosStats(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int32), ("nconv", ctypes.c_int32),
                ("restarts", ctypes.c_int32), ("matvecs", ctypes.c_int32),
                ("matvecs_f32", ctypes.c_int32), ("matvecs_f64", ctypes.c_int32),
                ("residual", ctypes.c_double * 64), ("time_total_ms", ctypes.c_double),
                ("time_spmv_ms", ctypes.c_double), ("time_ortho_ms", ctypes.c_double),
                ("reorth_steps", ctypes.c_int32), ("f32_cycles_discarded", ctypes.c_int32)]


class _BatchOpts(ctypes.Structure):
    _fields_ = [("nev", ctypes.c_int32), ("ncv", ctypes.c_int32), ("max_restarts", ctypes.c_int32),
                ("tol", ctypes.c_double), ("seed", ctypes.c_uint64), ("dense_below", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("full_reorth", ctypes.c_int32)]


class _BatchStats(ctypes.Structure):
    _fields_ = [("problems", ctypes.c_int32), ("lanczos_problems", ctypes.c_int32),
                ("dense_problems", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("restarts", ctypes.c_int32), ("steps", ctypes.c_int32),
                ("max_residual", ctypes.c_double), ("time_total_ms", ctypes.c_double),
                ("reorth_steps", ctypes.c_int32)]


_lib = None

# (name, restype, argtypes) of every C entry point the package binds
_P, _I32, _I64, _U32, _U64, _D = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64,
                                  ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double)
_PP = ctypes.POINTER(ctypes.c_void_p)
SIGNATURES = {
    "hd_last_error": (ctypes.c_char_p, []),
    "hd_version": (ctypes.c_int, []),
    "hd_malloc": (ctypes.c_int, [_PP, ctypes.c_size_t]),
    "hd_free": (ctypes.c_int, [_P]),
    "hd_memcpy_h2d": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "hd_memcpy_d2h": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "hd_memset": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_size_t]),
    "hd_synchronize": (ctypes.c_int, []),
    "hd_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "hd_basis_create": (ctypes.c_int, [ctypes.POINTER(_Sector), _U32, _U32, _PP]),
    "hd_basis_create_from_states": (ctypes.c_int, [_U32, _U32, _P, _I64, _PP]),
    "hd_basis_destroy": (ctypes.c_int, [_P]),
    "hd_basis_dim": (ctypes.c_int, [_P, ctypes.POINTER(_I64)]),
    "hd_basis_copy_states": (ctypes.c_int, [_P, _I64, _I64, _P]),
    "hd_basis_device_states": (ctypes.c_int, [_P, _PP]),
    "hd_basis_find_index": (ctypes.c_int, [_P, _U64, _P, _I64, _P]),
    "hd_sphere_ktable": (ctypes.c_int, [_I32, _P]),
    "hd_sphere_vtable": (ctypes.c_int, [_I32, _P, _I32, _P]),
    "hd_tables_create": (ctypes.c_int, [_I32, _U32, _P, _P, _I32, _P, _PP]),
    "hd_tables_create_two_level": (ctypes.c_int, [_U32, _P, _P, _P, _P, _P, _P, _PP]),
    "hd_tables_create_hash": (ctypes.c_int, [_U32, _I32, _I32, _P, _P, _P, _I32, _P, _P, _P, _PP]),
    "hd_ofl_wannier": (ctypes.c_int, [_I32, _I32, _D, _P, _P, _P, _P, _D, _P, _P, _P, _P, _P, _P,
                                      ctypes.POINTER(_D)]),
    "hd_ofl_ktable": (ctypes.c_int, [_I32, _I32, _P, _P, _P]),
    "hd_ofl_vtable": (ctypes.c_int, [_I32, _I32, _I32, _P, _P, _P]),
    "hd_tables_destroy": (ctypes.c_int, [_P]),
    "hd_ham_build": (ctypes.c_int, [_P, _P, _I32, _D, _I32, _PP]),
    "hd_ham_destroy": (ctypes.c_int, [_P]),
    "hd_ham_info": (ctypes.c_int, [_P, ctypes.POINTER(_HamInfo)]),
    "hd_ham_copy_crs": (ctypes.c_int, [_P, _I64, _I64, _I32, _P, _P, _P]),
    "hd_ham_device_crs": (ctypes.c_int, [_P, _PP, _PP, _PP, _PP]),
    "hd_ham_device_coded": (ctypes.c_int, [_P, _PP, _PP, _PP, ctypes.POINTER(_I32)]),
    "hd_ham_tile": (ctypes.c_int, [_P, _I32, _I32]),
    "hd_ham_tile_info": (ctypes.c_int, [_P, ctypes.POINTER(_TileInfo)]),
    "hd_ham_tile_order": (ctypes.c_int, [_P, ctypes.c_int64]),
    "hd_ham_device_tiled": (ctypes.c_int, [_P, _PP, _PP, _PP, _PP, _PP]),
    "hd_obs_density": (ctypes.c_int, [_P, _I32, _P, _P]),
    "hd_obs_density_density": (ctypes.c_int, [_P, _I32, _P, _P]),
    "hd_obs_translational_dd": (ctypes.c_int, [_P, _I32, _P, _I32, _P, _P, _P]),
    "hd_obs_rotational_dd": (ctypes.c_int, [_P, _I32, _P, _P, _P]),
    "hd_obs_participation_ratio": (ctypes.c_int, [_P, _I32, _P, ctypes.POINTER(_D)]),
    "hd_obs_most_probable": (ctypes.c_int, [_P, _I32, _P, _I32, _P, _P]),
    "hd_spmv": (ctypes.c_int, [_P, _I32, _P, _P]),
    "hd_spmv_host": (ctypes.c_int, [_P, _P, _P]),
    "hd_spmv_host_noalloc": (ctypes.c_int, [_P, _P, _P]),
    "hd_spmv_time": (ctypes.c_int, [_P, _I32, _I32, ctypes.POINTER(_D)]),
    "hd_spmv_get_schedule": (ctypes.c_int, [_P, _I32, ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    "hd_stream_copy": (ctypes.c_int, [_P, _P, ctypes.c_size_t, _I32, ctypes.POINTER(_D)]),
    "hd_spmv_set_lanes": (ctypes.c_int, [_P, _I32]),
    "hd_spmv_get_lanes": (ctypes.c_int, [_P, ctypes.POINTER(_I32)]),
    "hd_lanczos": (ctypes.c_int, [_P, ctypes.POINTER(_LanczosOpts), _P,
                                  ctypes.POINTER(_LanczosStats)]),
    "hd_lanczos_vectors": (ctypes.c_int, [_P, _I32, _P, _P]),
    "hd_full_diagonalize": (ctypes.c_int, [_P, _I32, _P]),
    "hd_mlp_create": (ctypes.c_int, [_I32, _I32, _U64, _PP]),
    "hd_mlp_create_states": (ctypes.c_int, [_I32, _I32, _I32, _U64, _PP]),
    "hd_mlp_num_states": (ctypes.c_int, [_P, ctypes.POINTER(_I32)]),
    "hd_mlp_destroy": (ctypes.c_int, [_P]),
    "hd_mlp_num_weights": (ctypes.c_int, [_P, ctypes.POINTER(_I64)]),
    "hd_mlp_get_weights": (ctypes.c_int, [_P, _P]),
    "hd_mlp_set_weights": (ctypes.c_int, [_P, _P]),
    "hd_mlp_forward": (ctypes.c_int, [_P, _P, _P, _I64, _P]),
    "hd_mlp_gradients": (ctypes.c_int, [_P, _P, _P, _P, _I64, ctypes.c_float,
                                        ctypes.POINTER(ctypes.c_float), _P]),
    "hd_mlp_train_step": (ctypes.c_int, [_P, _P, _P, _P, _I64, ctypes.c_float, ctypes.c_float,
                                         ctypes.POINTER(ctypes.c_float)]),
    "hd_mlp_start_vector": (ctypes.c_int, [_P, _P, _P, _P]),
    "hd_mlp_gemm_check": (ctypes.c_int, [_I32, _I32, _I32, _P, _P, _P]),
    "hd_sym_eig": (ctypes.c_int, [_I32, _P, _P, _P]),
    "hd_batch_create": (ctypes.c_int, [_P, _I32, _P, _I32, _D, _P]),
    "hd_batch_destroy": (ctypes.c_int, [_P]),
    "hd_batch_info": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "hd_batch_copy_pattern": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "hd_batch_spmm": (ctypes.c_int, [_P, _P, _P]),
    "hd_batch_spmm_time": (ctypes.c_int, [_P, _I32, ctypes.POINTER(_D)]),
    "hd_batch_solve": (ctypes.c_int, [_P, ctypes.POINTER(_BatchOpts), _P, _P, ctypes.POINTER(_BatchStats)]),
}


def lib():
    """Load libhdiag_hip.so (raises HDError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HDError(-1, f"{LIB_PATH} is missing: run __graft_entry__.build() "
                              "(make -C hamiltonian_diagonalization_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(code):
    if code != HD_OK:
        msg = lib().hd_last_error()
        raise HDError(code, msg.decode() if msg else "")


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return ctypes.c_void_p(a.ctypes.data)


# ------------------------------------------------------------------ device memory

class DeviceBuffer:
    """A raw HBM allocation (hd_malloc); freed on deletion."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(lib().hd_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p

    @classmethod
    def from_array(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        _check(lib().hd_memcpy_h2d(b.ptr, _ptr(a), a.nbytes))
        return b

    def to_array(self, dtype, count=None):
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(count, dtype=dtype)
        _check(lib().hd_memcpy_d2h(_ptr(out), self.ptr, out.nbytes))
        return out

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(lib().hd_memcpy_h2d(self.ptr, _ptr(a), a.nbytes))

    def __del__(self):
        if getattr(self, "ptr", None) is not None and _lib is not None:
            _lib.hd_free(self.ptr)
            self.ptr = None


# --------------------------------------------------------------------- sectors

class Sector:
    """Descriptor replacing the reference's sector predicate functions."""

    def __init__(self, kind, a=0, b=0, c=0, d=0):
        self.s = _Sector(kind, a, b, c, d)

    @classmethod
    def none(cls):
        return cls(SECTOR_NONE)

    @classmethod
    def lz(cls, lz_tot, max_lz):
        return cls(SECTOR_LZ, lz_tot, max_lz)

    @classmethod
    def k2d(cls, kx, ky, dim_x, dim_y):
        return cls(SECTOR_K2D, kx, ky, dim_x, dim_y)

    @classmethod
    def k1d(cls, ky, dim_y):
        return cls(SECTOR_K1D, ky, dim_y)

    @classmethod
    def lz2(cls, lz_tot, n1):
        """TestAngularMomentumSector2LL with max = m_maxLz2 = n1 + 1 over the
        combined orbitals of n1 first-level and n1 + 2 second-level orbitals."""
        return cls(SECTOR_LZ2, lz_tot, n1 + 1, n1)


class FockBasis:
    """FermionFockBasis::GenerateFockSpace on the device."""

    def __init__(self, sector: Sector, n_particles: int, n_orbitals: int):
        h = ctypes.c_void_p()
        _check(lib().hd_basis_create(ctypes.byref(sector.s), n_particles, n_orbitals,
                                     ctypes.byref(h)))
        self.h = h
        self.sector = sector
        self.n_particles = n_particles
        self.n_orbitals = n_orbitals
        d = ctypes.c_int64()
        _check(lib().hd_basis_dim(h, ctypes.byref(d)))
        self.dim = d.value
        self.is_list = False

    @classmethod
    def from_states(cls, states, n_particles: int, n_orbitals: int):
        """GenerateFockSpace with a predicate evaluated by the caller: the kept
        states, ascending (fermion_fock_basis.cpp:266-307); the device ranks
        them by binary search and Hamiltonians over them take the general
        build."""
        self = cls.__new__(cls)
        st = np.ascontiguousarray(states, dtype=np.uint64)
        h = ctypes.c_void_p()
        _check(lib().hd_basis_create_from_states(n_particles, n_orbitals, _ptr(st), len(st), ctypes.byref(h)))
        self.h = h
        self.sector = Sector(SECTOR_LIST)
        self.n_particles, self.n_orbitals = n_particles, n_orbitals
        self.dim = len(st)
        self.is_list = True
        return self

    def states(self, first=0, count=None) -> np.ndarray:
        count = self.dim - first if count is None else count
        out = np.empty(count, dtype=np.uint64)
        _check(lib().hd_basis_copy_states(self.h, first, count, _ptr(out)))
        return out

    def find_index(self, states: np.ndarray, offset: int = 0) -> np.ndarray:
        """FindFockStateIndex for many states (UINT64_MAX when absent)."""
        st = np.ascontiguousarray(states, dtype=np.uint64)
        din = DeviceBuffer.from_array(st)
        dout = DeviceBuffer(max(st.nbytes, 8))
        _check(lib().hd_basis_find_index(self.h, offset, din.ptr, len(st), dout.ptr))
        return dout.to_array(np.uint64, len(st))

    # ---- observables of a vector over this basis (struct Observables,
    # src/observables/fermion_observables.hpp:38-626), computed on the device

    def _dvec(self, vec):
        if isinstance(vec, DeviceBuffer):
            raise TypeError("pass (DeviceBuffer, is_complex) as a tuple, or a numpy array")
        if isinstance(vec, tuple):
            return vec[0], int(vec[1])
        v = np.ascontiguousarray(vec)
        if len(v) != self.dim:
            raise ValueError("vector length differs from the basis dimension")
        cplx = np.iscomplexobj(v)
        return DeviceBuffer.from_array(v.astype(np.complex128 if cplx else np.float64)), int(cplx)

    def density(self, vec) -> np.ndarray:
        """DensityFunction: <c+_k c_k> for every orbital k."""
        d, c = self._dvec(vec)
        out = np.zeros(self.n_orbitals)
        _check(lib().hd_obs_density(self.h, c, d.ptr, _ptr(out)))
        return out

    def density_density(self, vec) -> np.ndarray:
        """DensityDensityFunction: <n_k n_q> as an Norb x Norb array."""
        d, c = self._dvec(vec)
        out = np.zeros(self.n_orbitals ** 2)
        _check(lib().hd_obs_density_density(self.h, c, d.ptr, _ptr(out)))
        return out.reshape(self.n_orbitals, self.n_orbitals)

    def translational_dd(self, vec, add: np.ndarray, sub: np.ndarray) -> np.ndarray:
        """TranslationalDensityDensityFunction with StateAddition / StateSubtraction
        tables add[k, q] = k + q, sub[k, q] = k - q (shape Norb x nq)."""
        d, c = self._dvec(vec)
        a = np.ascontiguousarray(add, dtype=np.uint16)
        b = np.ascontiguousarray(sub, dtype=np.uint16)
        nq = a.shape[1]
        out = np.zeros(nq)
        _check(lib().hd_obs_translational_dd(self.h, c, d.ptr, nq, _ptr(a), _ptr(b), _ptr(out)))
        return out

    def rotational_dd(self, vec, inv_r60: np.ndarray) -> np.ndarray:
        """RotationalDensityDensityFunction: R(m), m = 0..5, with InverseR60 as a
        table inv_r60[k, alpha] (shape Norb x 6)."""
        d, c = self._dvec(vec)
        t = np.ascontiguousarray(inv_r60, dtype=np.uint16)
        out = np.zeros(12)
        _check(lib().hd_obs_rotational_dd(self.h, c, d.ptr, _ptr(t), _ptr(out)))
        return out[0::2] + 1j * out[1::2]

    def participation_ratio(self, vec) -> float:
        """GetParticipationRatio: 1 / sum |psi|^4."""
        d, c = self._dvec(vec)
        v = ctypes.c_double()
        _check(lib().hd_obs_participation_ratio(self.h, c, d.ptr, ctypes.byref(v)))
        return v.value

    def most_probable(self, vec, n: int):
        """GetMostProbable for one vector: (states, amplitudes) of the n largest |psi|."""
        d, c = self._dvec(vec)
        n = min(n, self.dim)
        st = np.zeros(n, dtype=np.uint64)
        amp = np.zeros(n, dtype=np.complex128 if c else np.float64)
        _check(lib().hd_obs_most_probable(self.h, c, d.ptr, n, _ptr(st), _ptr(amp)))
        return st, amp

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.hd_basis_destroy(self.h)
            self.h = None


class TermTables:
    """Quartic (+ optional diagonal quadratic) term tables on the device."""

    def __init__(self, layout: int, kmax: int, ktable: np.ndarray, vtable: np.ndarray,
                 eps: Optional[np.ndarray] = None):
        self.is_complex = np.iscomplexobj(vtable)
        kt = np.ascontiguousarray(ktable, dtype=np.uint16)
        vt = np.ascontiguousarray(vtable, dtype=np.complex128 if self.is_complex else np.float64)
        vtd = vt.view(np.float64)
        ep = None
        if eps is not None:
            ep = np.ascontiguousarray(eps, dtype=np.complex128 if self.is_complex else np.float64)
            ep = ep.view(np.float64)
        h = ctypes.c_void_p()
        _check(lib().hd_tables_create(layout, kmax, _ptr(kt), _ptr(vtd), int(self.is_complex),
                                      _ptr(ep), ctypes.byref(h)))
        self.h = h
        self.layout, self.kmax = layout, kmax
        self.ktable, self.vtable, self.eps = kt, vt, eps

    @classmethod
    def two_level(cls, n1: int, ktable1, vtable1, ktable2, vtable2, eps1=None, eps2=None):
        """Both Landau levels of the two-level sphere model as one table over
        the 2 n1 + 2 combined orbitals (level 0 on orbitals n1+2.., level 1
        on 0..n1+1)."""
        self = cls.__new__(cls)
        self.is_complex = False
        k1 = np.ascontiguousarray(ktable1, dtype=np.uint16)
        k2 = np.ascontiguousarray(ktable2, dtype=np.uint16)
        v1 = np.ascontiguousarray(vtable1, dtype=np.float64)
        v2 = np.ascontiguousarray(vtable2, dtype=np.float64)
        if len(k1) != n1 ** 3 or len(v1) != n1 ** 3 or len(k2) != (n1 + 2) ** 3 or len(v2) != (n1 + 2) ** 3:
            raise ValueError("two-level tables: n1^3 and (n1+2)^3 entries expected")
        e1 = None if eps1 is None else np.ascontiguousarray(eps1, dtype=np.float64)
        e2 = None if eps2 is None else np.ascontiguousarray(eps2, dtype=np.float64)
        if (e1 is not None and len(e1) != n1) or (e2 is not None and len(e2) != n1 + 2):
            raise ValueError("two-level energies: n1 and n1+2 entries expected")
        h = ctypes.c_void_p()
        _check(lib().hd_tables_create_two_level(n1, _ptr(k1), _ptr(v1), _ptr(e1), _ptr(k2), _ptr(v2), _ptr(e2),
                                                ctypes.byref(h)))
        self.h = h
        self.layout, self.kmax = LAYOUT_SPHERE, 2 * n1 + 2
        self.ktable, self.vtable, self.eps = (k1, k2), (v1, v2), (e1, e2)
        return self

    @classmethod
    def hash(cls, kmax: int, nk1, k1, v, nq1=None, q1=None, e=None):
        """General term tables as the lists TermHashTables' GetK1 returns
        (term_hash_tables.cpp:51-130): for entry (k4*kmax + k3)*kmax + k2,
        nk1[entry] labels k1[entry, :nk1] with values v[entry, :nk1]
        (complex or real, 0 where the hash table holds no value); for each k2,
        nq1[k2] quadratic labels q1[k2, :nq1] with values e (None: no
        quadratic terms).  k1 / v have shape (kmax^3, max_k1)."""
        self = cls.__new__(cls)
        v = np.asarray(v)
        self.is_complex = np.iscomplexobj(v) or (e is not None and np.iscomplexobj(e))
        vdt = np.complex128 if self.is_complex else np.float64
        n = kmax ** 3
        nk = np.ascontiguousarray(nk1, dtype=np.uint8).reshape(n)
        kk = np.ascontiguousarray(k1, dtype=np.uint16).reshape(n, -1)
        vv = np.ascontiguousarray(v, dtype=vdt).reshape(n, -1)
        if kk.shape != vv.shape:
            raise ValueError("k1 and v lists must have the same shape")
        max_k1 = kk.shape[1]
        qn = qq = ee = None
        max_q1 = 0
        if nq1 is not None:
            qn = np.ascontiguousarray(nq1, dtype=np.uint8).reshape(kmax)
            qq = np.ascontiguousarray(q1, dtype=np.uint16).reshape(kmax, -1)
            ee = np.ascontiguousarray(e, dtype=vdt).reshape(kmax, -1)
            if qq.shape != ee.shape:
                raise ValueError("q1 and e lists must have the same shape")
            max_q1 = qq.shape[1]
        h = ctypes.c_void_p()
        _check(lib().hd_tables_create_hash(kmax, int(self.is_complex), max_k1, _ptr(nk), _ptr(kk),
                                           _ptr(vv.view(np.float64)), max_q1, _ptr(qn), _ptr(qq),
                                           None if ee is None else _ptr(ee.view(np.float64)), ctypes.byref(h)))
        self.h = h
        self.layout, self.kmax = LAYOUT_HASH, kmax
        self.ktable, self.vtable, self.eps = (nk, kk), vv, (qn, qq, ee)
        return self

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.hd_tables_destroy(self.h)
            self.h = None


def hash_lists_from_array(layout: int, kmax: int, ktable, vtable, eps=None) -> dict:
    """ConvertTableFormat (interacting_ofl_model.cpp:597-619): the array
    tables as hash lists, one k1 per (k2,k3,k4) from GetK1 with
    V = GetVkkkk(k1,k2,k3,k4) (SetQuarticFromArray,
    term_hash_tables_base.hpp:231-252; the folded layout's k3 >= k4 symmetry,
    term_tables.cpp:95-175), and k1 = k2 with E_k2 for the quadratic terms
    (SetQuadraticFromArray :264-275).  Sphere tables list no k1 where
    GetK1 returns none (k1 >= kMax).  Keyword arguments of TermTables.hash."""
    K = kmax
    kt = np.asarray(ktable, dtype=np.int64)
    vt = np.asarray(vtable)
    k4, k3, k2 = np.meshgrid(np.arange(K), np.arange(K), np.arange(K), indexing="ij")
    k4, k3, k2 = k4.ravel(), k3.ravel(), k2.ravel()  # entry (k4*K + k3)*K + k2
    if layout == LAYOUT_SPHERE:
        i = (k4 * K + k3) * K + k2
        k1 = kt[i]
        nk1 = (k1 < K).astype(np.uint8)
        v = np.where(nk1 > 0, vt[i], 0)
    elif layout == LAYOUT_OFL:
        t1, t2 = np.minimum(k3, k4), np.maximum(k3, k4)
        row = (t1 * K + t2 - t1 * (t1 + 1) // 2) * K
        k1 = kt[row + k2]
        v = vt[row + np.where(k4 > k3, k1, k2)]
        nk1 = np.ones(K ** 3, dtype=np.uint8)
    else:
        raise ValueError("array layouts only")
    out = {"nk1": nk1, "k1": np.where(nk1 > 0, k1, 0).reshape(-1, 1), "v": np.asarray(v).reshape(-1, 1)}
    if eps is not None:
        ep = np.asarray(eps)
        out.update(nq1=np.ones(K, dtype=np.uint8), q1=np.arange(K).reshape(-1, 1), e=ep.reshape(-1, 1))
    return out


def ofl_wannier(dim_x: int, dim_y: int, offset_x: float, ktable, vtable, eps, u, tol: float,
                timing: bool = False):
    """ChangeToWannierBasis (interacting_ofl_model.cpp:728-923): the hybrid
    Wannier tables as hash lists (keyword arguments of TermTables.hash) from
    the folded Bloch-basis tables, the band energies eps (K, or None), the
    K x K change of basis u and the tolerance; the quartic transform runs on
    the device (kernel time in ms with timing=True)."""
    K = dim_x * dim_y
    kt = np.ascontiguousarray(ktable, dtype=np.uint16)
    vt = np.ascontiguousarray(vtable, dtype=np.complex128)
    uu = np.ascontiguousarray(u, dtype=np.complex128)
    if kt.size != K * K * (K + 1) // 2 or vt.size != kt.size or uu.size != K * K:
        raise ValueError("folded tables of K^2 (K+1)/2 entries and a K x K matrix expected")
    nk1 = np.zeros(K ** 3, dtype=np.uint8)
    k1 = np.zeros((K ** 3, dim_x), dtype=np.uint16)
    v = np.zeros((K ** 3, dim_x), dtype=np.complex128)
    ep = nq1 = q1 = e = None
    if eps is not None:
        ep = np.ascontiguousarray(eps, dtype=np.complex128)
        if ep.size != K:
            raise ValueError("eps needs K values")
        nq1 = np.zeros(K, dtype=np.uint8)
        q1 = np.zeros((K, dim_x), dtype=np.uint16)
        e = np.zeros((K, dim_x), dtype=np.complex128)
    ms = ctypes.c_double()
    _check(lib().hd_ofl_wannier(dim_x, dim_y, float(offset_x), _ptr(kt), _ptr(vt.view(np.float64)),
                                None if ep is None else _ptr(ep.view(np.float64)), _ptr(uu.view(np.float64)),
                                float(tol), _ptr(nk1), _ptr(k1), _ptr(v.view(np.float64)), _ptr(nq1), _ptr(q1),
                                None if e is None else _ptr(e.view(np.float64)), ctypes.byref(ms)))
    out = {"nk1": nk1, "k1": k1, "v": v, "nq1": nq1, "q1": q1, "e": e}
    return (out, ms.value) if timing else out


def sphere_ktable(max_lz: int) -> np.ndarray:
    kt = np.zeros((max_lz + 1) ** 3, dtype=np.uint16)
    _check(lib().hd_sphere_ktable(max_lz, _ptr(kt)))
    return kt


def ofl_ktable(dim_x: int, dim_y: int, with_g: bool = False):
    """Generate4KTable of the flux-lattice model (folded layout); with_g also
    returns the reciprocal-lattice shifts (gx, gy) of every entry."""
    K = dim_x * dim_y
    n = K * K * (K + 1) // 2
    kt = np.zeros(n, dtype=np.uint16)
    gx = np.zeros(n, dtype=np.int32)
    gy = np.zeros(n, dtype=np.int32)
    _check(lib().hd_ofl_ktable(dim_x, dim_y, _ptr(kt), _ptr(gx), _ptr(gy)))
    return (kt, gx, gy) if with_g else kt


def ofl_vtable(dim_x: int, dim_y: int, cutoff: int, bloch: np.ndarray, timing: bool = False):
    """GenerateQuarticTerms on the device.  bloch: complex [K][(2c+1)^2][2]
    (band 0, spin fastest).  Returns the folded complex V table (and the
    kernel time in ms with timing=True)."""
    K = dim_x * dim_y
    b = np.ascontiguousarray(bloch, dtype=np.complex128)
    if b.size != K * 2 * (2 * cutoff + 1) ** 2:
        raise ValueError("bloch must hold K * 2 * (2c+1)^2 coefficients")
    out = np.zeros(K * K * (K + 1) // 2, dtype=np.complex128)
    ms = ctypes.c_double()
    _check(lib().hd_ofl_vtable(dim_x, dim_y, cutoff, _ptr(b), _ptr(out), ctypes.byref(ms)))
    return (out, ms.value) if timing else out


def sphere_vtable(max_lz: int, pseudopotentials: Sequence[float]) -> np.ndarray:
    vm = np.ascontiguousarray(pseudopotentials, dtype=np.float64)
    vt = np.zeros((max_lz + 1) ** 3, dtype=np.float64)
    _check(lib().hd_sphere_vtable(max_lz, _ptr(vm), len(vm), _ptr(vt)))
    return vt


class LanczosResult:
    def __init__(self, evals, stats: _LanczosStats, nev):
        self.evals = evals
        self.converged = bool(stats.converged)
        self.nconv = stats.nconv
        self.restarts = stats.restarts
        self.matvecs = stats.matvecs
        self.matvecs_f32 = stats.matvecs_f32
        self.matvecs_f64 = stats.matvecs_f64
        self.residuals = np.array(stats.residual[:nev])
        self.time_ms = stats.time_total_ms
        self.time_spmv_ms = stats.time_spmv_ms
        self.time_ortho_ms = stats.time_ortho_ms
        self.reorth_steps = stats.reorth_steps
        self.f32_cycles_discarded = stats.f32_cycles_discarded

    def __repr__(self):
        return (f"LanczosResult(evals={self.evals}, matvecs={self.matvecs} "
                f"(f32 {self.matvecs_f32}, discarded {self.f32_cycles_discarded}), restarts={self.restarts}, "
                f"reorth={self.reorth_steps}, "
                f"max_resid={self.residuals.max():.2e}, {self.time_ms:.1f} ms, "
                f"spmv {self.time_spmv_ms:.1f} ms, ortho {self.time_ortho_ms:.1f} ms)")


class Hamiltonian:
    """SpinlessFermionHamiltonian build + CRS + SpMV + Lanczos on the device."""

    def __init__(self, basis: FockBasis, tables: TermTables, storage=STORAGE_FULL,
                 zero_tol=ZERO_TOL, values=None):
        if values is None:
            values = {STORAGE_FULL: (F64, F32, CODED), STORAGE_MATRIX_FREE: (F64,)}.get(storage, (F64, F32))
            if storage == STORAGE_FULL and (tables.layout == LAYOUT_HASH or getattr(basis, "is_list", False)):
                values = (F64, F32)  # general builds keep no coded copy
        mask = (1 if F64 in values else 0) | (2 if F32 in values else 0) | (4 if CODED in values else 0)
        h = ctypes.c_void_p()
        _check(lib().hd_ham_build(basis.h, tables.h, storage, zero_tol, mask, ctypes.byref(h)))
        self.h = h
        self.basis, self.tables = basis, tables
        self.storage = storage
        self.is_complex = tables.is_complex
        self.dim = basis.dim

    def info(self) -> dict:
        i = _HamInfo()
        _check(lib().hd_ham_info(self.h, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in _HamInfo._fields_}

    @property
    def nnz(self):
        return self.info()["nnz"]

    def crs(self, row0=0, row1=None, precision=F64):
        """(rowptr int64 relative to row0, col int32, values) host copies."""
        row1 = self.dim if row1 is None else row1
        rp = np.zeros(row1 - row0 + 1, dtype=np.int64)
        _check(lib().hd_ham_copy_crs(self.h, row0, row1, precision, _ptr(rp), None, None))
        n = int(rp[-1])
        col = np.zeros(max(n, 1), dtype=np.int32)
        if precision == F64:
            vdt = np.complex128 if self.is_complex else np.float64
        else:
            vdt = np.complex64 if self.is_complex else np.float32
        val = np.zeros(max(n, 1), dtype=vdt)
        _check(lib().hd_ham_copy_crs(self.h, row0, row1, precision, _ptr(rp), _ptr(col), _ptr(val)))
        return rp, col[:n], val[:n]

    def device_crs(self):
        """Device pointers (rowptr, col, val64, val32) as integers (0 if absent)."""
        p = [ctypes.c_void_p() for _ in range(4)]
        _check(lib().hd_ham_device_crs(self.h, *[ctypes.byref(x) for x in p]))
        return tuple(x.value or 0 for x in p)

    def device_coded(self):
        """Device pointers (code, diag, dict) as integers and the dictionary size."""
        p = [ctypes.c_void_p() for _ in range(3)]
        nd = ctypes.c_int32()
        _check(lib().hd_ham_device_coded(self.h, *[ctypes.byref(x) for x in p], ctypes.byref(nd)))
        return tuple(x.value or 0 for x in p) + (nd.value,)

    def tile(self, rows_per_group: int = 0, entry_bytes: int = 0) -> dict:
        """Builds the tiled layout of the coded operator (precision TILED);
        rows_per_group 128 or 64, 0 = 128 with a fallback to 64; entry_bytes
        3 (packed, read as 12-byte quads: real matrices with dictionaries of
        at most 2047 values), 4, or 0 = 3 where it applies.  Real and complex
        matrices."""
        _check(lib().hd_ham_tile(self.h, rows_per_group, entry_bytes))
        return self.tile_info()

    def tile_order(self, groups_per_xcd: int = -1) -> int:
        """Block order of the tiled product: each window of 8*w workgroups
        gives every XCD w consecutive row groups (0 = launch order, -1 = time
        the candidates and keep the fastest).  Returns the order in use."""
        _check(lib().hd_ham_tile_order(self.h, groups_per_xcd))
        return self.tile_info()["block_window"]

    def tile_info(self) -> dict:
        i = _TileInfo()
        _check(lib().hd_ham_tile_info(self.h, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in _TileInfo._fields_}

    def device_tiled(self):
        """Device pointers (bases, entries, lines, offsets, diagonal) of the tiled layout."""
        p = [ctypes.c_void_p() for _ in range(5)]
        _check(lib().hd_ham_device_tiled(self.h, *[ctypes.byref(x) for x in p]))
        return tuple(x.value or 0 for x in p)

    def vec_dtype(self, precision=F64):
        """dtype of x for a product in `precision` (y is FP64 except for F32)."""
        if precision in (F64, CODED, TILED):
            return np.complex128 if self.is_complex else np.float64
        return np.complex64 if self.is_complex else np.float32

    def out_dtype(self, precision=F64):
        return self.vec_dtype(F32 if precision == F32 else F64)

    def spmv(self, x: np.ndarray, precision=F64) -> np.ndarray:
        """y = H x (host arrays in, host array out; device SpMV in between)."""
        dt, dto = self.vec_dtype(precision), self.out_dtype(precision)
        xx = np.ascontiguousarray(x, dtype=dt)
        dx = DeviceBuffer.from_array(xx)
        dy = DeviceBuffer(self.dim * np.dtype(dto).itemsize)
        _check(lib().hd_spmv(self.h, precision, dx.ptr, dy.ptr))
        _check(lib().hd_synchronize())
        return dy.to_array(dto, self.dim)

    def spmv_device(self, dx: DeviceBuffer, dy: DeviceBuffer, precision=F64):
        _check(lib().hd_spmv(self.h, precision, dx.ptr, dy.ptr))

    def simple(self, x: np.ndarray) -> np.ndarray:
        """MatrixVectorFunction::Simple drop-in over host buffers (FP64)."""
        dt = self.vec_dtype(F64)
        xx = np.ascontiguousarray(x, dtype=dt)
        y = np.zeros_like(xx)
        _check(lib().hd_spmv_host(self.h, _ptr(xx), _ptr(y)))
        return y

    def simple_noalloc(self, x: np.ndarray) -> np.ndarray:
        """Simple() over host buffers staged through the library's fixed device
        array (hd_spmv_host_noalloc): no device allocation; dim <= 4,194,304
        real or 2,097,152 complex."""
        dt = self.vec_dtype(F64)
        xx = np.ascontiguousarray(x, dtype=dt)
        y = np.zeros_like(xx)
        _check(lib().hd_spmv_host_noalloc(self.h, _ptr(xx), _ptr(y)))
        return y

    def spmv_time_ms(self, precision=F64, iters=10) -> float:
        """Mean SpMV kernel time (HIP events on the library stream)."""
        v = ctypes.c_double()
        _check(lib().hd_spmv_time(self.h, precision, iters, ctypes.byref(v)))
        return v.value

    def spmv_bytes(self, precision=F64) -> int:
        """Algorithmic HBM bytes of one SpMV: nnz*(4 + sizeof(value)) +
        dim*(8 + sizeof(x) + sizeof(y)), x counted once."""
        z = 2 if self.is_complex else 1
        if precision == TILED:
            return self.tile_info()["spmv_bytes"]
        if precision == TILED_X32:  # x read in FP32 instead of FP64
            return self.tile_info()["spmv_bytes"] - 4 * self.dim
        if precision == CODED:
            # column + 16-bit code per entry; rowptr, diagonal, x, y per row
            return self.nnz * 6 + self.dim * (8 + 3 * 8 * z)
        vs = (8 if precision == F64 else 4) * z
        return self.nnz * (4 + vs) + self.dim * (8 + 2 * vs)

    def set_lanes(self, lanes: int):
        _check(lib().hd_spmv_set_lanes(self.h, lanes))

    def lanes(self) -> int:
        v = ctypes.c_int32()
        _check(lib().hd_spmv_get_lanes(self.h, ctypes.byref(v)))
        return v.value

    def schedule(self, precision=F64):
        """(lanes per row, threads per block) of the SpMV kernel for one value format."""
        lanes, threads = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().hd_spmv_get_schedule(self.h, precision, ctypes.byref(lanes), ctypes.byref(threads)))
        return lanes.value, threads.value

    def lanczos(self, nev=5, tol=1e-10, ncv=0, precision=F64, seed=0, start=None,
                f32_budget=0.0, max_restarts=0, raise_on_fail=True, block=1,
                full_reorth=False) -> LanczosResult:
        """precision F64, CODED or TILED: that exact operator throughout; F32:
        mixed, restart cycles in FP64 until the FP32-input operator's rounding
        fits f32_budget * tol (0 = 1).  block > 1 selects the block Lanczos
        (4, 8 or 16 vectors per SpMM); full_reorth orthogonalises every new
        vector against the whole basis (default: partial reorthogonalisation)."""
        o = _LanczosOpts()
        o.nev, o.ncv, o.max_restarts, o.tol = nev, ncv, max_restarts, tol
        o.seed, o.precision, o.f32_budget, o.want_vectors = seed, precision, f32_budget, 1
        o.block, o.full_reorth = block, int(bool(full_reorth))
        dstart = None
        if start is None:
            o.start = START_RANDOM
        elif isinstance(start, DeviceBuffer):
            o.start, o.d_start = START_USER, start.ptr
        else:
            dstart = DeviceBuffer.from_array(np.ascontiguousarray(start, dtype=self.vec_dtype(F64)))
            o.start, o.d_start = START_USER, dstart.ptr
        nev_eff = min(nev, self.dim)
        ev = np.zeros(nev_eff, dtype=np.float64)
        st = _LanczosStats()
        code = lib().hd_lanczos(self.h, ctypes.byref(o), _ptr(ev), ctypes.byref(st))
        if code != HD_OK and (raise_on_fail or code != HD_ERR_NOT_CONVERGED):
            _check(code)
        return LanczosResult(ev, st, nev_eff)

    def eigenvector(self, i: int) -> np.ndarray:
        out = np.zeros(self.dim, dtype=self.vec_dtype(F64))
        _check(lib().hd_lanczos_vectors(self.h, i, None, _ptr(out)))
        return out

    def full_diagonalize(self, nev: int) -> np.ndarray:
        ev = np.zeros(min(nev, self.dim), dtype=np.float64)
        _check(lib().hd_full_diagonalize(self.h, nev, _ptr(ev)))
        return ev

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.hd_ham_destroy(self.h)
            self.h = None


def header_functions():
    """Function names declared in include/hdiag.h."""
    import re
    path = os.path.join(os.path.dirname(HERE), "include", "hdiag.h")
    txt = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|char\s*\*|const char\s*\*)\s*\**\s*(hd_\w+)\s*\(",
                                 txt, re.M)))


class BatchProblem:
    """Sectors x parameter points solved together (hd_batch_*): one coded CRS
    pattern per sector shared by all points, point-fastest vectors."""

    def __init__(self, bases: Sequence[FockBasis], tables: Sequence[TermTables], zero_tol: float = ZERO_TOL):
        self.bases, self.tables = list(bases), list(tables)
        bp = (ctypes.c_void_p * len(self.bases))(*[b.h.value for b in self.bases])
        tp = (ctypes.c_void_p * len(self.tables))(*[t.h.value for t in self.tables])
        h = ctypes.c_void_p()
        _check(lib().hd_batch_create(bp, len(self.bases), tp, len(self.tables), zero_tol, ctypes.byref(h)))
        self.h = h
        r, n, nd, pp = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        _check(lib().hd_batch_info(h, ctypes.byref(r), ctypes.byref(n), ctypes.byref(nd), ctypes.byref(pp)))
        self.rows, self.nnz, self.ndict, self.points_padded = r.value, n.value, nd.value, pp.value
        self.points = len(self.tables)
        self.row0 = np.cumsum([0] + [b.dim for b in self.bases])[:-1]

    def pattern(self):
        """(rowptr, col, code, dict[ndict][Pp], diag[rows][Pp]) host copies."""
        rp = np.zeros(self.rows + 1, dtype=np.int64)
        col = np.zeros(self.nnz, dtype=np.int32)
        code = np.zeros(self.nnz, dtype=np.uint16)
        dic = np.zeros((max(self.ndict, 1), self.points_padded), dtype=np.float64)
        dg = np.zeros((self.rows, self.points_padded), dtype=np.float64)
        _check(lib().hd_batch_copy_pattern(self.h, _ptr(rp), _ptr(col), _ptr(code), _ptr(dic), _ptr(dg)))
        return rp, col, code, dic[:self.ndict], dg

    def spmm(self, x: np.ndarray) -> np.ndarray:
        """Y[:, p] = H_p X[:, p] for host X of shape (rows, points)."""
        xp = np.zeros((self.rows, self.points_padded), dtype=np.float64)
        xp[:, :self.points] = x
        dx, dy = DeviceBuffer.from_array(xp), DeviceBuffer(xp.nbytes)
        _check(lib().hd_batch_spmm(self.h, dx.ptr, dy.ptr))
        return dy.to_array(np.float64, xp.size).reshape(xp.shape)[:, :self.points]

    def spmm_time_ms(self, iters: int = 5) -> float:
        v = ctypes.c_double()
        _check(lib().hd_batch_spmm_time(self.h, iters, ctypes.byref(v)))
        return v.value

    def solve(self, nev=3, tol=1e-10, ncv=0, seed=0, dense_below=0, use_graph=True, max_restarts=0,
              raise_on_fail=True, full_reorth=False):
        """(evals[sector][point][nev], residuals, stats dict)."""
        o = _BatchOpts()
        o.nev, o.ncv, o.max_restarts, o.tol, o.seed = nev, ncv, max_restarts, tol, seed
        o.dense_below, o.use_graph = dense_below, 1 if use_graph else 0
        o.full_reorth = 1 if full_reorth else 0
        S, P = len(self.bases), self.points
        ev = np.zeros((S, P, nev), dtype=np.float64)
        rs = np.zeros((S, P, nev), dtype=np.float64)
        st = _BatchStats()
        code = lib().hd_batch_solve(self.h, ctypes.byref(o), _ptr(ev), _ptr(rs), ctypes.byref(st))
        stats = {f: getattr(st, f) for f, _ in _BatchStats._fields_}
        if code != HD_OK and (raise_on_fail or code != HD_ERR_NOT_CONVERGED):
            _check(code)
        return ev, rs, stats

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.hd_batch_destroy(self.h)
            self.h = None


def sym_eig(a: np.ndarray):
    """(eigenvalues ascending, eigenvectors as columns) of a symmetric matrix by
    the library's host eigensolver (hd_sym_eig)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    w = np.zeros(n)
    z = np.zeros((n, n))
    _check(lib().hd_sym_eig(n, _ptr(a), _ptr(w), _ptr(z)))
    return w, z
