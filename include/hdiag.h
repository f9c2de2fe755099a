/*
 * hdiag.h -- C ABI of libhdiag_hip.so, the MI355X (gfx950) exact-diagonalization
 * hot path: Fock basis, Hamiltonian build into CRS, SpMV and thick-restart
 * Lanczos, plus the learned start-vector network.
 *
 * Conventions
 *  - Every entry point returns an hd_status (0 = HD_OK) and never throws.
 *    hd_last_error() returns a thread-local message for the last failure.
 *  - Pointers named d_* are device (HBM) pointers; h_* are host pointers.
 *  - Handles are opaque and bound to the device that was current when they
 *    were created.  Calls are synchronous with respect to the host unless
 *    stated otherwise; all device work runs on the handle's own stream.
 *  - Types follow the reference (src/utilities/general/orbital_and_state_defs.hpp:36-58):
 *    fock_t = uint64_t, kState_t = uint16_t, iSize_t = uint32_t.
 *
 * Each declaration names the reference interface it replaces
 * (paths relative to the reference repository root).
 */
#ifndef HDIAG_H
#define HDIAG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    HD_OK = 0,
    HD_ERR_ARG = 1,           /* invalid argument (reference: mpi.m_exitFlag = true) */
    HD_ERR_HIP = 2,           /* HIP runtime error */
    HD_ERR_NOMEM = 3,         /* device allocation failed */
    HD_ERR_UNSUPPORTED = 4,   /* valid input outside what the kernels implement */
    HD_ERR_NOT_CONVERGED = 5, /* Lanczos hit its iteration limit (ARPACK INFO=1) */
    HD_ERR_EMPTY = 6          /* zero-dimensional sector (GenerateFockSpace returns true) */
} hd_status;

/* Sector predicate descriptors; they replace the std::function<bool(fock_t)>
 * constraint passed to FermionFockBasis::GenerateFockSpace
 * (src/basis/fermion_fock_basis.cpp:241-354). */
typedef enum {
    HD_SECTOR_NONE = 0, /* [](fock_t){return true;}  (interacting_ofl_model.cpp:1010) */
    HD_SECTOR_LZ = 1,   /* TestAngularMomentumSector(state, a=lzTot, b=maxLz)
                           (angular_momentum_constraint.hpp:41-62) */
    HD_SECTOR_K2D = 2,  /* TestLinearMomentumSector2(state, a=kx, b=ky, c=dimX, d=dimY)
                           (linear_momentum_constraints_2D.cpp:38-79) */
    HD_SECTOR_K1D = 3,  /* TestLinearMomentumSector1(state, a=ky, b=dimY)
                           (linear_momentum_constraints_2D.cpp:87-110) */
    HD_SECTOR_LZ2 = 4,  /* TestAngularMomentumSector2LL(state, a=lzTot, b=max) with
                           c = n1 first-level orbitals
                           (two_level_angular_momentum_constraint.hpp:41-73), on the
                           combined state (m_first << (n1+2)) | m_second: the
                           TwoLevelFockState order is ascending combined order
                           (two_level_fermion_fock_basis.cpp:40-45) */
    HD_SECTOR_LIST = 5  /* the states of any std::function predicate, enumerated by
                           the caller (hd_basis_create_from_states) */
} hd_sector_kind;

typedef struct {
    int32_t kind;
    int32_t a, b, c, d;
} hd_sector;

/* Quartic table layouts (QuarticTermTables flavours). */
typedef enum {
    HD_LAYOUT_SPHERE = 0, /* index (k4*kMax+k3)*kMax+k2; k1 >= kMax means "no term"
                             (src/fqhe_sphere/pseudopotential_model/term_tables.cpp:109-149) */
    HD_LAYOUT_OFL = 1,    /* folded on k3>=k4: (t1*kMax+t2-t1(t1+1)/2)*kMax+t3
                             (src/optical_flux_lattice/interacting_ofl_model/term_tables.cpp:95-175) */
    HD_LAYOUT_HASH = 2    /* general lists of k1 per (k2,k3,k4) and per k2 (hd_tables_create_hash;
                             src/hamiltonians/term_hash_tables_base.hpp:44-276) */
} hd_layout;

/* Value formats of the device CRS / SpMV:
 *   HD_F64    FP64 (complex FP64) values, FP64 vectors
 *   HD_F32    FP32 (complex FP32) values, FP32 vectors, FP32 accumulation
 *   HD_CODED  16-bit codes into an FP64 dictionary of the distinct
 *             off-diagonal values (sign in bit 15, 0xFFFF = diagonal, kept in a
 *             separate FP64 array): the exact FP64 operator in 6 bytes per
 *             entry; FP64 vectors and accumulation.
 *   HD_TILED  the coded operator re-laid out in row groups whose x lines
 *             are staged through LDS (hd_ham_tile): 3 or 4 bytes per entry, the
 *             same exact FP64 matrix; real and complex matrices.
 *   HD_TILED_X32  the tiled operator with x in FP32 (half the staged x bytes)
 *             and y in FP64: FP64 dictionary values, FP64 products and FP64
 *             accumulation, so y is the exact FP64 operator applied to x
 *             (which is FP32) with the rounding of HD_TILED; real matrices. */
typedef enum { HD_F64 = 0, HD_F32 = 1, HD_CODED = 2, HD_TILED = 3, HD_TILED_X32 = 4 } hd_precision;

typedef enum {
    HD_STORAGE_FULL = 0,     /* both triangles, int64 row pointers (GPU SpMV layout) */
    HD_STORAGE_UPPER_REF = 1, /* the reference's upper-triangle CRS: M[i][j] = <j|H|i>, j >= i
                                (src/hamiltonians/spinless_fermion_hamiltonian.hpp:362,
                                 src/utilities/data_structures/crs_sparse_matrix.hpp:127-133) */
    HD_STORAGE_MATRIX_FREE = 2 /* no CRS: the pair tables and the diagonal (dim values) stay
                                  resident and every SpMV recomputes the full-storage rows
                                  from the basis (FP64 only; the matrix-free product the
                                  reference declares as MatrixVectorFunction::Recalculate,
                                  src/hamiltonians/matrix_vector_routines.hpp:106-143, and
                                  never calls), for sectors whose CRS does not fit in HBM;
                                  sector bases with pair-structure tables only (a list basis
                                  or hash tables: HD_ERR_UNSUPPORTED) */
} hd_storage;

typedef struct hd_basis hd_basis;
typedef struct hd_tables hd_tables;
typedef struct hd_hamiltonian hd_hamiltonian;

/* ----------------------------------------------------------------- runtime */

const char* hd_last_error(void);
int hd_version(void);
/* Device buffers for hosts without their own device allocator. */
int hd_malloc(void** d_ptr, size_t bytes);
int hd_free(void* d_ptr);
int hd_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes);
int hd_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes);
int hd_memset(void* d_dst, int value, size_t bytes);
int hd_synchronize(void);
int hd_device_count(int* n);

/* ------------------------------------------------------------------- basis */

/* FermionFockBasis::GenerateFockSpace (src/basis/fermion_fock_basis.cpp:241-354):
 * the sector basis in ascending numeric order, generated on the device by
 * combinatorial unranking.  Returns HD_ERR_EMPTY (and no handle) for a
 * zero-dimensional sector, mirroring the reference's `true` return. */
int hd_basis_create(const hd_sector* sector, uint32_t n_particles, uint32_t n_orbitals,
                    hd_basis** out);
/* FermionFockBasis::GenerateFockSpace with a predicate the device cannot
 * evaluate (fermion_fock_basis.cpp:266-307 walks every N-particle state and
 * keeps those the std::function accepts): the caller passes the kept states,
 * ascending, each with n_particles bits below n_orbitals.  The basis ranks by
 * binary search (FindFockStateIndex, fermion_fock_basis.cpp:134-161).
 * hd_ham_build over such a basis runs the general build (array tables are
 * read as one-k1 lists), full or upper storage, FP64/FP32 values; out states
 * that are not in the list are skipped, as the reference's search miss. */
int hd_basis_create_from_states(uint32_t n_particles, uint32_t n_orbitals, const uint64_t* h_states, int64_t dim,
                                hd_basis** out);
int hd_basis_destroy(hd_basis* b);
/* FermionFockBasis::GetStoredDimension (fermion_fock_basis.cpp:222-225). */
int hd_basis_dim(const hd_basis* b, int64_t* dim);
/* FermionFockBasis::GetFockBasis (fermion_fock_basis.cpp:360-372): copies
 * states [first, first+count) to host memory. */
int hd_basis_copy_states(const hd_basis* b, int64_t first, int64_t count, uint64_t* h_out);
/* Device pointer to the basis (dim uint64 states), owned by the handle. */
int hd_basis_device_states(const hd_basis* b, const uint64_t** d_states);
/* FermionFockBasis::FindFockStateIndex (fermion_fock_basis.cpp:134-161) for a
 * batch of device states: d_index[i] = offset + index, or UINT64_MAX
 * (_SEARCH_ERROR_CODE_) when the state is not in the sector. */
int hd_basis_find_index(const hd_basis* b, uint64_t offset, const uint64_t* d_states,
                        int64_t n, uint64_t* d_index);

/* ------------------------------------------------------------------ tables */

/* BaseGenerate4KTable (src/fqhe_sphere/pseudopotential_model/pseudopotential_model_base.hpp:97-139):
 * h_ktable receives (maxLz+1)^3 entries. */
int hd_sphere_ktable(int32_t max_lz, uint16_t* h_ktable);
/* BaseTermsFromPseudopotentials (pseudopotential_model_base.hpp:150-229):
 * V_{k1k2k3k4} = -sum_m V_m CG(S,m1;S,m2|J,M) CG(S,m3;S,m4|J,M), J = 2S - m,
 * with exact-rational Clebsch-Gordan coefficients rounded once to double.
 * h_vtable receives (maxLz+1)^3 doubles. */
int hd_sphere_vtable(int32_t max_lz, const double* h_pseudopotentials, int32_t n_pseudo,
                     double* h_vtable);

/* QuarticTermTables / QuadraticTermTables (term_tables_base.hpp:44-174) moved
 * to the device.  h_vtable holds doubles (interleaved re,im when is_complex);
 * h_eps is the optional diagonal quadratic table E_k (kMax entries, interleaved
 * when complex) or NULL when Add_CdC_Terms is not called.  The tables must
 * conserve the pair structure (k1 = f(k2,k3,k4) only produces diagonal terms or
 * two-particle hops), which both reference layouts do; other tables give
 * HD_ERR_UNSUPPORTED. */
int hd_tables_create(int32_t layout, uint32_t kmax, const uint16_t* h_ktable,
                     const double* h_vtable, int32_t is_complex, const double* h_eps,
                     hd_tables** out);
/* The two Landau levels of SphereTwoLevelPseudopotentialModel
 * (two_level_pseudopotential_model.cpp:61-110, 188-200) as one table over the
 * combined orbitals of TwoLevelFockState (m_first << (n1+2)) | m_second:
 * level-0 sphere tables over n1 orbitals (maxLz = n1-1: m_quarticTables,
 * m_quadraticTables), level-1 tables over n1+2 orbitals (maxLz = n1+1:
 * m_quarticTables2LL, m_quadraticTables2LL), real values, no terms across
 * levels (TwoLevelSpinlessFermionHamiltonian accepts only same-level terms,
 * two_level_spinless_fermion_hamiltonian.hpp:155-164, 317-326).  A k-table label
 * >= the level's orbital count means no k1 (GetK1, term_tables.cpp:109-124), so
 * the FindM1 error label (0xFFFE in level 0, 0xFFFF in level 1) and
 * hd_sphere_ktable's 0xFFFF act alike.  h_eps1 / h_eps2 (n1 and n1+2 energies,
 * SetEnergyLevels :91-110) may be NULL.  hd_ham_build with this table and an
 * HD_SECTOR_LZ2 or HD_SECTOR_NONE basis over 2n1+2 orbitals follows the
 * model's call order: E terms of level 0 then 1, quartic terms of level 0, its
 * Trim, quartic terms of level 1, final Trim.  Not accepted by hd_batch_create. */
int hd_tables_create_two_level(uint32_t n1, const uint16_t* h_ktable1, const double* h_vtable1,
                               const double* h_eps1, const uint16_t* h_ktable2,
                               const double* h_vtable2, const double* h_eps2, hd_tables** out);
/* TermHashTables (src/hamiltonians/term_hash_tables_base.hpp:44-276) as read by
 * QuarticTermHashTables / QuadraticTermHashTables::GetK1, GetVkkkk and GetEkk
 * (src/optical_flux_lattice/interacting_ofl_model/term_hash_tables.cpp:51-130):
 * general term tables, given as the lists those calls return.
 *   Quartic: for every (k2,k3,k4), entry e = (k4*kmax + k3)*kmax + k2 holds
 *   h_nk1[e] <= max_k1 labels k1 = h_k1[e*max_k1 + j] (GetK1's buffer, in its
 *   order) and V(k1,k2,k3,k4) = h_v[e*max_k1 + j] (interleaved re,im when
 *   is_complex; 0.0 where the hash table holds no value, GetVkkkk's empty
 *   return).  A (k2,k3,k4) the k table does not hold has h_nk1[e] = 0.
 *   Quadratic: h_nq1[k2] <= max_q1 labels k1 = h_q1[k2*max_q1 + j] with
 *   E(k1,k2) = h_e[k2*max_q1 + j]; h_nq1 = NULL when Add_CdC_Terms is not
 *   called.  Off-diagonal quadratic terms (k1 != k2) are one-body hops.
 * kmax <= 64; every label < kmax.  hd_ham_build with these tables runs the
 * general device build (Add_CdC_Terms then Add_CdCdCC_Terms in the reference's
 * loop order with these lists, out states outside the sector skipped as
 * FindFockStateIndex's miss, duplicate columns of a row summed in term order
 * by a per-row stable sort, then Trim): full or upper storage, FP64 and FP32
 * value copies (no coded copy, so no tiled layout; not accepted by
 * hd_batch_create). */
int hd_tables_create_hash(uint32_t kmax, int32_t is_complex, int32_t max_k1, const uint8_t* h_nk1,
                          const uint16_t* h_k1, const double* h_v, int32_t max_q1, const uint8_t* h_nq1,
                          const uint16_t* h_q1, const double* h_e, hd_tables** out);
int hd_tables_destroy(hd_tables* t);

/* ------------------------------------------------- flux-lattice term tables */

/* InteractingOflModel::Generate4KTable (interacting_ofl_model.cpp:258-296) with
 * FindK1 (:374-407): the folded k table over k4 <= k3 (k2 fastest),
 * (K+1)K^2/2 entries for K = dim_x*dim_y <= 64, and the reciprocal-lattice
 * shifts (h_gx, h_gy may be NULL). */
int hd_ofl_ktable(int32_t dim_x, int32_t dim_y, uint16_t* h_ktable, int32_t* h_gx, int32_t* h_gy);
/* InteractingOflModel::GenerateQuarticTerms (interacting_ofl_model.cpp:457-590)
 * on the device, bit-identical to the reference's sequential FP64 (Kahan sums
 * of ArrayConvolution2D, array_convolution.hpp:136-230).  h_bloch holds, for
 * each k, the band-0 block of NonInteractingOflModel::m_blochCoefficients:
 * 2(2c+1)^2 complex values (interleaved re,im) with the spin index fastest
 * (GetBlochCoefficients, noninteracting_ofl_model.cpp:550-582); the reference
 * generates them with one cutoff c in both directions
 * (NonInteractingOflModelGrid::GenerateBlochWaveFunctionTable).  Writes the
 * (K+1)K^2/2 complex V (interleaved) in the folded order, before
 * SetCoefficient(0.5*interactionStrength).  kernel_ms (may be NULL) receives the
 * kernel time. */
int hd_ofl_vtable(int32_t dim_x, int32_t dim_y, int32_t cutoff, const double* h_bloch, double* h_vtable,
                  double* kernel_ms);
/* InteractingOflModel::ChangeToWannierBasis (interacting_ofl_model.cpp:728-923)
 * with GenerateWannier2KTable / GenerateWannier4KTable (:300-355): the tables
 * of the hybrid Wannier basis |x,ky> in the list format of
 * hd_tables_create_hash (complex, max_k1 = max_q1 = dim_x), K = dim_x*dim_y.
 * Inputs: the Bloch-basis array tables (h_ktable from hd_ofl_ktable, h_vtable
 * the folded complex V after SetCoefficient), the complex band energies h_eps
 * (K values, m_quadraticTables), the change-of-basis matrix h_u (K x K
 * complex, row x*dim_y+ky, column kx*dim_y+ky',
 * NonInteractingOflModelGrid::GenerateWannierCoefficients), offset_x and tol
 * (matrixElementTol).
 *   Quartic, on the device: for (k2,k3,k4) the list is k1 = x1*dim_y + ky1,
 *   x1 ascending, ky1 = FindK1(k2y,k3y,k4y) (:410-432), and V'(k1,k2,k3,k4) is
 *   the sum over k4x, k3x, k2x (k2x fastest) of
 *   V(q1,q2,q3,q4) conj(U[k1,q1] U[k2,q2]) U[k3,q3] U[k4,q4] with q1 from the
 *   array k table, products left to right, each complex product
 *   (ac - bd, ad + bc) with every real product rounded, summed from 0 in FP64:
 *   bit-identical to the sequential C++.  Stored where |sum| > tol, else 0.0
 *   (no hash entry, GetVkkkk's empty return).
 *   Quadratic, on the host: for k2 = (x2,ky) the list is (x1,ky), x1
 *   ascending, E'(k1,k2) = exp(2 pi i |x2-x1| offset_x/dim_x) out[ky][|x2-x1|]
 *   with out[ky] the unnormalised backward DFT (sign +1) of E(kx,ky)/dim_x
 *   over kx.  The reference takes the DFT from FFTW
 *   (DiscreteFourierTransform1D, fftw_wrapper.cpp:36-70); here it is the
 *   direct sum in long double, rounded once, so it agrees with FFTW to a few
 *   ulp, not bit for bit.
 * Outputs (host): h_nk1[K^3], h_k1[K^3 dim_x], h_v[K^3 dim_x] complex,
 * h_nq1[K], h_q1[K dim_x], h_e[K dim_x] complex; kernel_ms may be NULL. */
int hd_ofl_wannier(int32_t dim_x, int32_t dim_y, double offset_x, const uint16_t* h_ktable, const double* h_vtable,
                   const double* h_eps, const double* h_u, double tol, uint8_t* h_nk1, uint16_t* h_k1,
                   double* h_v, uint8_t* h_nq1, uint16_t* h_q1, double* h_e, double* kernel_ms);

/* ------------------------------------------------------------------- build */

typedef struct {
    int64_t dim;
    int64_t nnz;          /* stored entries (storage dependent) */
    int32_t storage;      /* hd_storage */
    int32_t is_complex;
    int32_t has_f32;      /* an FP32 (complex FP32) value copy is resident */
    int32_t has_f64;
    int64_t n_trimmed;    /* entries removed by Trim(|v| < zero_tol) */
    int64_t max_row;      /* longest row */
    double t_count_ms, t_scan_ms, t_fill_ms, t_sort_ms; /* device time per stage */
    size_t device_bytes;  /* HBM held by the handle */
} hd_ham_info_t;

/* SpinlessFermionHamiltonian<T>::Add_CdC_Terms + Add_CdCdCC_Terms
 * (src/hamiltonians/spinless_fermion_hamiltonian.hpp:143-436) followed by
 * MappedSparseMatrix::Trim and CrsSparseMatrix::Initialize(map)
 * (mapped_sparse_matrix.hpp:423-462, crs_sparse_matrix.hpp:193-254), all on the
 * device: count, exclusive scan, fill, per-row sort.  Values are accumulated
 * in FP64 in the reference's term order; value_mask selects which value
 * copies stay resident (bit 0 = FP64, bit 1 = FP32, bit 2 = coded; full
 * storage only for the coded copy). */
int hd_ham_build(const hd_basis* b, const hd_tables* t, int32_t storage, double zero_tol,
                 int32_t value_mask, hd_hamiltonian** out);
int hd_ham_destroy(hd_hamiltonian* h);
int hd_ham_info(const hd_hamiltonian* h, hd_ham_info_t* info);
/* CrsSparseMatrix accessors (GetRowStarts/GetCols/GetData, crs_sparse_matrix.hpp:297-330)
 * as host copies of rows [row0, row1): h_rowptr gets row1-row0+1 entries relative
 * to the first row; precision selects the FP64 or FP32 value copy. */
int hd_ham_copy_crs(const hd_hamiltonian* h, int64_t row0, int64_t row1, int32_t precision,
                    int64_t* h_rowptr, int32_t* h_col, void* h_val);

/* Zero-copy view of the device CRS (CrsSparseMatrix::GetRowStarts / GetCols /
 * GetData, crs_sparse_matrix.hpp:297-330): int64 row pointers (dim+1), int32
 * columns, FP64 and FP32 values (NULL when not resident).  Owned by the handle. */
int hd_ham_device_crs(const hd_hamiltonian* h, const int64_t** d_rowptr, const int32_t** d_col,
                      const void** d_val64, const void** d_val32);
/* Zero-copy view of the coded values (no reference counterpart; the coded
 * format is this library's): per-entry uint16 codes, the FP64 diagonal (dim)
 * and the FP64 dictionary (ndict entries).  NULL when not resident. */
int hd_ham_device_coded(const hd_hamiltonian* h, const uint16_t** d_code, const void** d_diag,
                        const void** d_dict, int32_t* ndict);

/* ------------------------------------------------------------ observables */

/* Observables of one eigenvector (struct Observables,
 * src/observables/fermion_observables.hpp:38-626), computed on the device from
 * the basis handle's states and a device vector d_vec of dim entries (double,
 * or interleaved complex when is_complex).  Deterministic reductions. */

/* DensityFunction (fermion_observables.hpp:47-106): h_out[k] = <c+_k c_k>, Norb values. */
int hd_obs_density(const hd_basis* b, int32_t is_complex, const void* d_vec, double* h_out);
/* DensityDensityFunction (:115-198): h_out[k*Norb+q] = <n_k n_q>, Norb^2 values. */
int hd_obs_density_density(const hd_basis* b, int32_t is_complex, const void* d_vec, double* h_out);
/* TranslationalDensityDensityFunction (:216-317): h_out[q] =
 * 1/N sum_{k1,k2} <c+_{k2-q} c_{k2} c+_{k1+q} c_{k1}> for q < nq, with the
 * model's StateAddition / StateSubtraction as tables h_add[k*nq+q] = k+q and
 * h_sub[k*nq+q] = k-q (linear_momentum_constraints_2D.cpp:128-168). */
int hd_obs_translational_dd(const hd_basis* b, int32_t is_complex, const void* d_vec, int32_t nq,
                            const uint16_t* h_add, const uint16_t* h_sub, double* h_out);
/* RotationalDensityDensityFunction (:334-428): h_out[2m], h_out[2m+1] = Re, Im of
 * R(m), m = 0..5, with InverseR60 as a table h_inv_r60[k*6+alpha]
 * (linear_momentum_constraints_2D.cpp:221-264). */
int hd_obs_rotational_dd(const hd_basis* b, int32_t is_complex, const void* d_vec, const uint16_t* h_inv_r60,
                         double* h_out);
/* GetParticipationRatio (:528-563): 1 / sum |psi_i|^4. */
int hd_obs_participation_ratio(const hd_basis* b, int32_t is_complex, const void* d_vec, double* h_out);
/* GetMostProbable (:437-517) for one vector: the n (<= 1024) largest |psi_i|,
 * descending (ties by ascending state index), as states and amplitudes
 * (double, or interleaved complex). */
int hd_obs_most_probable(const hd_basis* b, int32_t is_complex, const void* d_vec, int32_t n, uint64_t* h_states,
                         void* h_amplitudes);

/* ---------------------------------------------------------------- mat-vec */

/* MatrixVectorFunction<T,L1,L2>::Simple (src/hamiltonians/matrix_vector_routines.hpp:94-104):
 * y = H x on device vectors.  precision = HD_F64: FP64 values, x, y (complex
 * FP64 for complex tables).  HD_F32: FP32 values, x, y, FP32 accumulation.
 * HD_CODED, HD_TILED: FP64 x, y.  HD_TILED_X32: FP32 x, FP64 y.  Full storage
 * only.  The call is asynchronous on the handle's stream. */
int hd_spmv(const hd_hamiltonian* h, int32_t precision, const void* d_x, void* d_y);
/* Simple() drop-in over host buffers (FP64 / complex FP64): copies in, multiplies,
 * copies out. */
int hd_spmv_host(const hd_hamiltonian* h, const void* h_x, void* h_y);
/* Simple() over host buffers without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, the reverse-communication loop
 * of an ARPACK driver calling Simple() hundreds of times).  x and y are staged
 * through a fixed 64 MiB device array of the library (32 MiB per vector, held
 * by every GPU the library runs on until the process ends), so dim is at most
 * 4,194,304 (real) or 2,097,152 (complex); larger handles get HD_ERR_ARG.  The
 * handle's device must be current.  All calls on one GPU share that array:
 * calls that overlap on one GPU race (serialise them). */
int hd_spmv_host_noalloc(const hd_hamiltonian* h, const void* h_x, void* h_y);
/* Measurement helper (no reference counterpart): runs `iters` SpMVs on zero
 * vectors allocated by the library and reports the mean kernel time from HIP
 * events recorded on the handle's stream. */
int hd_spmv_time(hd_hamiltonian* h, int32_t precision, int32_t iters, double* avg_ms);
/* Measurement helper (no reference counterpart): `iters` launches of a plain
 * streaming copy kernel d_dst[i] = d_src[i] over `bytes` (a multiple of 16),
 * the calibration reference for hardware byte counters; mean kernel time from
 * HIP events on the default stream. */
int hd_stream_copy(const void* d_src, void* d_dst, size_t bytes, int32_t iters, double* avg_ms);
/* Rows per wavefront schedule of the SpMV (lanes per row 4..64, 0 = autotune). */
int hd_spmv_set_lanes(hd_hamiltonian* h, int32_t lanes_per_row);
int hd_spmv_get_lanes(const hd_hamiltonian* h, int32_t* lanes_per_row);
/* Tiled layout of the coded operator (no reference counterpart; it replaces
 * the column gathers of SymmetricMatrixVectorMultiply,
 * src/utilities/mathematics/sparse_linear_algebra.hpp:218-312, for the
 * product path): built on the device from the coded values, after which
 * precision HD_TILED is available to hd_spmv and hd_lanczos.  Real and complex
 * (Hermitian) matrices; a complex line holds 8 values and its dictionary is read
 * from L2.  Fails with HD_ERR_UNSUPPORTED when a group reads more than 5120
 * distinct 128-byte lines of x, or for 3-byte entries of a complex matrix.  rows_per_group: 128 or 64, or 0 for
 * 128 with a fallback to 64.  entry_bytes: 4 (line slot, word, sign and a
 * 19-bit dictionary index), 3 (the same with an 11-bit index, for
 * dictionaries of at most 2047 values), or 0 for 4. */
int hd_ham_tile(hd_hamiltonian* h, int32_t rows_per_group, int32_t entry_bytes);
typedef struct {
    int64_t groups;         /* row groups */
    int64_t entries;        /* stored entries (entry_bytes each): the off-diagonal
                               entries plus at most 3 padding entries per chunk */
    int64_t lines;          /* staged x lines, summed over groups */
    int64_t chunks;         /* chunks of 256 lines */
    int64_t rows_per_group; /* 128, or 64 when a 128-row group reads too many lines */
    int64_t spmv_bytes;     /* algorithmic HBM bytes of one product */
    double build_ms;
    int64_t entry_bytes;    /* 3 or 4 */
    int64_t block_window;   /* block order of the product: groups per XCD window
                               (0 = launch order), see hd_ham_tile_order */
} hd_tile_info_t;
int hd_ham_tile_info(const hd_hamiltonian* h, hd_tile_info_t* info);
/* Block order of the tiled product (no reference counterpart; a schedule knob
 * like hd_spmv_set_lanes).  Workgroups go to the 8 XCDs in turn and each XCD
 * has its own L2; with groups_per_xcd = w > 0 each window of 8w blocks gives
 * every XCD w consecutive row groups, so neighbouring groups, which read
 * mostly the same x lines, share an L2.  0 keeps the launch order; -1 times
 * the candidates (0, 16, 64, 256, one window over the grid) and keeps the
 * fastest, which hd_ham_tile also does after building the layout. */
int hd_ham_tile_order(hd_hamiltonian* h, int64_t groups_per_xcd);
/* Zero-copy view of the tiled layout: group bases (3 x (groups+1) int64:
 * entries, lines, offset table), entries (uint32 words, or with entry_bytes 3
 * the packed little-endian 3-byte entries: slot bits 16-23, word 12-15, sign
 * 11, index 0-10), line lists, offset tables and the FP64 diagonal.  Owned by
 * the handle. */
int hd_ham_device_tiled(const hd_hamiltonian* h, const int64_t** d_base, const uint32_t** d_entries,
                        const uint32_t** d_lines, const uint32_t** d_offsets, const double** d_diag);
/* The SpMV schedule of one value format: lanes per row and threads per block
 * (each format is autotuned separately; no reference counterpart). */
int hd_spmv_get_schedule(const hd_hamiltonian* h, int32_t precision, int32_t* lanes_per_row, int32_t* threads);

/* ---------------------------------------------------------------- Lanczos */

/* The host eigensolver of the projected matrices (Householder tridiagonal
 * reduction + implicit QL; the role of LAPACK dsyevr in the reference's dense
 * path, lapack_wrapper.hpp:566-649): h_a is an n x n symmetric row-major
 * matrix, h_w receives the eigenvalues ascending, h_z (optional) the
 * eigenvectors as columns, row-major.  Runs on the CPU, no device needed. */
int hd_sym_eig(int32_t n, const double* h_a, double* h_w, double* h_z);

typedef enum {
    HD_START_RANDOM = 0, /* normal(0,1) entries from a counter-based generator */
    HD_START_USER = 1    /* caller's device vector (the reference's --use-initial file) */
} hd_start;

typedef struct {
    int32_t nev;          /* eigenpairs wanted (clamped to dim, hamiltonian_base.hpp:605-623) */
    int32_t ncv;          /* Krylov basis size; 0 = min(dim, 2*nev+21) (arpack_wrapper.hpp:702) */
    int32_t max_restarts; /* 0 = 2000 (arpack_wrapper.cpp:44) */
    double tol;           /* residual ||Hx - lx|| <= tol * max(1,|l|) */
    int32_t start;        /* hd_start */
    uint64_t seed;
    const void* d_start;  /* HD_START_USER: dim FP64 (complex FP64) device vector */
    int32_t precision;    /* HD_F64, HD_CODED, HD_TILED: that exact FP64 operator throughout.
                             HD_F32: mixed; each restart cycle uses the exact FP64 operator
                             (tiled when built) or the FP32-input one (HD_TILED_X32 for real
                             matrices with a tiled layout, else the FP32-value CRS); a cycle
                             takes the FP32 operator only while the rounding it adds to every
                             wanted residual, estimated from the Ritz coefficients on the
                             cycle's new vectors, stays within f32_budget of tol: predicted
                             before the cycle, and checked after it, when a cycle that
                             overshoots is discarded and run again in FP64.  The
                             eigenpairs are checked with the FP64 operator; when a residual
                             misses tol the solve is repeated with the FP64 operator. */
    double f32_budget;    /* HD_F32: share of tol the FP32 cycles may take; 0 = 1 */
    int32_t want_vectors; /* keep the eigenvectors on the device (hd_lanczos_vectors) */
    int32_t block;        /* 0 or 1: single-vector Lanczos.  4, 8 or 16: block Lanczos with
                             that many vectors per SpMM (real matrices, FP64 or coded operator;
                             ncv then counts columns, rounded up to whole blocks) */
    int32_t full_reorth;  /* 0: partial reorthogonalisation (orthogonality estimated by the
                             omega recurrence, the basis kept semi-orthogonal, sqrt(eps));
                             1: every new vector against the whole basis */
} hd_lanczos_opts;

typedef struct {
    int32_t converged;
    int32_t nconv;
    int32_t restarts;
    int32_t matvecs;      /* all operator applications */
    int32_t matvecs_f32;
    int32_t matvecs_f64;
    double residual[64];  /* ||H x - l x|| / max(1,|l|), FP64 operator, per eigenpair */
    double time_total_ms;
    double time_spmv_ms;
    double time_ortho_ms;
    int32_t reorth_steps; /* steps that orthogonalised against the whole basis */
    int32_t f32_cycles_discarded; /* HD_F32: FP32 cycles run again in FP64 (their rounding
                                     estimate exceeded the budget once their coefficients
                                     were known) */
} hd_lanczos_stats;

/* HamiltonianBase::LanczosDiagonalize -> ArpackWrapper::DiagonalizeSymmetricMatrix
 * (src/hamiltonians/hamiltonian_base.hpp:769-931, src/utilities/wrappers/arpack_wrapper.hpp:595-937):
 * the nev lowest eigenvalues (WHICH='SR'), ascending, by a device-resident
 * thick-restart Lanczos with FP64 recurrence and reductions.  By default
 * (full_reorth = 0) the orthogonality of the basis is tracked by Simon's omega
 * recurrence and a new vector is orthogonalised against the whole basis only
 * when an estimate exceeds min(sqrt(eps), tol) (the first and last step of
 * every restart cycle always are); full_reorth = 1 orthogonalises at every
 * step.  Once a cycle starts with every wanted Ritz estimate within 1e4 tol,
 * the Ritz pairs are tested after each step and the solve ends at the first
 * step where all meet tol (not at the end of the cycle).  Every returned
 * eigenpair's residual is recomputed with the FP64
 * operator.  When one misses tol, the solve is repeated once from the start,
 * with the FP64 operator if precision was HD_F32, else with full_reorth = 1;
 * the returned statistics then add up both runs, so such a call can cost
 * about twice the operator applications.  Eigenvalues go to h_evals; with
 * want_vectors the eigenvectors stay resident and are read with
 * hd_lanczos_vectors. */
int hd_lanczos(hd_hamiltonian* h, const hd_lanczos_opts* opts, double* h_evals,
               hd_lanczos_stats* stats);
/* Copies eigenvector i (dim FP64 / complex FP64) of the last hd_lanczos call to
 * device memory d_out, or to host memory when d_out is NULL and h_out is set. */
int hd_lanczos_vectors(const hd_hamiltonian* h, int32_t i, void* d_out, void* h_out);
/* Dense fallback for tiny sectors (HamiltonianBase::FullDiagonalize with LAPACK
 * dsyevr/zheevr, hamiltonian_base.hpp:634-758, lapack_wrapper.hpp:566-649):
 * the full CRS is expanded on the host and diagonalised by Jacobi rotations. */
int hd_full_diagonalize(hd_hamiltonian* h, int32_t nev, double* h_evals);

/* ------------------------------------------------ batched sectors x points */

/* No single reference counterpart: the reference diagonalises one L_z sector
 * per pass of its sector loop (src/fqhe_sphere/main/
 * fqhe_sphere_spinless_fermions_diagonalization.cpp:255-291 ->
 * pseudopotential_model_base.hpp:235-285) and one parameter point per program
 * run.  A batch holds n_sectors bases (same N and orbital count) and n_points
 * real term tables (same layout and k table, different values); every
 * (sector, point) pair is one problem.  The CRS pattern of each sector is built
 * once (coded format, codes shared by all points), with an FP64 dictionary and
 * diagonal per point that are bit-identical to single builds. */
typedef struct hd_batch hd_batch;
int hd_batch_create(const hd_basis* const* bases, int32_t n_sectors, const hd_tables* const* tables,
                    int32_t n_points, double zero_tol, hd_batch** out);
int hd_batch_destroy(hd_batch* b);
/* Rows (sum of the sector dimensions), stored entries, dictionary tuples and
 * the padded point count Pp (16, 32 or 64) of the point-fastest vectors. */
int hd_batch_info(const hd_batch* b, int64_t* rows, int64_t* nnz, int32_t* ndict, int32_t* points_padded);
/* Host copies of the concatenated pattern: row pointers (rows+1), columns
 * (global rows), codes (bit 15 sign, low 15 bits tuple index, 0xFFFF
 * diagonal), the dictionary (ndict x Pp) and the diagonals (rows x Pp).  Any
 * pointer may be NULL. */
int hd_batch_copy_pattern(const hd_batch* b, int64_t* h_rowptr, int32_t* h_col, uint16_t* h_code, double* h_dict,
                          double* h_diag);
/* Y[r][p] = (H_p X)[r] for every point on device arrays of rows x Pp doubles. */
int hd_batch_spmm(hd_batch* b, const double* d_x, double* d_y);
/* Measurement helper: mean time of one batched SpMM over all points. */
int hd_batch_spmm_time(hd_batch* b, int32_t iters, double* avg_ms);

typedef struct {
    int32_t nev;          /* eigenvalues per problem (sectors smaller than nev give NaN beyond their dimension) */
    int32_t ncv;          /* Krylov basis size, 0 = 2*nev+21; at most 64 */
    int32_t max_restarts; /* 0 = 2000 */
    double tol;           /* residual ||Hx - lx|| <= tol * max(1,|l|); 0 = 1e-10 */
    uint64_t seed;        /* random-normal start of every problem */
    int32_t dense_below;  /* sectors below this dimension are diagonalised densely; 0 = 60
                             (BaseDiagonalize, pseudopotential_model_base.hpp:243-252); at most 65 */
    int32_t use_graph;    /* replay the restart cycle as a captured hipGraph */
    int32_t full_reorth;  /* 0: partial reorthogonalisation (per-problem omega estimates; a
                             pass against the whole basis only where one is needed);
                             1: a pass at every step */
} hd_batch_opts;

typedef struct {
    int32_t problems, lanczos_problems, dense_problems, converged;
    int32_t restarts;     /* restart cycles after the first */
    int32_t steps;        /* Lanczos steps (operator applications per problem) */
    double max_residual;  /* largest true residual over the Lanczos problems */
    double time_total_ms;
    int32_t reorth_steps; /* steps whose full pass ran (for at least one problem) */
} hd_batch_stats;

/* The nev lowest eigenvalues of every problem, h_evals[(s*n_points + p)*nev + i];
 * h_resid (optional) the true residuals ||Hx - lx||/max(1,|l|) (0 on the
 * dense path).  HD_ERR_NOT_CONVERGED when a problem misses tol. */
int hd_batch_solve(hd_batch* b, const hd_batch_opts* opts, double* h_evals, double* h_resid, hd_batch_stats* stats);

/* ------------------------------------------------------ learned start vector */

/* No reference counterpart: the reference's warm start is an ARPACK initial
 * vector read from a file (src/utilities/wrappers/arpack_wrapper.hpp:753-774).
 * The amplitude network maps (occupations of n_bits orbitals, n_params
 * Hamiltonian parameters) -> 256 -> 256 -> 256 (SiLU) -> (log|psi|, sign logit),
 * BF16 weights and activations on the MFMA units with FP32 accumulation and an
 * FP32 master copy trained with Adam.  Weights are exchanged as one flat FP32
 * array: per layer the padded weight matrix (rows x cols: 256x64, 256x256,
 * 256x256, 64x256) followed by its bias (rows). */
typedef struct hd_mlp hd_mlp;
int hd_mlp_create(int32_t n_bits, int32_t n_params, uint64_t seed, hd_mlp** out);
/* The same network predicting n_states (1..16) amplitudes: 2 n_states outputs
 * (log|psi_s|, sign logit_s), trained on the n_states lowest eigenvectors;
 * hd_mlp_create is n_states = 1. */
int hd_mlp_create_states(int32_t n_bits, int32_t n_params, int32_t n_states, uint64_t seed, hd_mlp** out);
int hd_mlp_num_states(const hd_mlp* m, int32_t* n_states);
int hd_mlp_destroy(hd_mlp* m);
int hd_mlp_num_weights(const hd_mlp* m, int64_t* count);
int hd_mlp_get_weights(hd_mlp* m, float* h_weights);
int hd_mlp_set_weights(hd_mlp* m, const float* h_weights);
/* Network outputs (n x 2 n_states: log|psi_s|, sign logit_s) for host states
 * and per-sample parameters (n x n_params). */
int hd_mlp_forward(hd_mlp* m, const uint64_t* h_states, const float* h_params, int64_t n, float* h_out);
/* Loss and gradients (flat layout above) without an update, host inputs;
 * targets are n x 3 n_states (per amplitude: log|psi|, sign +-1, weight);
 * lambda weights the sign term. */
int hd_mlp_gradients(hd_mlp* m, const uint64_t* h_states, const float* h_params, const float* h_targets,
                     int64_t n, float lambda, float* loss, float* h_grads);
/* One Adam step on a device batch. */
int hd_mlp_train_step(hd_mlp* m, const uint64_t* d_states, const float* d_params, const float* d_targets,
                      int64_t n, float lambda, float lr, float* loss);
/* The network's amplitudes over a whole sector basis as an FP64 Lanczos start
 * vector (d_vec: dim entries; pass it as hd_lanczos_opts.d_start):
 * exp(o0 - max o0) sign(o1) for one amplitude; with several, the sum of the
 * amplitudes each normalised to unit length, so the vector leans on all of
 * the wanted eigenvectors. */
int hd_mlp_start_vector(hd_mlp* m, const hd_basis* b, const float* h_params, double* d_vec);
/* The MFMA GEMM on host data: C (M x N, FP32) = A (M x K) . B (N x K)^T with
 * BF16 bit patterns; M, N multiples of 64, K a multiple of 16. */
int hd_mlp_gemm_check(int32_t M, int32_t N, int32_t K, const uint16_t* h_a, const uint16_t* h_b, float* h_c);

#ifdef __cplusplus
}
#endif
#endif
