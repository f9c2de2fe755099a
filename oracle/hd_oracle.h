/*
 * hd_oracle.h -- CPU restatement of the reference exact-diagonalization hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, its C++
 * host layer or the Python package) may link, load or call this code.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and
 * only as the checker / the timed CPU baseline.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Pinning status of each function (what
 * independent check fixes it) is listed in the header comment of hd_oracle.c.
 */
#ifndef HD_ORACLE_H
#define HD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Sector predicate kinds (the std::function<bool(fock_t)> constraints used by
 * the reference's callers). */
enum { OC_SECTOR_NONE = 0, OC_SECTOR_LZ = 1, OC_SECTOR_K2D = 2, OC_SECTOR_K1D = 3 };

typedef struct {
    int kind;
    int a, b, c, d; /* LZ: (lzTot, maxLz); K2D: (kx, ky, dimX, dimY); K1D: (ky, dimY) */
} oc_sector;

/* Table layouts (QuarticTermTables flavours). */
enum { OC_LAYOUT_SPHERE = 0, OC_LAYOUT_OFL = 1 };

uint64_t oc_binomial(int n, int k);
int oc_popcount(uint64_t x);
uint64_t oc_next_hamming(uint64_t x);
uint64_t oc_generate_hamming(int weight, uint64_t index);
uint64_t oc_index_hamming(int weight, uint64_t x);
int oc_test_lz(uint64_t state, int lzTot, int maxLz);
int oc_test_k2d(uint64_t state, int kx, int ky, int dimX, int dimY);
int oc_test_k1d(uint64_t state, int ky, int dimY);
int oc_sector_test(const oc_sector* s, uint64_t state);

/* Fock basis in ascending order.  Writes min(dim, cap) states to out (out may
 * be NULL when cap == 0).  Returns the sector dimension. */
int64_t oc_basis(const oc_sector* s, int N, int Norb, uint64_t* out, int64_t cap);
/* Binary search; returns -1 when absent. */
int64_t oc_find_index(const uint64_t* basis, int64_t dim, uint64_t state);

/* Evaluate c+_k1 c+_k2 c_k3 c_k4 and c+_k1 c_k2 on a state (sign 0 => zero). */
uint64_t oc_cdcdcc(int* sign, uint64_t b1, uint64_t b2, uint64_t b3, uint64_t b4,
                   uint64_t state, uint64_t highest);
uint64_t oc_cdc(int* sign, uint64_t b1, uint64_t b2, uint64_t state, uint64_t highest);

/* Sphere k-table (BaseGenerate4KTable). kTable has kMax^3 entries. */
void oc_sphere_ktable(int maxLz, uint16_t* kTable);

/*
 * Upper-triangle CRS build of rows [row0, row1) (Add_CdC_Terms then
 * Add_CdCdCC_Terms, hash accumulation, Trim(1e-14), sort into CRS).
 *  layout      OC_LAYOUT_SPHERE or OC_LAYOUT_OFL
 *  vtab        quartic values (double, or interleaved complex when is_complex)
 *  eps         quadratic (diagonal) energies, kMax entries, or NULL for none
 *  full        0: reference upper triangle (out >= in); 1: both triangles
 *  rowptr      (row1-row0+1) int64, relative to the first row; NULL to count only
 *  col/val     capacity cap; val interleaved complex when is_complex
 * Returns total nnz of the range, or -1 on error (capacity, bad table).
 */
int64_t oc_build(int N, int Norb, const uint64_t* basis, int64_t dim,
                 int64_t row0, int64_t row1,
                 int layout, int kMax, const uint16_t* ktab, const double* vtab,
                 const double* eps, int is_complex, double tol, int full,
                 int64_t* rowptr, int32_t* col, double* val, int64_t cap,
                 int64_t* n_trimmed);

/* The same build with general (hash) term tables: k1 lists per (k2,k3,k4) at
 * entry (k4*K + k3)*K + k2 (nk1 lengths, k1 labels, v values, max_k1 slots per
 * entry) and per k2 for the quadratic terms (nq1 == NULL: none).  Arguments and
 * return as oc_build. */
int64_t oc_build_general(int Norb, const uint64_t* basis, int64_t dim, int64_t row0, int64_t row1, int K,
                         int is_complex, int max_k1, const uint8_t* nk1, const uint16_t* k1, const double* v,
                         int max_q1, const uint8_t* nq1, const uint16_t* q1, const double* e, double tol,
                         int full, int64_t* rowptr, int32_t* col, double* val, int64_t cap, int64_t* n_trimmed);

/* y = U x + (U - D)^H x for the reference's upper-triangle CRS (naive loop). */
void oc_spmv_sym(int64_t dim, const int64_t* rowptr, const int32_t* col,
                 const double* val, int is_complex, const double* x, double* y);
/* Same, parallel over row blocks with nthreads threads (private y copies). */
void oc_spmv_sym_mt(int64_t dim, const int64_t* rowptr, const int32_t* col,
                    const double* val, int is_complex, const double* x, double* y,
                    int nthreads);
/* Rows [row0, row0+nrows) of the naive symmetric product (rowptr relative to
 * row0), accumulated into the full-length y: the per-row work of oc_spmv_sym
 * on a row block, used to time a bounded sample of a large matrix. */
void oc_spmv_sym_block(int64_t row0, int64_t nrows, const int64_t* rowptr, const int32_t* col,
                       const double* val, int is_complex, const double* x, double* y);
/* Row-parallel build over nthreads threads; the same outputs as oc_build
 * (row0 = 0, row1 = dim), used for the CPU baseline timing. */
int64_t oc_build_mt(int N, int Norb, const uint64_t* basis, int64_t dim,
                    int layout, int kMax, const uint16_t* ktab, const double* vtab,
                    const double* eps, int is_complex, double tol, int full,
                    int64_t* rowptr, int32_t* col, double* val, int64_t cap,
                    int nthreads);

/* Two-level (LLL + second Landau level) basis and Hamiltonian: see
 * hd_oracle.c.  Pairs (first, second) sorted lexicographically. */
int oc_test_lz2(uint64_t first, uint64_t second, int lzTot, int max);
int64_t oc_two_level_basis(int N, int n1, int n2, int use_sector, int lzTot, int max, uint64_t* first,
                           uint64_t* second, int64_t cap);
int64_t oc_two_level_build(int n1, int n2, const uint64_t* first, const uint64_t* second, int64_t dim,
                           int64_t row0, int64_t row1, const uint16_t* kt1, const double* vt1, const double* eps1,
                           const uint16_t* kt2, const double* vt2, const double* eps2, double tol, int full,
                           int64_t* rowptr, int32_t* col, double* val, int64_t cap, int64_t* n_trimmed);

/* Flux-lattice V from Bloch coefficients (GenerateQuarticTerms): see hd_oracle.c. */
int oc_ofl_find_k1(int dimX, int dimY, int k2, int k3, int k4, int* gx, int* gy);
void oc_conv2d(const double* right, const double* left, int X, int Y, double* out, int top, int leftOff,
               int bottom, int rightOff);
void oc_ofl_vtable(int dimX, int dimY, int cutX, int cutY, const double* bloch, int64_t e0, int64_t e1,
                   double* out);

/* Hybrid Wannier quartic terms (ChangeToWannierBasis): see hd_oracle.c.
 * kt / vt: folded Bloch-basis tables; u: K x K complex change of basis; slots
 * [s0, s1) of the (k4*K + k3)*K + k2, x1 enumeration. */
void oc_ofl_wannier(int dimX, int dimY, const uint16_t* kt, const double* vt, const double* u, double tol,
                    int64_t s0, int64_t s1, uint16_t* k1_out, double* v_out);

#ifdef __cplusplus
}
#endif
#endif
