/*
 * hd_oracle.c -- CPU restatement of the reference's Fock-basis, Hamiltonian
 * build and symmetric mat-vec stages.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed CPU
 * baseline.  The product never links or loads this file.
 *
 * Reference files restated (paths relative to the reference repository root):
 *   src/utilities/mathematics/binary_number_tools.cpp:95-331   bit tools
 *   src/utilities/mathematics/binomials.cpp:34-64              C(n,k)
 *   src/fqhe_sphere/pseudopotential_model/angular_momentum_constraint.hpp:41-62
 *   src/optical_flux_lattice/interacting_ofl_model/linear_momentum_constraints_2D.cpp:38-110
 *   src/basis/fermion_fock_basis.cpp:134-161, 241-354          basis + lookup
 *   src/utilities/algorithms/binary_search.hpp:50-81
 *   src/utilities/mathematics/fermions.hpp:52-227              c+c+cc, c+c
 *   src/fqhe_sphere/pseudopotential_model/term_tables.cpp:36-169     sphere tables
 *   src/fqhe_sphere/pseudopotential_model/pseudopotential_model_base.hpp:62-139  k-table
 *   src/optical_flux_lattice/interacting_ofl_model/term_tables.cpp:36-198 folded tables
 *   src/hamiltonians/spinless_fermion_hamiltonian.hpp:143-436   the build loops
 *   src/utilities/data_structures/mapped_sparse_matrix.hpp:355-462  Insert / Trim
 *   src/utilities/data_structures/crs_sparse_matrix.hpp:193-254     map -> CRS
 *   src/utilities/mathematics/sparse_linear_algebra.hpp:142-165     naive SpMV
 *
 * Pinning (what fixes each function independently of this file; tests/):
 *   oc_binomial, oc_next_hamming, oc_generate_hamming, oc_index_hamming:
 *       Python math.comb and itertools.combinations order (test_oracle_bits).
 *   oc_test_*, oc_basis: brute force over all bitstrings with the predicate
 *       written from its definition; sector dimensions from an independent
 *       subset-sum count (8,512 / 7,764,392 / 1,001,402; test_oracle_basis).
 *   oc_cdcdcc, oc_cdc: Jordan-Wigner Kronecker-product operator matrices for
 *       Norb <= 8 (test_oracle_fermions).
 *   oc_build: dense H assembled from the Jordan-Wigner matrices for Norb <= 12
 *       (sphere and flux-lattice tables), Hermiticity, trace identity, and the
 *       Laughlin zero mode of V1 at 2S = 3(N-1) (test_oracle_build).
 *   oc_spmv_sym*: dense numpy product on the expanded matrix (test_oracle_build).
 *   The reference itself cannot be compiled here (boost, ARPACK, sparsehash are
 *   absent), so no function is pinned against reference binaries.
 */
#include "hd_oracle.h"

#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- bit tools */

/* C(n,k) exactly in 64 bits (binomials.cpp:34-64 restated; the reference's
 * table is exact for every n <= 64 used on this path). */
uint64_t oc_binomial(int n, int k)
{
    if (k < 0 || n < 0 || k > n) return 0;
    if (k > n - k) k = n - k;
    unsigned __int128 r = 1;
    for (int i = 1; i <= k; ++i) r = r * (unsigned)(n - k + i) / (unsigned)i;
    return (uint64_t)r;
}

int oc_popcount(uint64_t x) { return __builtin_popcountll(x); }

/* NextHammingNumber64 (binary_number_tools.cpp:255-261). */
uint64_t oc_next_hamming(uint64_t x)
{
    if (x == 0) return x;
    uint64_t t = x & (0 - x);
    return (((x ^ (x + t)) / t) >> 2) | (x + t);
}

/* GenerateHammingNumber (binary_number_tools.cpp:297-331): the index-th
 * weight-`weight` number in ascending order. */
uint64_t oc_generate_hamming(int weight, uint64_t index)
{
    if (index == 0) return (weight >= 64) ? ~0ull : ((1ull << weight) - 1ull);
    uint64_t out = 0, cur = 0;
    for (int p = weight - 1; p >= 0; --p) {
        int m = p;
        uint64_t inc;
        do {
            inc = oc_binomial(m, p);
            cur += inc;
            ++m;
        } while (index >= cur);
        cur -= inc;
        out |= 1ull << (m - 1);
    }
    return out;
}

/* IndexHammingNumber64 (binary_number_tools.cpp:270-287). */
uint64_t oc_index_hamming(int weight, uint64_t x)
{
    uint64_t out = 0, t = x;
    for (int p = weight; p > 0; --p) {
        int m = 63 - __builtin_clzll(t);
        if (p <= m) out += oc_binomial(m, p);
        t ^= 1ull << m;
    }
    return out;
}

/* ------------------------------------------------------- sector predicates */

/* TestAngularMomentumSector (angular_momentum_constraint.hpp:41-62). */
int oc_test_lz(uint64_t state, int lzTot, int maxLz)
{
    int16_t cum = 0; /* mState_t accumulator */
    while (state) {
        uint64_t r = state & (0 - state);
        cum = (int16_t)(cum + 2 * (int)__builtin_ctzll(r) - maxLz);
        state ^= r;
    }
    return cum == (int16_t)lzTot;
}

static int oc_mod(int a, int n) { int r = a % n; return r < 0 ? r + n : r; }

/* TestLinearMomentumSector2 (linear_momentum_constraints_2D.cpp:38-79). */
int oc_test_k2d(uint64_t state, int kx, int ky, int dimX, int dimY)
{
    int x = 0, y = 0;
    while (state) {
        uint64_t r = state & (0 - state);
        int pos = __builtin_ctzll(r);
        x += (int)floor((double)pos / dimY);
        y += oc_mod(pos, dimY);
        state ^= r;
    }
    x = oc_mod(x, dimX);
    y = oc_mod(y, dimY);
    return x == kx && y == ky;
}

/* TestLinearMomentumSector1 (linear_momentum_constraints_2D.cpp:89-110). */
int oc_test_k1d(uint64_t state, int ky, int dimY)
{
    int y = 0;
    while (state) {
        uint64_t r = state & (0 - state);
        y += oc_mod(__builtin_ctzll(r), dimY);
        state ^= r;
    }
    return oc_mod(y, dimY) == ky;
}

int oc_sector_test(const oc_sector* s, uint64_t state)
{
    switch (s->kind) {
    case OC_SECTOR_NONE: return 1;
    case OC_SECTOR_LZ: return oc_test_lz(state, s->a, s->b);
    case OC_SECTOR_K2D: return oc_test_k2d(state, s->a, s->b, s->c, s->d);
    case OC_SECTOR_K1D: return oc_test_k1d(state, s->a, s->b);
    }
    return 0;
}

/* ------------------------------------------------------------------- basis */

/* GenerateFockSpace, single rank (fermion_fock_basis.cpp:241-354): walk all
 * C(Norb,N) weight-N numbers in ascending order from GenerateHammingNumber(N,0)
 * with NextHammingNumber64 and keep those passing the predicate. */
int64_t oc_basis(const oc_sector* s, int N, int Norb, uint64_t* out, int64_t cap)
{
    uint64_t full = oc_binomial(Norb, N);
    uint64_t st = oc_generate_hamming(N, 0);
    int64_t cnt = 0;
    for (uint64_t i = 0; i < full; ++i) {
        if (oc_sector_test(s, st)) {
            if (cnt < cap) out[cnt] = st;
            ++cnt;
        }
        st = oc_next_hamming(st);
    }
    return cnt;
}

/* BinarySearch (binary_search.hpp:50-81) as used by FindFockStateIndex. */
int64_t oc_find_index(const uint64_t* a, int64_t dim, uint64_t v)
{
    if (dim <= 0) return -1;
    int64_t lo = 0, hi = dim - 1;
    if (v < a[lo] || v > a[hi]) return -1;
    while (hi >= lo) {
        int64_t mid = (hi + lo) / 2;
        uint64_t m = a[mid];
        if (v == m) return mid;
        if (v > m) lo = mid + 1;
        else {
            if (hi == 0) break;
            hi = mid - 1;
        }
    }
    return -1;
}

/* --------------------------------------------------------- fermion algebra */

/* Evaluate_CdCdCC (fermions.hpp:52-144). */
uint64_t oc_cdcdcc(int* sign, uint64_t b1, uint64_t b2, uint64_t b3, uint64_t b4,
                   uint64_t state, uint64_t highest)
{
    if (b3 > highest || b4 > highest || (state & b3) == 0 || (state & b4) == 0 ||
        (b1 & b2) || (b3 & b4) ||
        (!(b1 & b3) && !(b1 & b4) && (state & b1) != 0) ||
        (!(b2 & b3) && !(b2 & b4) && (state & b2) != 0)) {
        *sign = 0;
        return 0;
    }
    int c = oc_popcount(state & (b4 - 1));
    uint64_t out = state ^ b4;
    c += oc_popcount(out & (b3 - 1));
    out ^= b3;
    c += oc_popcount(out & (b2 - 1));
    out |= b2;
    c += oc_popcount(out & (b1 - 1));
    out |= b1;
    *sign = (c & 1) ? -1 : 1;
    return out;
}

/* Evaluate_CdC (fermions.hpp:155-227). */
uint64_t oc_cdc(int* sign, uint64_t b1, uint64_t b2, uint64_t state, uint64_t highest)
{
    if (b2 > highest || (state & b2) == 0 || (!(b1 & b2) && (state & b1) != 0)) {
        *sign = 0;
        return 0;
    }
    if (b1 & b2) {
        *sign = 1;
        return state;
    }
    int c = oc_popcount(state & (b2 - 1));
    uint64_t out = state ^ b2;
    c += oc_popcount(out & (b1 - 1));
    out |= b1;
    *sign = (c & 1) ? -1 : 1;
    return out;
}

/* ------------------------------------------------------------- term tables */

/* BaseGenerate4KTable + FindM1 + ConvertStateLabel
 * (pseudopotential_model_base.hpp:62-139): m4, m3, m2 descend from maxLz in
 * steps of 2; k = (maxLz - m)/2; an out-of-range m1 stores the label of
 * maxLz + 2, i.e. (uint16)-1. */
void oc_sphere_ktable(int maxLz, uint16_t* kt)
{
    int64_t it = 0;
    for (int m4 = maxLz; m4 >= -maxLz; m4 -= 2)
        for (int m3 = maxLz; m3 >= -maxLz; m3 -= 2)
            for (int m2 = maxLz; m2 >= -maxLz; m2 -= 2, ++it) {
                int m1 = m3 + m4 - m2;
                if (!(m1 <= maxLz && m1 >= -maxLz)) m1 = maxLz + 2;
                kt[it] = (uint16_t)((-m1 + maxLz) / 2);
            }
}

typedef struct {
    int layout, K;
    const uint16_t* kt;
    const double* vt;
    int cplx;
} oc_tables;

/* GetK1: sphere term_tables.cpp:109-124 (count 0 when k1 >= kMax);
 * flux lattice term_tables.cpp:95-112 (folded on k3 >= k4, always one k1). */
static int oc_get_k1(const oc_tables* t, int k2, int k3, int k4, int* k1)
{
    if (t->layout == OC_LAYOUT_SPHERE) {
        int v = t->kt[((int64_t)k4 * t->K + k3) * t->K + k2];
        *k1 = v;
        return v < t->K;
    }
    int t1 = k4, t2 = k3;
    if (k4 > k3) { t1 = k3; t2 = k4; }
    *k1 = t->kt[((int64_t)t1 * t->K + t2 - (t1 * (t1 + 1)) / 2) * t->K + k2];
    return 1;
}

/* GetVkkkk: sphere term_tables.cpp:137-149; flux lattice term_tables.cpp:150-175. */
static int64_t oc_v_index(const oc_tables* t, int k1, int k2, int k3, int k4)
{
    if (t->layout == OC_LAYOUT_SPHERE) return ((int64_t)k4 * t->K + k3) * t->K + k2;
    int t1 = k4, t2 = k3, t3 = k2;
    if (k4 > k3) { t1 = k3; t2 = k4; t3 = k1; }
    return ((int64_t)t1 * t->K + t2 - (t1 * (t1 + 1)) / 2) * t->K + t3;
}

/* ---------------------------------------------------- row accumulator (map) */

/* MappedSparseMatrix::Insert(i,j) += el (mapped_sparse_matrix.hpp:355-370):
 * a per-row open-addressed table keyed by column; values accumulate in
 * generation order starting from 0.0. */
typedef struct {
    int64_t* keys;
    int32_t* slot;  /* index into the entry arrays, -1 empty */
    int64_t* ecol;
    double complex* eval;
    int n, cap, mask;
} oc_rowacc;

static void acc_init(oc_rowacc* a)
{
    a->cap = 1 << 14;
    a->mask = (1 << 15) - 1;
    a->keys = (int64_t*)malloc(sizeof(int64_t) * (a->mask + 1));
    a->slot = (int32_t*)malloc(sizeof(int32_t) * (a->mask + 1));
    a->ecol = (int64_t*)malloc(sizeof(int64_t) * a->cap);
    a->eval = (double complex*)malloc(sizeof(double complex) * a->cap);
    for (int i = 0; i <= a->mask; ++i) a->slot[i] = -1;
    a->n = 0;
}
static void acc_free(oc_rowacc* a)
{
    free(a->keys); free(a->slot); free(a->ecol); free(a->eval);
}
static void acc_clear(oc_rowacc* a)
{
    for (int e = 0; e < a->n; ++e) {
        /* clear only the used slots */
        uint64_t h = ((uint64_t)a->ecol[e] * 0x9E3779B97F4A7C15ull) >> 49;
        int s = (int)(h & a->mask);
        while (a->slot[s] != -1) {
            a->slot[s] = -1;
            s = (s + 1) & a->mask;
        }
    }
    a->n = 0;
}
static int acc_add(oc_rowacc* a, int64_t col, double complex v)
{
    uint64_t h = ((uint64_t)col * 0x9E3779B97F4A7C15ull) >> 49;
    int s = (int)(h & a->mask);
    while (a->slot[s] != -1) {
        if (a->keys[s] == col) {
            a->eval[a->slot[s]] += v;
            return 0;
        }
        s = (s + 1) & a->mask;
    }
    if (a->n >= a->cap) return -1;
    a->keys[s] = col;
    a->slot[s] = a->n;
    a->ecol[a->n] = col;
    a->eval[a->n] = 0.0 + v; /* Insert() default-constructs 0, then += el */
    a->n++;
    return 0;
}

static double complex* acc_find(oc_rowacc* a, int64_t col)
{
    uint64_t h = ((uint64_t)col * 0x9E3779B97F4A7C15ull) >> 49;
    int s = (int)(h & a->mask);
    while (a->slot[s] != -1) {
        if (a->keys[s] == col) return &a->eval[a->slot[s]];
        s = (s + 1) & a->mask;
    }
    return NULL;
}

typedef struct { int64_t col; double complex v; } oc_entry;
static int cmp_entry(const void* x, const void* y)
{
    int64_t a = ((const oc_entry*)x)->col, b = ((const oc_entry*)y)->col;
    return (a > b) - (a < b);
}

/* ------------------------------------------------------------------- build */

/* One row of Add_CdC_Terms + Add_CdCdCC_Terms
 * (spinless_fermion_hamiltonian.hpp:143-257 and 281-403). */
static int oc_build_row(oc_rowacc* acc, int N, int Norb, const uint64_t* basis, int64_t dim,
                        int64_t row, const oc_tables* T, const double* eps, double tol,
                        int full)
{
    (void)N;
    const uint64_t in = basis[row];
    const uint64_t highest = 1ull << (Norb - 1);
    if (eps) {
        /* Quadratic terms: k1 = GetK1(k2) = k2, so only diagonal elements. */
        uint64_t t2 = in;
        while (t2) {
            uint64_t b2 = t2 & (0 - t2);
            int k2 = oc_popcount(b2 - 1);
            double complex el = T->cplx ? (eps[2 * k2] + I * eps[2 * k2 + 1]) : eps[k2];
            if (cabs(el) > tol && acc_add(acc, row, el)) return -1;
            t2 ^= b2;
        }
    }
    const uint64_t mask = (1ull << Norb) - 1;
    uint64_t t4 = in;
    while (t4) {
        uint64_t b4 = t4 & (0 - t4);
        int k4 = oc_popcount(b4 - 1);
        uint64_t t3 = in ^ b4;
        while (t3) {
            uint64_t b3 = t3 & (0 - t3);
            int k3 = oc_popcount(b3 - 1);
            if (k3 != k4) {
                uint64_t t2 = mask ^ (t3 ^ b3);
                while (t2) {
                    uint64_t b2 = t2 & (0 - t2);
                    int k2 = oc_popcount(b2 - 1);
                    int k1;
                    if (oc_get_k1(T, k2, k3, k4, &k1)) {
                        uint64_t b1 = 1ull << k1;
                        if ((!(b1 & in) || ((b1 & b3) || (b1 & b4))) &&
                            (!(b2 & in) || ((b2 & b3) || (b2 & b4))) && !(b1 & b2)) {
                            int sign;
                            uint64_t out = oc_cdcdcc(&sign, b1, b2, b3, b4, in, highest);
                            if (sign != 0 && (full || out >= in)) {
                                int64_t idx = oc_find_index(basis, dim, out);
                                if (idx >= 0) {
                                    int64_t vi = oc_v_index(T, k1, k2, k3, k4);
                                    double complex v = T->cplx
                                        ? (T->vt[2 * vi] + I * T->vt[2 * vi + 1])
                                        : T->vt[vi];
                                    double complex el = (double)sign * v;
                                    if (cabs(el) > tol && acc_add(acc, idx, el)) return -1;
                                }
                            }
                        }
                    }
                    t2 ^= b2;
                }
            }
            t3 ^= b3;
        }
        t4 ^= b4;
    }
    return 0;
}

typedef struct {
    int N, Norb;
    const uint64_t* basis;
    int64_t dim;
    oc_tables T;
    const double* eps;
    double tol;
    int full;
} oc_build_ctx;

/* Trim (mapped_sparse_matrix.hpp:423-462, the erase-current form: an entry is
 * dropped when |v| < tol) and sort by column (crs_sparse_matrix.hpp:193-254).
 * Returns the number of kept entries written to (col, val) starting at off. */
static int64_t oc_emit_row(oc_rowacc* acc, oc_entry* scratch, int cplx, double tol,
                           int32_t* col, double* val, int64_t off, int64_t cap,
                           int64_t* n_trim)
{
    int64_t k = 0;
    for (int e = 0; e < acc->n; ++e) {
        if (cabs(acc->eval[e]) < tol) {
            if (n_trim) (*n_trim)++;
            continue;
        }
        scratch[k].col = acc->ecol[e];
        scratch[k].v = acc->eval[e];
        ++k;
    }
    qsort(scratch, (size_t)k, sizeof(oc_entry), cmp_entry);
    if (col) {
        if (off + k > cap) return -1;
        for (int64_t e = 0; e < k; ++e) {
            col[off + e] = (int32_t)scratch[e].col;
            if (cplx) {
                val[2 * (off + e)] = creal(scratch[e].v);
                val[2 * (off + e) + 1] = cimag(scratch[e].v);
            } else {
                val[off + e] = creal(scratch[e].v);
            }
        }
    }
    return k;
}

static int oc_table_ok(int layout, int kMax, int Norb)
{
    if (kMax < Norb) return 0;
    if (layout != OC_LAYOUT_SPHERE && layout != OC_LAYOUT_OFL) return 0;
    return 1;
}

int64_t oc_build(int N, int Norb, const uint64_t* basis, int64_t dim,
                 int64_t row0, int64_t row1,
                 int layout, int kMax, const uint16_t* ktab, const double* vtab,
                 const double* eps, int is_complex, double tol, int full,
                 int64_t* rowptr, int32_t* col, double* val, int64_t cap,
                 int64_t* n_trimmed)
{
    if (!oc_table_ok(layout, kMax, Norb) || row0 < 0 || row1 > dim || row0 > row1) return -1;
    oc_tables T = {layout, kMax, ktab, vtab, is_complex};
    oc_rowacc acc;
    acc_init(&acc);
    oc_entry* scratch = (oc_entry*)malloc(sizeof(oc_entry) * acc.cap);
    int64_t nnz = 0;
    if (n_trimmed) *n_trimmed = 0;
    if (rowptr) rowptr[0] = 0;
    for (int64_t r = row0; r < row1; ++r) {
        acc_clear(&acc);
        if (oc_build_row(&acc, N, Norb, basis, dim, r, &T, eps, tol, full)) {
            nnz = -1;
            break;
        }
        int64_t k = oc_emit_row(&acc, scratch, is_complex, tol, rowptr ? col : NULL, val,
                                nnz, cap, n_trimmed);
        if (k < 0) { nnz = -1; break; }
        nnz += k;
        if (rowptr) rowptr[r - row0 + 1] = nnz;
    }
    free(scratch);
    acc_free(&acc);
    return nnz;
}

/* ------------------------------------------------- general (hash) tables */

/* The tables of TermHashTables (term_hash_tables_base.hpp:44-276) as the lists
 * that QuarticTermHashTables::GetK1 / QuadraticTermHashTables::GetK1 return
 * (term_hash_tables.cpp:51-120), with GetVkkkk / GetEkk per list slot. */
typedef struct {
    int K, cplx, max_k1, max_q1;
    const uint8_t* nk1;  /* [(k4*K + k3)*K + k2] */
    const uint16_t* k1;
    const double* v;
    const uint8_t* nq1;  /* NULL: no quadratic terms */
    const uint16_t* q1;
    const double* e;
} oc_general;

static double complex oc_gval(const double* a, int64_t i, int cplx)
{
    return cplx ? (a[2 * i] + I * a[2 * i + 1]) : a[i];
}

/* One row of Add_CdC_Terms (spinless_fermion_hamiltonian.hpp:172-228) and
 * Add_CdCdCC_Terms (:316-398) with k1 lists: every term goes into the row's
 * map in loop order; an out state missing from the basis is skipped. */
static int oc_build_row_general(oc_rowacc* acc, int Norb, const uint64_t* basis, int64_t dim, int64_t row,
                                const oc_general* G, double tol, int full)
{
    const uint64_t in = basis[row];
    const uint64_t highest = 1ull << (Norb - 1);
    if (G->nq1) {
        uint64_t t2 = in;
        while (t2) {
            uint64_t b2 = t2 & (0 - t2);
            int k2 = oc_popcount(b2 - 1);
            for (int j = 0; j < G->nq1[k2]; ++j) {
                int64_t slot = (int64_t)k2 * G->max_q1 + j;
                int k1 = G->q1[slot];
                double complex el = 0.0;
                int64_t idx = row;
                if (k1 == k2) {
                    el = oc_gval(G->e, slot, G->cplx);
                } else {
                    uint64_t b1 = 1ull << k1;
                    if (b1 & ~in) {
                        int sign;
                        uint64_t out = oc_cdc(&sign, b1, b2, in, highest);
                        if (sign != 0 && (full || out >= in)) {
                            el = oc_gval(G->e, slot, G->cplx) * (double)sign;
                            idx = oc_find_index(basis, dim, out);
                        }
                    }
                }
                if (idx >= 0 && cabs(el) > tol && acc_add(acc, idx, el)) return -1;
            }
            t2 ^= b2;
        }
    }
    const uint64_t mask = Norb >= 64 ? ~0ull : (1ull << Norb) - 1;
    uint64_t t4 = in;
    while (t4) {
        uint64_t b4 = t4 & (0 - t4);
        int k4 = oc_popcount(b4 - 1);
        uint64_t t3 = in ^ b4;
        while (t3) {
            uint64_t b3 = t3 & (0 - t3);
            int k3 = oc_popcount(b3 - 1);
            if (k3 != k4) {
                uint64_t t2 = mask ^ (t3 ^ b3);
                while (t2) {
                    uint64_t b2 = t2 & (0 - t2);
                    int k2 = oc_popcount(b2 - 1);
                    int64_t ent = ((int64_t)k4 * G->K + k3) * G->K + k2;
                    for (int j = 0; j < G->nk1[ent]; ++j) {
                        int64_t slot = ent * G->max_k1 + j;
                        uint64_t b1 = 1ull << G->k1[slot];
                        if ((!(b1 & in) || ((b1 & b3) || (b1 & b4))) &&
                            (!(b2 & in) || ((b2 & b3) || (b2 & b4))) && !(b1 & b2)) {
                            int sign;
                            uint64_t out = oc_cdcdcc(&sign, b1, b2, b3, b4, in, highest);
                            if (sign != 0 && (full || out >= in)) {
                                int64_t idx = oc_find_index(basis, dim, out);
                                if (idx >= 0) {
                                    double complex el = (double)sign * oc_gval(G->v, slot, G->cplx);
                                    if (cabs(el) > tol && acc_add(acc, idx, el)) return -1;
                                }
                            }
                        }
                    }
                    t2 ^= b2;
                }
            }
            t3 ^= b3;
        }
        t4 ^= b4;
    }
    return 0;
}

int64_t oc_build_general(int Norb, const uint64_t* basis, int64_t dim, int64_t row0, int64_t row1, int K,
                         int is_complex, int max_k1, const uint8_t* nk1, const uint16_t* k1, const double* v,
                         int max_q1, const uint8_t* nq1, const uint16_t* q1, const double* e, double tol,
                         int full, int64_t* rowptr, int32_t* col, double* val, int64_t cap, int64_t* n_trimmed)
{
    if (K < Norb || K > 64 || row0 < 0 || row1 > dim || row0 > row1) return -1;
    oc_general G = {K, is_complex, max_k1, max_q1, nk1, k1, v, nq1, q1, e};
    oc_rowacc acc;
    acc_init(&acc);
    oc_entry* scratch = (oc_entry*)malloc(sizeof(oc_entry) * acc.cap);
    int64_t nnz = 0;
    if (n_trimmed) *n_trimmed = 0;
    if (rowptr) rowptr[0] = 0;
    for (int64_t r = row0; r < row1; ++r) {
        acc_clear(&acc);
        if (oc_build_row_general(&acc, Norb, basis, dim, r, &G, tol, full)) {
            nnz = -1;
            break;
        }
        int64_t k = oc_emit_row(&acc, scratch, is_complex, tol, rowptr ? col : NULL, val, nnz, cap, n_trimmed);
        if (k < 0) { nnz = -1; break; }
        nnz += k;
        if (rowptr) rowptr[r - row0 + 1] = nnz;
    }
    free(scratch);
    acc_free(&acc);
    return nnz;
}

/* ------------------------------------------------------- multithreaded build */

typedef struct {
    oc_build_ctx* ctx;
    int64_t r0, r1;
    int64_t* counts; /* per row, relative to r0 in pass 1 */
    int64_t* rowptr; int32_t* col; double* val;
    int pass, err;
} oc_build_job;

static void* oc_build_worker(void* p)
{
    oc_build_job* j = (oc_build_job*)p;
    oc_build_ctx* c = j->ctx;
    oc_rowacc acc;
    acc_init(&acc);
    oc_entry* scratch = (oc_entry*)malloc(sizeof(oc_entry) * acc.cap);
    for (int64_t r = j->r0; r < j->r1; ++r) {
        acc_clear(&acc);
        if (oc_build_row(&acc, c->N, c->Norb, c->basis, c->dim, r, &c->T, c->eps, c->tol,
                         c->full)) { j->err = 1; break; }
        if (j->pass == 0) {
            j->counts[r] = oc_emit_row(&acc, scratch, c->T.cplx, c->tol, NULL, NULL, 0, 0, NULL);
        } else {
            oc_emit_row(&acc, scratch, c->T.cplx, c->tol, j->col, j->val, j->rowptr[r],
                        j->rowptr[r + 1], NULL);
        }
    }
    free(scratch);
    acc_free(&acc);
    return NULL;
}

int64_t oc_build_mt(int N, int Norb, const uint64_t* basis, int64_t dim,
                    int layout, int kMax, const uint16_t* ktab, const double* vtab,
                    const double* eps, int is_complex, double tol, int full,
                    int64_t* rowptr, int32_t* col, double* val, int64_t cap,
                    int nthreads)
{
    if (!oc_table_ok(layout, kMax, Norb)) return -1;
    if (nthreads < 1) nthreads = 1;
    oc_build_ctx ctx = {N, Norb, basis, dim, {layout, kMax, ktab, vtab, is_complex}, eps, tol, full};
    int64_t* counts = (int64_t*)calloc((size_t)dim + 1, sizeof(int64_t));
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    oc_build_job* jobs = (oc_build_job*)calloc((size_t)nthreads, sizeof(oc_build_job));
    int64_t nnz = 0;
    for (int pass = 0; pass < 2; ++pass) {
        for (int t = 0; t < nthreads; ++t) {
            jobs[t].ctx = &ctx;
            /* interleaved blocks of 256 rows would balance better; contiguous
             * blocks keep the code simple and the timing conservative */
            jobs[t].r0 = dim * t / nthreads;
            jobs[t].r1 = dim * (t + 1) / nthreads;
            jobs[t].counts = counts;
            jobs[t].rowptr = rowptr; jobs[t].col = col; jobs[t].val = val;
            jobs[t].pass = pass;
            jobs[t].err = 0;
            pthread_create(&th[t], NULL, oc_build_worker, &jobs[t]);
        }
        int err = 0;
        for (int t = 0; t < nthreads; ++t) {
            pthread_join(th[t], NULL);
            err |= jobs[t].err;
        }
        if (err) { nnz = -1; break; }
        if (pass == 0) {
            rowptr[0] = 0;
            for (int64_t r = 0; r < dim; ++r) rowptr[r + 1] = rowptr[r] + counts[r];
            nnz = rowptr[dim];
            if (nnz > cap || !col) break;
        }
    }
    free(jobs); free(th); free(counts);
    return nnz;
}

/* -------------------------------------------------------------------- SpMV */

/* SymmetricMatrixVectorMultiply, non-MKL branch with ALPHA=1, BETA=0
 * (sparse_linear_algebra.hpp:142-165): y = U x + (U - D)^H x. */
static void oc_spmv_rows(int64_t r0, int64_t r1, const int64_t* rp, const int32_t* col,
                         const double* val, int cplx, const double* x, double* y)
{
    if (!cplx) {
        for (int64_t i = r0; i < r1; ++i)
            for (int64_t j = rp[i]; j < rp[i + 1]; ++j) {
                int c = col[j];
                double v = val[j];
                y[i] += v * x[c];
                if (i != c) y[c] += v * x[i];
            }
    } else {
        const double complex* X = (const double complex*)x;
        double complex* Y = (double complex*)y;
        const double complex* V = (const double complex*)val;
        for (int64_t i = r0; i < r1; ++i)
            for (int64_t j = rp[i]; j < rp[i + 1]; ++j) {
                int c = col[j];
                double complex v = V[j];
                Y[i] += v * X[c];
                if (i != c) Y[c] += conj(v) * X[i];
            }
    }
}

void oc_spmv_sym(int64_t dim, const int64_t* rowptr, const int32_t* col,
                 const double* val, int is_complex, const double* x, double* y)
{
    memset(y, 0, sizeof(double) * (size_t)dim * (is_complex ? 2 : 1));
    oc_spmv_rows(0, dim, rowptr, col, val, is_complex, x, y);
}

void oc_spmv_sym_block(int64_t row0, int64_t nrows, const int64_t* rowptr, const int32_t* col,
                       const double* val, int is_complex, const double* x, double* y)
{
    /* rowptr/col/val hold rows [row0, row0+nrows) with rowptr relative to the
     * first of them; the same loop as oc_spmv_rows with the row index shifted */
    if (!is_complex) {
        for (int64_t r = 0; r < nrows; ++r) {
            const int64_t i = row0 + r;
            for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
                int c = col[j];
                double v = val[j];
                y[i] += v * x[c];
                if (i != c) y[c] += v * x[i];
            }
        }
    } else {
        const double complex* X = (const double complex*)x;
        double complex* Y = (double complex*)y;
        const double complex* V = (const double complex*)val;
        for (int64_t r = 0; r < nrows; ++r) {
            const int64_t i = row0 + r;
            for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
                int c = col[j];
                double complex v = V[j];
                Y[i] += v * X[c];
                if (i != c) Y[c] += conj(v) * X[i];
            }
        }
    }
}

typedef struct {
    int64_t r0, r1, dim;
    const int64_t* rp; const int32_t* col; const double* val;
    int cplx; const double* x; double* y;
} oc_spmv_job;

static void* oc_spmv_worker(void* p)
{
    oc_spmv_job* j = (oc_spmv_job*)p;
    memset(j->y, 0, sizeof(double) * (size_t)j->dim * (j->cplx ? 2 : 1));
    oc_spmv_rows(j->r0, j->r1, j->rp, j->col, j->val, j->cplx, j->x, j->y);
    return NULL;
}

void oc_spmv_sym_mt(int64_t dim, const int64_t* rowptr, const int32_t* col,
                    const double* val, int is_complex, const double* x, double* y,
                    int nthreads)
{
    if (nthreads <= 1) {
        oc_spmv_sym(dim, rowptr, col, val, is_complex, x, y);
        return;
    }
    size_t w = (size_t)dim * (is_complex ? 2 : 1);
    double* priv = (double*)malloc(sizeof(double) * w * (size_t)nthreads);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    oc_spmv_job* jobs = (oc_spmv_job*)malloc(sizeof(oc_spmv_job) * nthreads);
    /* balance by nnz: each thread takes a contiguous block of ~nnz/nthreads */
    int64_t nnz = rowptr[dim], r = 0;
    for (int t = 0; t < nthreads; ++t) {
        int64_t target = nnz * (t + 1) / nthreads;
        int64_t r0 = r;
        while (r < dim && rowptr[r] < target) ++r;
        if (t == nthreads - 1) r = dim;
        jobs[t] = (oc_spmv_job){r0, r, dim, rowptr, col, val, is_complex, x, priv + w * t};
        pthread_create(&th[t], NULL, oc_spmv_worker, &jobs[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    for (size_t i = 0; i < w; ++i) {
        double s = 0.0;
        for (int t = 0; t < nthreads; ++t) s += priv[w * t + i];
        y[i] = s;
    }
    free(priv); free(th); free(jobs);
}

/* ------------------------------------------------------------- two levels */

/* TestAngularMomentumSector2LL (two_level_angular_momentum_constraint.hpp:41-73):
 * first-level orbital k has m = 2k - (max - 2), second-level orbital k has
 * m = 2k - max. */
int oc_test_lz2(uint64_t first, uint64_t second, int lzTot, int max)
{
    const int maxLLL = max - 2;
    int lz = 0;
    for (uint64_t t = first; t; t &= t - 1) lz += 2 * __builtin_ctzll(t) - maxLLL;
    for (uint64_t t = second; t; t &= t - 1) lz += 2 * __builtin_ctzll(t) - max;
    return lz == lzTot;
}

static int cmp_pair(const void* x, const void* y)
{
    const uint64_t* a = (const uint64_t*)x;
    const uint64_t* b = (const uint64_t*)y;
    if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
    if (a[1] != b[1]) return a[1] < b[1] ? -1 : 1;
    return 0;
}

/* TwoLevelFermionFockBasis::GenerateFockSpace (two_level_fermion_fock_basis.cpp:214-337):
 * for n0 = 0..N particles in the first level, Gosper over the first level's
 * n1 orbitals (outer) and the second level's n2 (inner), filtered by the
 * sector test when use_sector, then sorted by (first, second).  Writes up to
 * cap pairs; returns the dimension. */
int64_t oc_two_level_basis(int N, int n1, int n2, int use_sector, int lzTot, int max, uint64_t* first,
                           uint64_t* second, int64_t cap)
{
    int64_t cnt = 0;
    uint64_t* buf = NULL;
    int64_t bcap = 0;
    for (int n0 = 0; n0 <= N; ++n0) {
        if (n0 > n1 || N - n0 > n2) continue;
        uint64_t d1 = oc_binomial(n1, n0), d2 = oc_binomial(n2, N - n0);
        uint64_t s1 = oc_generate_hamming(n0, 0);
        for (uint64_t i = 0; i < d1; ++i) {
            uint64_t s2 = oc_generate_hamming(N - n0, 0);
            for (uint64_t j = 0; j < d2; ++j) {
                if (!use_sector || oc_test_lz2(s1, s2, lzTot, max)) {
                    if (cnt >= bcap) {
                        bcap = bcap ? 2 * bcap : 1024;
                        buf = (uint64_t*)realloc(buf, sizeof(uint64_t) * 2 * bcap);
                    }
                    buf[2 * cnt] = s1;
                    buf[2 * cnt + 1] = s2;
                    ++cnt;
                }
                s2 = oc_next_hamming(s2);
            }
            s1 = oc_next_hamming(s1);
        }
    }
    if (cnt) qsort(buf, (size_t)cnt, 2 * sizeof(uint64_t), cmp_pair);
    for (int64_t i = 0; i < cnt && i < cap; ++i) {
        first[i] = buf[2 * i];
        second[i] = buf[2 * i + 1];
    }
    free(buf);
    return cnt;
}

/* FindFockStateIndex (two_level_fermion_fock_basis.cpp:155-185): binary search
 * in the (first, second)-sorted list. */
static int64_t oc_find_pair(const uint64_t* f, const uint64_t* s, int64_t dim, uint64_t a, uint64_t b)
{
    int64_t lo = 0, hi = dim - 1;
    while (lo <= hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (f[mid] == a && s[mid] == b) return mid;
        if (f[mid] < a || (f[mid] == a && s[mid] < b)) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

/* One row of TwoLevelSpinlessFermionHamiltonian: Add_CdC_Terms for level 0
 * then level 1 (when eps given), Add_CdCdCC_Terms for level 0 then level 1
 * (two_level_spinless_fermion_hamiltonian.hpp:128-262, 280-480), with the
 * model's call order (two_level_pseudopotential_model.cpp:188-200).  Sphere
 * tables per level: index (k4*kMax+k3)*kMax+k2, k1 >= kMax = no term. */
static int oc_two_level_row(oc_rowacc* acc, int n1, int n2, int Ntot, const uint64_t* F, const uint64_t* S,
                            int64_t dim, int64_t row, const uint16_t* kt[2], const double* vt[2],
                            const double* eps[2], double tol, int full)
{
    const int nl[2] = {n1, n2};
    const uint64_t highest = 1ull << (Ntot - 1);
    for (int L = 0; L < 2; ++L) {
        if (!eps[L]) continue;
        const uint64_t in = L ? S[row] : F[row];
        for (uint64_t t = in; t; t &= t - 1) {
            const int k2 = __builtin_ctzll(t);
            const double el = eps[L][k2];
            if (fabs(el) > tol && acc_add(acc, row, el)) return -1;
        }
    }
    for (int L = 0; L < 2; ++L) {
        const int K = nl[L];
        const uint64_t in = L ? S[row] : F[row];
        const uint64_t other = L ? F[row] : S[row];
        const uint64_t mask = (1ull << K) - 1;
        for (uint64_t t4 = in; t4; t4 &= t4 - 1) {
            const uint64_t b4 = t4 & (0 - t4);
            const int k4 = __builtin_ctzll(b4);
            uint64_t t3 = in ^ b4;
            while (t3) {
                const uint64_t b3 = t3 & (0 - t3);
                const int k3 = __builtin_ctzll(b3);
                if (k3 != k4) {
                    uint64_t t2 = mask ^ (t3 ^ b3);
                    while (t2) {
                        const uint64_t b2 = t2 & (0 - t2);
                        const int k2 = __builtin_ctzll(b2);
                        const int64_t vi = ((int64_t)k4 * K + k3) * K + k2;
                        const int k1 = kt[L][vi];
                        if (k1 < K) {
                            const uint64_t b1 = 1ull << k1;
                            if ((!(b1 & in) || ((b1 & b3) || (b1 & b4))) &&
                                (!(b2 & in) || ((b2 & b3) || (b2 & b4))) && !(b1 & b2)) {
                                int sign;
                                const uint64_t out = oc_cdcdcc(&sign, b1, b2, b3, b4, in, highest);
                                if (sign != 0 && (full || out >= in)) {
                                    const int64_t idx = L ? oc_find_pair(F, S, dim, other, out)
                                                          : oc_find_pair(F, S, dim, out, other);
                                    if (idx >= 0) {
                                        const double el = (double)sign * vt[L][vi];
                                        if (fabs(el) > tol && acc_add(acc, idx, el)) return -1;
                                    }
                                }
                            }
                        }
                        t2 ^= b2;
                    }
                }
                t3 ^= b3;
            }
        }
        if (L == 0) {
            /* Add_CdCdCC_Terms ends with Trim (two_level_spinless_fermion_hamiltonian.hpp:458-467):
             * after the first level only the diagonal can still receive
             * terms, so dropping it when |v| < tol restarts its sum at 0. */
            double complex* d = acc_find(acc, row);
            if (d && cabs(*d) < tol) *d = 0.0;
        }
    }
    return 0;
}

/* The two-level upper-triangle (or full) CRS for rows [row0, row1), real
 * values; Trim and column order as oc_build. */
int64_t oc_two_level_build(int n1, int n2, const uint64_t* first, const uint64_t* second, int64_t dim,
                           int64_t row0, int64_t row1, const uint16_t* kt1, const double* vt1, const double* eps1,
                           const uint16_t* kt2, const double* vt2, const double* eps2, double tol, int full,
                           int64_t* rowptr, int32_t* col, double* val, int64_t cap, int64_t* n_trimmed)
{
    if (row0 < 0 || row1 > dim || row0 > row1 || n1 <= 0 || n2 <= 0 || n1 + n2 > 64) return -1;
    const uint16_t* kt[2] = {kt1, kt2};
    const double* vt[2] = {vt1, vt2};
    const double* eps[2] = {eps1, eps2};
    oc_rowacc acc;
    acc_init(&acc);
    oc_entry* scratch = (oc_entry*)malloc(sizeof(oc_entry) * acc.cap);
    int64_t nnz = 0;
    if (n_trimmed) *n_trimmed = 0;
    if (rowptr) rowptr[0] = 0;
    for (int64_t r = row0; r < row1; ++r) {
        acc_clear(&acc);
        if (oc_two_level_row(&acc, n1, n2, n1 + n2, first, second, dim, r, kt, vt, eps, tol, full)) {
            nnz = -1;
            break;
        }
        int64_t k = oc_emit_row(&acc, scratch, 0, tol, rowptr ? col : NULL, val, nnz, cap, n_trimmed);
        if (k < 0) {
            nnz = -1;
            break;
        }
        nnz += k;
        if (rowptr) rowptr[r - row0 + 1] = nnz;
    }
    free(scratch);
    acc_free(&acc);
    return nnz;
}

/* ------------------------------------------------ flux-lattice V from Bloch */

/* FindK1 (interacting_ofl_model.cpp:374-407): k1 = k3 + k4 - k2 per grid
 * component, folded into the zone; g = (folded - unfolded) / dim per component
 * (exact: the difference is a multiple of dim). */
int oc_ofl_find_k1(int dimX, int dimY, int k2, int k3, int k4, int* gx, int* gy)
{
    const int fx = k3 / dimY + k4 / dimY - k2 / dimY;
    const int fy = k3 % dimY + k4 % dimY - k2 % dimY;
    const int sx = ((fx % dimX) + dimX) % dimX, sy = ((fy % dimY) + dimY) % dimY;
    *gx = (sx - fx) / dimX;
    *gy = (sy - fy) / dimY;
    return sx * dimY + sy;
}

typedef struct { double re, im; } oc_c;

static inline oc_c c_mul(oc_c a, oc_c b)
{
    /* GCC's inline complex product without -ffast-math: (ac - bd, ad + bc),
     * each product rounded before the sum (no contraction on x86-64 -O3; the
     * oracle is built with -ffp-contract=off) */
    const double ac = a.re * b.re, bd = a.im * b.im, ad = a.re * b.im, bc = a.im * b.re;
    oc_c r = {ac - bd, ad + bc};
    return r;
}

/* KahanAccumulation<dcmplx> += (kahan_arithmetic.hpp:67-76), componentwise */
static inline void kahan_add(oc_c* sum, oc_c* corr, oc_c b)
{
    const double tv_re = b.re - corr->re, tv_im = b.im - corr->im;
    const double ts_re = sum->re + tv_re, ts_im = sum->im + tv_im;
    corr->re = (ts_re - sum->re) - tv_re;
    corr->im = (ts_im - sum->im) - tv_im;
    sum->re = ts_re;
    sum->im = ts_im;
}

/* ArrayConvolution2D<dcmplx,dcmplx> (array_convolution.hpp:136-230): output
 * (i, j) of the (2X-1) x (2Y-1) grid is the Kahan sum over the overlap, x
 * outer and y inner, of left[..]*right[..] with the four index branches of the
 * reference; outside [top, 2X-1-bottom) x [leftOff, 2Y-1-rightOff) it is 0. */
void oc_conv2d(const double* right_, const double* left_, int X, int Y, double* out_, int top, int leftOff,
               int bottom, int rightOff)
{
    const oc_c* right = (const oc_c*)right_;
    const oc_c* left = (const oc_c*)left_;
    oc_c* out = (oc_c*)out_;
    const int cX = 2 * X - 1, cY = 2 * Y - 1;
    for (int i = 0; i < cX; ++i) {
        const int kX = i + 1 < cX - i ? i + 1 : cX - i;
        for (int j = 0; j < cY; ++j) {
            oc_c s = {0, 0}, c = {0, 0};
            if (i >= top && i < cX - bottom && j >= leftOff && j < cY - rightOff) {
                const int kY = j + 1 < cY - j ? j + 1 : cY - j;
                for (int x = 0; x < kX; ++x)
                    for (int y = 0; y < kY; ++y) {
                        const int xl = X > i ? x : x + X - kX, xr = X > i ? x + X - kX : x;
                        const int yl = Y > j ? y : y + Y - kY, yr = Y > j ? y + Y - kY : y;
                        kahan_add(&s, &c, c_mul(left[xl * Y + yl], right[xr * Y + yr]));
                    }
            }
            out[i * cY + j] = s;
        }
    }
}

/* GenerateQuarticTerms (interacting_ofl_model.cpp:457-590) for table entries
 * [e0, e1) of the folded order (k4 < kMax, k3 >= k4, k2 < kMax).  bloch holds,
 * per k, the band-0 block of NonInteractingOflModel::m_blochCoefficients:
 * 2 (2cx+1)(2cy+1) complex values with the spin index fastest
 * (GetBlochCoefficients, noninteracting_ofl_model.cpp:550-582).  Writes
 * complex V (interleaved) for each entry. */
void oc_ofl_vtable(int dimX, int dimY, int cutX, int cutY, const double* bloch_, int64_t e0, int64_t e1,
                   double* out_)
{
    const int X = 2 * cutX + 1, Y = 2 * cutY + 1, n = X * Y;
    const int kMax = dimX * dimY;
    const int cdim = (2 * Y - 1) * (2 * X - 1);
    const int cao = X + Y - 1; /* GetArrayConvolution1DSize(X, Y) */
    const oc_c* bloch = (const oc_c*)bloch_;
    oc_c* out = (oc_c*)out_;
    oc_c* arr = (oc_c*)malloc(sizeof(oc_c) * 8 * n);
    oc_c* conv = (oc_c*)malloc(sizeof(oc_c) * 4 * cdim);
    oc_c *Au = arr, *Ad = arr + n, *Bu = arr + 2 * n, *Bd = arr + 3 * n, *Cu = arr + 4 * n, *Cd = arr + 5 * n,
         *Du = arr + 6 * n, *Dd = arr + 7 * n;
    int64_t counter = 0;
    for (int k4 = 0; k4 < kMax; ++k4)
        for (int k3 = k4; k3 < kMax; ++k3)
            for (int k2 = 0; k2 < kMax; ++k2, ++counter) {
                if (counter < e0 || counter >= e1) continue;
                int gx, gy;
                const int k1 = oc_ofl_find_k1(dimX, dimY, k2, k3, k4, &gx, &gy);
                for (int i = 0; i < n; ++i) {
                    const oc_c* b1 = bloch + (size_t)k1 * 2 * n + 2 * i;
                    const oc_c* b2 = bloch + (size_t)k2 * 2 * n + 2 * i;
                    const oc_c* b3 = bloch + (size_t)k3 * 2 * n + 2 * i;
                    const oc_c* b4 = bloch + (size_t)k4 * 2 * n + 2 * i;
                    Au[i].re = b1[0].re; Au[i].im = -b1[0].im; Ad[i].re = b1[1].re; Ad[i].im = -b1[1].im;
                    Bu[i].re = b2[0].re; Bu[i].im = -b2[0].im; Bd[i].re = b2[1].re; Bd[i].im = -b2[1].im;
                    Cu[i] = b3[0]; Cd[i] = b3[1];
                    Du[i] = b4[0]; Dd[i] = b4[1];
                }
                const int o1 = gx < 0 ? -gx : 0, o2 = gy < 0 ? -gy : 0, o3 = gx > 0 ? gx : 0, o4 = gy > 0 ? gy : 0;
                oc_c *c1 = conv, *c2 = conv + cdim, *c3 = conv + 2 * cdim, *c4 = conv + 3 * cdim;
                oc_conv2d((const double*)Bd, (const double*)Du, X, Y, (double*)c1, 0, 0, 0, 0);
                oc_conv2d((const double*)Cd, (const double*)Au, X, Y, (double*)c2, o1, o2, o3, o4);
                oc_conv2d((const double*)Bu, (const double*)Dd, X, Y, (double*)c3, 0, 0, 0, 0);
                oc_conv2d((const double*)Cu, (const double*)Ad, X, Y, (double*)c4, o1, o2, o3, o4);
                oc_c s = {0, 0}, c = {0, 0};
                const int lim = cdim - o1 * cao - o2 - o3 * cao - o4;
                for (int i = 0; i < lim; ++i) {
                    kahan_add(&s, &c, c_mul(c1[i + o3 * cao + o4], c2[i + o1 * cao + o2]));
                    kahan_add(&s, &c, c_mul(c3[i + o3 * cao + o4], c4[i + o1 * cao + o2]));
                }
                out[counter - e0] = s;
            }
    free(arr);
    free(conv);
}

/* ------------------------------------------------- hybrid Wannier tables */

/* ChangeToWannierBasis quartic terms (interacting_ofl_model.cpp:820-905) for
 * the list slots [s0, s1) of entry (k4*K + k3)*K + k2, slot x1: k1 = x1*dimY +
 * ky1, ky1 = FindK1 along y (:410-432); V' = sum over k4x, k3x, k2x (k2x
 * fastest) of V(q1..q4) * conj(U1 * U2) * U3 * U4 from 0, products left to
 * right; V' is kept where |V'| > tol, else 0 (no hash entry). */
void oc_ofl_wannier(int dimX, int dimY, const uint16_t* kt, const double* vt_, const double* u_, double tol,
                    int64_t s0, int64_t s1, uint16_t* k1_out, double* v_out)
{
    const int K = dimX * dimY;
    const oc_c* vt = (const oc_c*)vt_;
    const oc_c* u = (const oc_c*)u_;
    for (int64_t idx = s0; idx < s1; ++idx) {
        const int64_t e = idx / dimX;
        const int x1 = (int)(idx - e * dimX);
        const int k2 = (int)(e % K), k3 = (int)((e / K) % K), k4 = (int)(e / ((int64_t)K * K));
        const int k2y = k2 % dimY, k3y = k3 % dimY, k4y = k4 % dimY;
        const int k1y = oc_mod(k3y + k4y - k2y, dimY);
        const int k1 = x1 * dimY + k1y;
        oc_c sum = {0.0, 0.0};
        for (int k4x = 0; k4x < dimX; ++k4x)
            for (int k3x = 0; k3x < dimX; ++k3x)
                for (int k2x = 0; k2x < dimX; ++k2x) {
                    const int q2 = k2x * dimY + k2y, q3 = k3x * dimY + k3y, q4 = k4x * dimY + k4y;
                    int t1 = q4, t2 = q3;
                    if (q4 > q3) { t1 = q3; t2 = q4; }
                    const int64_t row = ((int64_t)t1 * K + t2 - (int64_t)t1 * (t1 + 1) / 2) * K;
                    const int q1 = kt[row + q2];
                    const oc_c V = vt[row + (q4 > q3 ? q1 : q2)];
                    oc_c p = c_mul(u[(int64_t)k1 * K + q1], u[(int64_t)k2 * K + q2]);
                    p.im = -p.im;
                    oc_c t = c_mul(V, p);
                    t = c_mul(t, u[(int64_t)k3 * K + q3]);
                    t = c_mul(t, u[(int64_t)k4 * K + q4]);
                    sum.re = sum.re + t.re;
                    sum.im = sum.im + t.im;
                }
        k1_out[idx - s0] = (uint16_t)k1;
        const int keep = hypot(sum.re, sum.im) > tol;
        v_out[2 * (idx - s0)] = keep ? sum.re : 0.0;
        v_out[2 * (idx - s0) + 1] = keep ? sum.im : 0.0;
    }
}
